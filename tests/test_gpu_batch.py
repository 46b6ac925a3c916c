"""Batches of independent scenes (lf_batch_*) on the GPU.

A scene stepped inside a batch must give bitwise the same state as the same
scene stepped alone in its own lf_scene: same cell table, same sorted order,
same neighbour walks, same arithmetic.  These tests check that for uneven
scenes with every force term on or off, through admission, removal,
compaction, pausing, graph replay versus eager launches and autotuning, and
check one batched scene against the FP64 oracle directly.
"""
import numpy as np
import pytest

from libfluid_amd import _lib as L
from libfluid_amd import configs
from oracle.pyoracle import PIN_F32
from tests.helpers import normwise

pytestmark = pytest.mark.gpu
TOL = 1e-5
STATE = ("pos", "vel", "acc", "colour")


def mixed_scenes(k=6, seed=5):
    """Uneven scenes: some with the reference's dam-break coefficients
    (viscosity, tension, adhesion on), some plain."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for j in range(k):
        n = int(rng.integers(700, 5000))
        cfg = configs.block_scene(n, rng.uniform(1.5, 2.5, 3), rng.normal(size=3), seed * 100 + j, f"mixed{j}")
        if j % 2:
            cfg.params.viscosity, cfg.params.surface_tension, cfg.params.adhesion = 0.2, 1.0, 0.1
        out.append(cfg)
    return out


def solo(gpu, cfg, steps, graphs=True):
    sc = gpu.Scene(cfg.params, cfg.n)
    try:
        sc.set_graphs(graphs)
        sc.upload(cfg.pos, cfg.vel)
        sc.run(cfg.dt, steps)
        sc.sync()
        return sc.download(L.LF_ORDER_ID)
    finally:
        sc.close()


def batch_for(gpu, cfgs, extra_slots=0, extra_cells=0, extra_scenes=0):
    slots = sum((c.n + 511) // 512 * 512 for c in cfgs) + extra_slots
    cells = sum(c.params.ncells + 1 for c in cfgs) + extra_cells
    return gpu.Batch(slots, cells, len(cfgs) + extra_scenes)


def assert_same(a, b, label):
    for k in STATE:
        assert np.array_equal(a[k], b[k]), f"{label}: {k} differs"


def test_batch_matches_solo_scenes(gpu):
    cfgs = mixed_scenes()
    steps = 15
    ref = [solo(gpu, c, steps) for c in cfgs]
    with batch_for(gpu, cfgs) as b:
        ids = [b.add(c.params, c.n, c.dt) for c in cfgs]
        for i, c in zip(ids, cfgs):
            b.upload(i, c.pos, c.vel)
        b.step(steps)
        b.sync()
        for j, (i, c) in enumerate(zip(ids, cfgs)):
            assert b.scene_info(i)["steps"] == steps
            assert_same(b.download(i), ref[j], f"scene {j}")


@pytest.mark.parametrize("graphs", [True, False])
def test_ragged_batch_with_tiny_scenes(gpu, graphs):
    """Scenes of 1, 2, 511, 512 and 513 particles (no neighbours at all, a
    chunk filled exactly, one slot into the next chunk) beside a larger one:
    every batched scene equals its solo run bit for bit, with graph replay
    and with eager launches."""
    rng = np.random.Generator(np.random.PCG64(11))
    sizes = (1, 2, 511, 512, 513, 3000)
    cfgs = [configs.block_scene(n, rng.uniform(1.5, 2.5, 3), rng.normal(size=3), 1100 + j, f"ragged{j}")
            for j, n in enumerate(sizes)]
    steps = 6
    ref = [solo(gpu, c, steps, graphs) for c in cfgs]
    with batch_for(gpu, cfgs) as b:
        b.set_graphs(graphs)
        ids = [b.add(c.params, c.n, c.dt) for c in cfgs]
        for i, c in zip(ids, cfgs):
            b.upload(i, c.pos, c.vel)
        b.step(steps)
        b.sync()
        for j, (i, c) in enumerate(zip(ids, cfgs)):
            out = b.download(i)
            assert np.all(np.isfinite(out["pos"])), f"{c.n} particles: non-finite positions"
            assert_same(out, ref[j], f"scene of {c.n} particles")


def test_batch_graphs_and_eager_agree(gpu):
    cfgs = mixed_scenes(4, seed=6)
    outs = []
    for graphs in (True, False):
        with batch_for(gpu, cfgs) as b:
            b.set_graphs(graphs)
            ids = [b.add(c.params, c.n, c.dt) for c in cfgs]
            for i, c in zip(ids, cfgs):
                b.upload(i, c.pos, c.vel)
            for _ in range(10):
                b.step(1)
            b.sync()
            outs.append([b.download(i) for i in ids])
    for j in range(len(cfgs)):
        assert_same(outs[0][j], outs[1][j], f"scene {j} graph vs eager")


def test_admission_removal_compaction_and_pause(gpu):
    """Scenes leave, a paused scene stays put, and a late arrival only fits
    after compaction; every survivor still matches its solo run."""
    cfgs = mixed_scenes(5, seed=7)  # 10, 7, 10, 6, 10 chunks of 512 slots
    late = configs.block_scene(5500, (1.6, 1.6, 1.6), (0.0, -1.0, 0.0), 77, "late")  # 11 chunks
    assert (late.n + 511) // 512 > max((cfgs[1].n + 511) // 512, (cfgs[3].n + 511) // 512)
    with batch_for(gpu, cfgs, extra_scenes=1) as b:
        ids = [b.add(c.params, c.n, c.dt) for c in cfgs]
        for i, c in zip(ids, cfgs):
            b.upload(i, c.pos, c.vel)
        b.step(5)
        paused = b.download(ids[4])
        b.set_active(ids[4], False)
        b.remove(ids[1])
        b.remove(ids[3])
        # room for `late` exists only once the holes left by scenes 1 and 3 are merged
        lid = b.try_add(late.params, late.n, late.dt)
        assert lid is not None, "the late scene should fit after compaction"
        b.upload(lid, late.pos, late.vel)
        b.step(7)
        b.sync()
        assert_same(b.download(ids[0]), solo(gpu, cfgs[0], 12), "scene 0 after churn")
        assert_same(b.download(ids[2]), solo(gpu, cfgs[2], 12), "scene 2 after churn")
        assert_same(b.download(ids[4]), paused, "paused scene")
        assert b.scene_info(ids[4])["steps"] == 5
        assert_same(b.download(lid), solo(gpu, late, 7), "late arrival")
        b.set_active(ids[4], True)
        b.step(3)
        b.sync()
        assert_same(b.download(ids[4]), solo(gpu, cfgs[4], 8), "resumed scene")


def test_spare_capacity_spans_and_long_scans(gpu):
    """A pool with far more room than its scenes.  A filler scene (100,000
    slots, a 150^3 grid) is admitted first and scene A after it, so A lies
    high in both slot and bucket space once the filler leaves; B then takes
    the bottom (first fit).  Stepping launches cover only the active span,
    whose bucket part here needs more than 2,048 scan tiles (the tile-prefix
    pass); pausing A shrinks the span and resuming it grows it again (each a
    new capture).  Every scene matches its solo run bit for bit."""
    a, bcfg = mixed_scenes(2, seed=12)
    filler = configs.Config("filler", configs._plain_params((3.0, 3.0, 3.0), (150, 150, 150)),
                            np.zeros((100_000, 3), np.float32), np.zeros((100_000, 3), np.float32), 1e-4, 1)
    assert filler.params.ncells + 1 > 2048 * 1024
    with batch_for(gpu, [filler, a, bcfg], extra_slots=64 * 512, extra_cells=1_000_000) as b:
        fid = b.add(filler.params, filler.n, filler.dt)
        ida = b.add(a.params, a.n, a.dt)
        b.remove(fid)
        idb = b.add(bcfg.params, bcfg.n, bcfg.dt)
        b.upload(ida, a.pos, a.vel)
        b.upload(idb, bcfg.pos, bcfg.vel)
        b.step(4)
        b.sync()
        assert_same(b.download(ida), solo(gpu, a, 4), "high scene")
        assert_same(b.download(idb), solo(gpu, bcfg, 4), "low scene")
        b.set_active(ida, False)
        b.step(2)
        b.sync()
        assert_same(b.download(idb), solo(gpu, bcfg, 6), "low scene, high one paused")
        assert_same(b.download(ida), solo(gpu, a, 4), "paused high scene")
        b.set_active(ida, True)
        b.step(3)
        b.sync()
        assert_same(b.download(ida), solo(gpu, a, 7), "high scene resumed")
        assert_same(b.download(idb), solo(gpu, bcfg, 9), "low scene")


def test_batch_full_reports_full(gpu):
    cfgs = mixed_scenes(2, seed=8)
    with batch_for(gpu, cfgs) as b:
        for c in cfgs:
            b.add(c.params, c.n, c.dt)
        extra = configs.block_scene(1000, (2.0, 2.0, 2.0), (0, -1, 0), 9)
        assert b.try_add(extra.params, extra.n, extra.dt) is None
        u = b.usage()
        assert u["scenes"] == 2 and u["slots_used"] <= u["slot_cap"]


def test_autotune_and_block_sizes_do_not_change_results(gpu):
    cfgs = mixed_scenes(3, seed=9)
    outs = []
    for mode in ("default", "autotune", "small", "large", "rows"):
        with batch_for(gpu, cfgs) as b:
            ids = [b.add(c.params, c.n, c.dt) for c in cfgs]
            for i, c in zip(ids, cfgs):
                b.upload(i, c.pos, c.vel)
            b.step(2)
            if mode == "autotune":
                ms = b.autotune(2)
                assert ms.shape == (4, 3) and np.all(ms[[0, 2, 3]] > 0)
                assert all(x in (128, 256, 512) for x in b.tuning())
                assert all(f in (0, 1, 2) for f in b.walks())
            elif mode == "rows":
                b.set_walks((0, 0, 0))
                assert b.walks() == (0, 0, 0)
            elif mode == "small":
                b.set_tuning((128, 128, 128, 128))
            elif mode == "large":
                b.set_tuning((512, 512, 512, 512))
            b.step(8)
            b.sync()
            outs.append([b.download(i) for i in ids])
    for m in range(1, 5):
        for j in range(len(cfgs)):
            assert_same(outs[0][j], outs[m][j], f"mode {m} scene {j}")


def test_batched_scene_against_oracle(gpu, oracle_omp):
    """A batched scene after 10 substeps against the free-running FP64 oracle."""
    cfgs = mixed_scenes(3, seed=10)
    steps = 10
    with batch_for(gpu, cfgs) as b:
        ids = [b.add(c.params, c.n, c.dt) for c in cfgs]
        for i, c in zip(ids, cfgs):
            b.upload(i, c.pos, c.vel)
        b.step(steps)
        b.sync()
        for i, c in zip(ids, cfgs):
            g = b.download(i)
            p = np.array(c.pos, np.float64)
            v = np.zeros_like(p)
            a = np.zeros_like(p)
            for _ in range(steps):
                orc = oracle_omp.step(c.params, float(np.float32(c.dt)), p, v, a, PIN_F32)
            e = normwise(g["pos"], p)
            print(f"\n[{c.name}] n={c.n} pos normwise {e:.3e}, density range "
                  f"{orc['density'].min():.1f}..{orc['density'].max():.1f}")
            assert e <= TOL
