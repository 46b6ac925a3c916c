"""The data formats either side of the stepping path, on the GPU.

* checkpoint: a scene saved after k substeps and loaded into a new scene (or
  into a batch) continues bitwise as the original does;
* device-side setup: the GPU lattice generators write exactly the positions
  the host generators produce;
* asynchronous .geo export: files written by the background exporter while
  stepping continues are byte-identical to download + write at the same
  state.
"""
import filecmp
import os

import numpy as np
import pytest

from libfluid_amd import _lib as L
from libfluid_amd import configs
from libfluid_amd.scene import fluid_block, jittered_lattice, write_geo

pytestmark = pytest.mark.gpu
STATE = ("pos", "vel", "acc", "colour")


def equal_states(a, b):
    return all(np.array_equal(a[k], b[k]) for k in STATE)


@pytest.mark.parametrize("case", ["c1", "double_dam"])
def test_checkpoint_resume_is_exact(gpu, tmp_path, case):
    cfg = configs.dambreak_c1() if case == "c1" else configs.double_dam(n=4096, res=16)
    path = str(tmp_path / "scene.lfck")
    sc = gpu.Scene(cfg.params, cfg.n)
    try:
        sc.upload(cfg.pos, cfg.vel)
        sc.run(cfg.dt, 10)
        sc.sync()
        sc.checkpoint_save(path)
        sc.run(cfg.dt, 10)
        sc.sync()
        ref = sc.download(L.LF_ORDER_ID)
    finally:
        sc.close()
    re = gpu.Scene.checkpoint_load(path)
    try:
        assert re.step_count() == 10 and re.n == cfg.n
        assert bytes(re.params) == bytes(cfg.params)
        re.run(cfg.dt, 10)
        re.sync()
        got = re.download(L.LF_ORDER_ID)
    finally:
        re.close()
    assert equal_states(got, ref), "resumed trajectory differs"


def test_batch_checkpoint_into_batch(gpu, tmp_path):
    cfgs = [configs.block_scene(n, (2.0, 1.8, 2.2), g, 40 + k) for k, (n, g) in
            enumerate([(3000, (0, -1, 0)), (5000, (1, -1, 0)), (2000, (0, -1, 1))])]
    slots = sum((c.n + 511) // 512 * 512 for c in cfgs)
    cells = sum(c.params.ncells + 1 for c in cfgs)
    path = str(tmp_path / "b.lfck")
    with gpu.Batch(slots, cells, 3) as b:
        ids = [b.add(c.params, c.n, c.dt) for c in cfgs]
        for i, c in zip(ids, cfgs):
            b.upload(i, c.pos, c.vel)
        b.step(6)
        b.checkpoint_save(ids[1], path)
        b.step(6)
        b.sync()
        ref = b.download(ids[1])
    with gpu.Batch(slots, cells, 3) as b2:
        other = b2.add(cfgs[0].params, cfgs[0].n, cfgs[0].dt)  # a neighbour in the new batch
        b2.upload(other, cfgs[0].pos, cfgs[0].vel)
        sid = b2.checkpoint_load(path)
        assert b2.scene_info(sid)["steps"] == 6
        b2.step(6)
        b2.sync()
        assert equal_states(b2.download(sid), ref)


def test_device_setup_matches_host_generators(gpu):
    p = configs.dambreak_c1().params
    n = 20 * 18 * 16 + 1000
    with gpu.Scene(p, n) as sc:
        sc.setup_jittered_lattice(20, 18, 16, 0.01, (0.01, 0.02, 0.03), 0.0005, 1234)
        w, m, s = sc.setup_fluid_block(1000, (0.5, 0.1, 0.1), (0.7, 0.3, 0.25), first=20 * 18 * 16)
        d = sc.download(L.LF_ORDER_ID)
    host = jittered_lattice(20, 18, 16, 0.01, (0.01, 0.02, 0.03), 0.0005, 1234)
    assert np.array_equal(d["pos"][: 20 * 18 * 16], host)
    pts, mass, psize = fluid_block(1000, (0.5, 0.1, 0.1), (0.7, 0.3, 0.25), p.rest_density)
    assert w == pts.shape[0] and np.float32(m) == np.float32(mass) and np.float32(s) == np.float32(psize)
    assert np.array_equal(d["pos"][20 * 18 * 16: 20 * 18 * 16 + w], pts)


def test_async_geo_export_matches_sync(gpu, tmp_path):
    cfg = configs.scaled_c1(16, 12, 10)
    a, b = str(tmp_path / "async"), str(tmp_path / "sync")
    with gpu.Scene(cfg.params, cfg.n) as sc:
        sc.upload(cfg.pos, cfg.vel)
        for k in range(1, 4):
            sc.run(cfg.dt, 3)
            sc.export_geo_async(a, k)  # snapshot of this state, written later
            d = sc.download(L.LF_ORDER_CELL)
            write_geo(b, k, d["pos"], d["colour"], d["vel"], float(np.float32(cfg.params.particle_size)))
            sc.run(cfg.dt, 2)  # steps continue while the file is written
        sc.export_wait()
    for k in range(1, 4):
        fa, fb = f"{a}.{k:04d}.geo", f"{b}.{k:04d}.geo"
        assert os.path.exists(fa) and filecmp.cmp(fa, fb, shallow=False), f"frame {k} differs"


def test_async_export_reports_io_errors(gpu, tmp_path):
    cfg = configs.scaled_c1(8, 8, 8)
    with gpu.Scene(cfg.params, cfg.n) as sc:
        sc.upload(cfg.pos, cfg.vel)
        sc.export_geo_async(str(tmp_path / "no_such_dir" / "x"), 1)
        with pytest.raises(L.LfError):
            sc.export_wait()
        sc.export_geo_async(str(tmp_path / "ok"), 1)
        sc.export_wait()  # the error was reported once
