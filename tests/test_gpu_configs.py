"""GPU: the batched and served workloads at the configurations they are
measured on.

* 256 independent scenes, particle counts log-uniform in [2048, 65536]
  (configs.batched_c4), stepped together in one batch: bitwise the same
  scenes stepped one after another, and two of them against the FP64
  oracle.
* A seeded Poisson stream of jobs of 4k-256k particles (serve.served_jobs),
  at reduced step counts, through the continuous-batching server with room
  for a few jobs at a time: every served job bitwise its solo run, and one
  served job against the FP64 oracle.
"""
import numpy as np
import pytest

from libfluid_amd import _lib as L
from libfluid_amd import configs
from libfluid_amd.serve import Server, served_jobs
from tests.helpers import normwise

pytestmark = pytest.mark.gpu
TOL = 1e-5
STATE = ("pos", "vel", "acc", "colour")
SUBSTEPS = 3


def solo(gpu, cfg, steps):
    with gpu.Scene(cfg.params, cfg.n) as sc:
        sc.upload(cfg.pos, cfg.vel)
        sc.run(cfg.dt, steps)
        sc.sync()
        return sc.download(L.LF_ORDER_ID)


def oracle_run(orc, cfg, steps):
    p = cfg.pos.astype(np.float64)
    v = cfg.vel.astype(np.float64)
    a = np.zeros_like(p)
    for _ in range(steps):
        orc.step(cfg.params, float(np.float32(cfg.dt)), p, v, a, 0, want=())
    return p, v


def test_batched_256_scenes(gpu, oracle_omp):
    bc = configs.batched_c4()
    assert len(bc.scenes) == 256 and 2048 <= min(s.n for s in bc.scenes) and max(s.n for s in bc.scenes) <= 65536
    slots = sum((s.n + 511) // 512 * 512 for s in bc.scenes)
    cells = sum(s.params.ncells + 1 for s in bc.scenes)
    with gpu.Batch(slots, cells, len(bc.scenes)) as b:
        ids = [b.add(s.params, s.n, s.dt) for s in bc.scenes]
        for i, s in zip(ids, bc.scenes):
            b.upload(i, s.pos, s.vel)
        b.step(SUBSTEPS)
        b.sync()
        got = [b.download(i) for i in ids]
    print(f"\n{len(bc.scenes)} scenes, {bc.n} particles, {SUBSTEPS} substeps")
    for k, s in enumerate(bc.scenes):
        ref = solo(gpu, s, SUBSTEPS)
        for f in STATE:
            assert np.array_equal(got[k][f], ref[f]), f"scene {k} ({s.n} particles): {f} differs from its solo run"
    for k in (int(np.argmin([s.n for s in bc.scenes])), int(np.argmax([s.n for s in bc.scenes]))):
        s = bc.scenes[k]
        p64, v64 = oracle_run(oracle_omp, s, SUBSTEPS)
        err = normwise(got[k]["pos"], p64)
        print(f"scene {k} ({s.n} particles): position {err:.2e} normwise vs the oracle")
        assert err <= TOL


def test_served_poisson_stream(gpu, oracle_omp):
    jobs = served_jobs(12, rate=200.0, seed=5, steps_lo=4, steps_hi=12)
    sizes = [j.cfg.n for j in jobs]
    assert min(sizes) >= 4096 and max(sizes) <= 262144
    # room for roughly three average jobs: later arrivals queue and fold in
    slots = 3 * 512 * ((int(np.mean(sizes)) + 511) // 512) + 262144
    cells = 3 * max(j.cfg.params.ncells + 1 for j in jobs)
    out = {}
    with Server(slots, cells, 64) as srv:
        ids = {}
        order = sorted(range(len(jobs)), key=lambda k: jobs[k].arrival)
        for k in order[: len(order) // 2]:
            ids[srv.submit(jobs[k].cfg.params, jobs[k].cfg.pos, jobs[k].cfg.vel, jobs[k].cfg.dt, jobs[k].steps)] = k
        ticks = 0
        while len(out) < len(jobs):
            if ticks == 3:  # the rest arrive while the first are on the GPU
                for k in order[len(order) // 2:]:
                    ids[srv.submit(jobs[k].cfg.params, jobs[k].cfg.pos, jobs[k].cfg.vel, jobs[k].cfg.dt,
                                   jobs[k].steps)] = k
            srv.tick()
            ticks += 1
            for j in srv.take_finished():
                out[ids[j]] = srv.result(j)
            assert ticks < 5000
        st = srv.stats()
    print(f"\n{len(jobs)} jobs of {min(sizes)}-{max(sizes)} particles, {st['ticks']} ticks, "
          f"mean occupancy {st['mean_occupancy']:.2f}")
    assert st["admitted"] == len(jobs) and st["finished"] == len(jobs)
    for k, j in enumerate(jobs):
        ref = solo(gpu, j.cfg, j.steps)
        for f in STATE:
            assert np.array_equal(out[k][f], ref[f]), f"job {k} ({j.cfg.n} particles): {f} differs from solo"
    k = int(np.argmax(sizes))
    p64, v64 = oracle_run(oracle_omp, jobs[k].cfg, jobs[k].steps)
    err = normwise(out[k]["pos"], p64)
    print(f"job {k} ({sizes[k]} particles, {jobs[k].steps} steps): position {err:.2e} normwise vs the oracle")
    assert err <= TOL
