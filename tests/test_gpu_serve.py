"""Continuous batching (lf_server_*) on the GPU.

Jobs arrive while others run, wait when memory is short, are admitted as
memory frees up and share every batched step.  Each job's final state must
be bitwise the state the same scene reaches alone in its own lf_scene, under
both admission policies.
"""
import numpy as np
import pytest

from libfluid_amd import _lib as L
from libfluid_amd import configs
from libfluid_amd.serve import Server, drive, served_jobs

pytestmark = pytest.mark.gpu
STATE = ("pos", "vel", "acc", "colour")


def small_jobs(k=10, seed=11):
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for j in range(k):
        n = int(rng.integers(800, 6000))
        cfg = configs.block_scene(n, rng.uniform(1.5, 2.5, 3), rng.normal(size=3), seed * 100 + j, f"sj{j}")
        if j % 3 == 1:
            cfg.params.viscosity, cfg.params.surface_tension, cfg.params.adhesion = 0.2, 1.0, 0.1
        out.append((cfg, int(rng.integers(3, 25))))
    return out


def solo(gpu, cfg, steps):
    sc = gpu.Scene(cfg.params, cfg.n)
    try:
        sc.upload(cfg.pos, cfg.vel)
        sc.run(cfg.dt, steps)
        sc.sync()
        return sc.download(L.LF_ORDER_ID)
    finally:
        sc.close()


def run_stream(jobs, backfill, slots, cells, run_ahead=2, graphs=True):
    """Half the jobs at tick 0, the rest after 4 ticks; ticks until all done."""
    srv = Server(slots, cells, 64, backfill=backfill)
    try:
        srv.configure(run_ahead, graphs)
        ids = {}
        half = len(jobs) // 2
        for k in range(half):
            cfg, steps = jobs[k]
            ids[srv.submit(cfg.params, cfg.pos, cfg.vel, cfg.dt, steps)] = k
        out, order = {}, []
        ticks = 0
        while len(out) < len(jobs):
            if ticks == 4:
                for k in range(half, len(jobs)):
                    cfg, steps = jobs[k]
                    ids[srv.submit(cfg.params, cfg.pos, cfg.vel, cfg.dt, steps)] = k
            srv.tick()
            ticks += 1
            for j in srv.take_finished():
                order.append(ids[j])
                out[ids[j]] = srv.result(j)
            assert ticks < 2000
        return out, order, srv.stats()
    finally:
        srv.close()


@pytest.mark.parametrize("backfill", [True, False])
def test_served_jobs_match_solo_runs(gpu, backfill):
    jobs = small_jobs()
    # room for about three jobs at a time: the rest queue and fold in later
    slots = 3 * 6144
    cells = 3 * max(c.params.ncells + 1 for c, _ in jobs)
    out, order, st = run_stream(jobs, backfill, slots, cells)
    for k, (cfg, steps) in enumerate(jobs):
        ref = solo(gpu, cfg, steps)
        for f in STATE:
            assert np.array_equal(out[k][f], ref[f]), f"job {k} ({cfg.n} particles, {steps} steps): {f} differs"
    assert st["submitted"] == st["finished"] == len(jobs) and st["queued"] == st["running"] == 0
    assert st["particle_steps"] == sum(c.n * s for c, s in jobs)
    assert 0 < st["step_p50_ms"] <= st["step_p99_ms"] and 0 < st["mean_occupancy"] <= 1
    assert st["wait_p99_ms"] > 0  # some job had to wait for memory
    print(f"\nbackfill={backfill}: finish order {order}, {st['step_ticks']} stepping ticks, "
          f"occupancy {st['mean_occupancy']:.2f}")


def test_submit_rejects_a_job_that_can_never_fit(gpu):
    cfg = configs.block_scene(4000, (2.0, 2.0, 2.0), (0, -1, 0), 3)
    with Server(2048, 100000, 4) as srv:
        with pytest.raises(L.LfError) as e:
            srv.submit(cfg.params, cfg.pos, cfg.vel, cfg.dt, 5)
        assert e.value.code == L.LF_ERR_FULL


def test_poisson_stream_drains(gpu):
    """A short real-time Poisson stream of the served workload's shape (scaled
    down): every job finishes and reports a latency."""
    jobs = served_jobs(8, rate=40.0, seed=2, n_lo=2048, n_hi=16384, steps_lo=5, steps_hi=30)
    cap = 3 * 16384
    cells = 3 * max(j.cfg.params.ncells + 1 for j in jobs)
    with Server(cap, cells, 32) as srv:
        r = drive(srv, jobs, keep_results=True)
        st = srv.stats()
    assert len(r["latency_ms"]) == len(jobs) and all(v > 0 for v in r["latency_ms"].values())
    assert st["finished"] == len(jobs)
    for k in (0, len(jobs) - 1):
        ref = solo(gpu, jobs[k].cfg, jobs[k].steps)
        assert np.array_equal(r["results"][k]["pos"], ref["pos"])


@pytest.mark.parametrize("run_ahead,graphs", [(0, False), (0, True), (2, False), (2, True)])
def test_results_do_not_depend_on_run_ahead_or_graphs(gpu, run_ahead, graphs):
    """Jobs that retire in the same tick each keep their own exported state
    (their pinned buffers are held until the state is moved out), whether
    ticks wait for the device or run ahead of it."""
    jobs = small_jobs(6, seed=12)
    for k in range(0, 6, 2):  # pairs of jobs with equal step counts retire together
        jobs[k + 1] = (jobs[k + 1][0], jobs[k][1])
    slots = 4 * 6144
    cells = 4 * max(c.params.ncells + 1 for c, _ in jobs)
    out, _, _ = run_stream(jobs, True, slots, cells, run_ahead=run_ahead, graphs=graphs)
    for k, (cfg, steps) in enumerate(jobs):
        ref = solo(gpu, cfg, steps)
        for f in STATE:
            assert np.array_equal(out[k][f], ref[f]), f"job {k}: {f} differs"
