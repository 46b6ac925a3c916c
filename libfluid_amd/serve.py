"""Serving: continuous batching of scene jobs (lf_server_*), the served
workload generator and a wall-clock driver.

The C++ server admits queued jobs into free device memory, advances every
running job one substep per tick in one batched step and retires finished
jobs.  `drive` feeds it a seeded Poisson arrival stream in real time.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import time

import numpy as np

from . import _lib as L
from . import configs
from ._lib import LfParams, LfServerStats, check, fptr, lib


class Server:
    """One continuous-batching server on the current HIP device."""

    def __init__(self, max_particles: int, max_cells: int, max_scenes: int, backfill: bool = True):
        self._h = C.c_void_p()
        policy = L.SERVE_BACKFILL if backfill else L.SERVE_FIFO
        check(lib.lf_server_create(int(max_particles), int(max_cells), int(max_scenes), policy, C.byref(self._h)),
              "lf_server_create")
        self._n = {}

    @classmethod
    def for_budget(cls, budget_bytes: int, cells_per_particle: float = 1.5, max_scenes: int = 4096,
                   backfill: bool = True) -> "Server":
        """Largest server whose device footprint fits budget_bytes, with
        cells_per_particle grid cells provisioned per particle slot."""
        per = Server.footprint(1 << 20, int((1 << 20) * cells_per_particle), 0) / float(1 << 20)
        slots = int(budget_bytes / per) // 512 * 512
        return cls(slots, int(slots * cells_per_particle), max_scenes, backfill)

    @staticmethod
    def footprint(max_particles: int, max_cells: int, max_scenes: int) -> int:
        b = C.c_uint64(0)
        check(lib.lf_batch_footprint(int(max_particles), int(max_cells), int(max_scenes), C.byref(b)),
              "lf_batch_footprint")
        return b.value

    def close(self) -> None:
        if self._h and self._h.value:
            lib.lf_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, params: LfParams, pos: np.ndarray, vel: np.ndarray | None, dt: float, steps: int) -> int:
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float32).reshape(-1, 3)
        job = C.c_uint64(0)
        check(lib.lf_server_submit(self._h, C.byref(params), pos.shape[0], fptr(pos), fptr(vel), C.c_float(dt),
                                   int(steps), C.byref(job)), "lf_server_submit")
        self._n[job.value] = pos.shape[0]
        return job.value

    def configure(self, run_ahead: int = 2, graphs: bool = True) -> None:
        check(lib.lf_server_configure(self._h, int(run_ahead), 1 if graphs else 0), "lf_server_configure")

    def tick(self) -> int:
        fin = C.c_uint32(0)
        check(lib.lf_server_tick(self._h, C.byref(fin)), "lf_server_tick")
        return fin.value

    def take_finished(self, cap: int = 4096) -> list:
        ids = (C.c_uint64 * cap)()
        cnt = C.c_uint32(0)
        check(lib.lf_server_take_finished(self._h, ids, cap, C.byref(cnt)), "lf_server_take_finished")
        return list(ids[: cnt.value])

    def status(self, job: int) -> tuple:
        st, done = C.c_int(0), C.c_uint32(0)
        check(lib.lf_server_job_status(self._h, int(job), C.byref(st), C.byref(done)), "lf_server_job_status")
        return st.value, done.value

    def result(self, job: int, want_state: bool = True) -> dict:
        n = self._n.pop(int(job))
        out = {k: np.empty((n, 3), np.float32) for k in ("pos", "colour", "vel", "acc")} if want_state else {}
        lat = C.c_double(0.0)
        check(lib.lf_server_job_result(self._h, int(job), fptr(out.get("pos")), fptr(out.get("colour")),
                                       fptr(out.get("vel")), fptr(out.get("acc")), C.byref(lat)),
              "lf_server_job_result")
        out["latency_ms"] = lat.value
        return out

    def stats(self) -> dict:
        s = LfServerStats()
        check(lib.lf_server_get_stats(self._h, C.byref(s)), "lf_server_get_stats")
        return s.as_dict()


@dataclasses.dataclass
class Job:
    cfg: configs.Config
    steps: int
    arrival: float  # seconds after the stream starts


def served_jobs(count: int, rate: float, seed: int = 5, n_lo: int = 4096, n_hi: int = 262144,
                steps_lo: int = 200, steps_hi: int = 2000) -> list:
    """The served workload: `count` jobs with Poisson arrivals at `rate` per
    second; particle counts log-uniform in [n_lo, n_hi], step counts uniform
    in [steps_lo, steps_hi], each a jittered block with its own box and
    gravity direction (configs.block_scene), all from one seed."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    jobs = []
    for k in range(count):
        t += float(rng.exponential(1.0 / rate))
        n = int(round(float(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi))))))
        steps = int(rng.integers(steps_lo, steps_hi + 1))
        cfg = configs.block_scene(n, rng.uniform(1.5, 2.5, 3), rng.normal(size=3), seed * 100000 + k, f"job{k}")
        jobs.append(Job(cfg, steps, t))
    return jobs


def drive(server: Server, jobs: list, keep_results: bool = False, log=None, sample_s: float = 0.05) -> dict:
    """Feeds `jobs` to the server at their arrival times (wall clock from the
    call) and ticks until every job finished.  Returns wall time, per-job
    latencies, a trace of (seconds, particle-steps so far, resident
    particles) sampled every sample_s, and, with keep_results, every job's
    final state by index."""
    order = sorted(range(len(jobs)), key=lambda k: jobs[k].arrival)
    ids = {}
    results = {}
    lat = {}
    trace = []
    nxt = 0
    t0 = time.perf_counter()
    last_log = t0
    last_sample = t0
    while len(lat) < len(jobs):
        now = time.perf_counter() - t0
        while nxt < len(order) and jobs[order[nxt]].arrival <= now:
            k = order[nxt]
            ids[server.submit(jobs[k].cfg.params, jobs[k].cfg.pos, jobs[k].cfg.vel, jobs[k].cfg.dt,
                              jobs[k].steps)] = k
            nxt += 1
        if nxt == len(lat):  # nothing queued or running: idle until the next arrival
            time.sleep(max(0.0, min(0.001, jobs[order[nxt]].arrival - now)))
            continue
        if server.tick():
            for j in server.take_finished():
                r = server.result(j, want_state=keep_results)
                lat[ids[j]] = r.pop("latency_ms")
                if keep_results:
                    results[ids[j]] = r
        now = time.perf_counter()
        if now - last_sample >= sample_s:
            last_sample = now
            st = server.stats()
            trace.append((now - t0, st["particle_steps"], st["resident_particles"]))
        if log and now - last_log > 5.0:
            last_log = now
            st = server.stats()
            log(f"[serve] t={last_log - t0:.1f}s submitted {nxt}/{len(jobs)} finished {len(lat)} "
                f"running {st['running']} queued {st['queued']} resident {st['resident_particles']}")
    wall = time.perf_counter() - t0
    st = server.stats()
    trace.append((wall, st["particle_steps"], st["resident_particles"]))
    return {"wall_s": wall, "latency_ms": lat, "results": results, "trace": trace,
            "last_arrival_s": max(j.arrival for j in jobs) if jobs else 0.0}


def steady_window(run: dict, slot_cap: int, jobs: list, skip: float = 0.125) -> dict:
    """Throughput and occupancy of a driven stream between the arrival of
    its first `skip` fraction of jobs (ramp-up) and its last arrival (after
    which the server only drains)."""
    arr = sorted(j.arrival for j in jobs)
    ta, tb = arr[int(skip * len(arr))], arr[-1]
    tr = [x for x in run["trace"] if ta <= x[0] <= tb]
    if len(tr) < 2:
        return {"t0_s": ta, "t1_s": tb, "value": None, "mean_occupancy": None}
    dt = tr[-1][0] - tr[0][0]
    return {"t0_s": tr[0][0], "t1_s": tr[-1][0], "value": (tr[-1][1] - tr[0][1]) / dt,
            "mean_occupancy": float(np.mean([x[2] for x in tr])) / slot_cap}
