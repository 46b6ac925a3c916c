"""Python handle over one lf_scene (device-resident WCSPH state).

Thin: every method is one C-ABI call; arrays cross as numpy float32 (N,3).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import LfParams, check, fptr, lib, uptr


def default_params(res: int) -> LfParams:
    """FluidSystem::init defaults on the unit cube with a res^3 grid."""
    p = LfParams()
    check(lib.lf_params_default(C.byref(p), int(res)), "lf_params_default")
    return p


def make_params(**kw) -> LfParams:
    """Params from keyword values on top of the res=1 defaults."""
    p = default_params(1)
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        cur = getattr(p, k)
        if hasattr(cur, "__len__"):
            for i, x in enumerate(v):
                cur[i] = x
        else:
            setattr(p, k, v)
    return p


class Scene:
    """One WCSPH scene on the current HIP device."""

    def __init__(self, params: LfParams, num_particles: int):
        self._h = C.c_void_p()
        check(lib.lf_create(C.byref(params), int(num_particles), C.byref(self._h)), "lf_create")
        self.n = int(num_particles)
        self.params = LfParams.from_buffer_copy(params)

    # -- lifetime -----------------------------------------------------------
    def close(self) -> None:
        if self._h and self._h.value:
            lib.lf_destroy(self._h)
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

    @property
    def handle(self):
        return self._h

    # -- parameters ---------------------------------------------------------
    def set_params(self, params: LfParams) -> None:
        check(lib.lf_set_params(self._h, C.byref(params)), "lf_set_params")
        self.params = LfParams.from_buffer_copy(params)

    # -- data ---------------------------------------------------------------
    def upload(self, pos: np.ndarray, vel: np.ndarray | None = None, acc: np.ndarray | None = None,
               first: int = 0) -> None:
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float32).reshape(-1, 3)
        acc = None if acc is None else np.ascontiguousarray(acc, dtype=np.float32).reshape(-1, 3)
        check(lib.lf_upload(self._h, int(first), pos.shape[0], fptr(pos), fptr(vel), fptr(acc)), "lf_upload")

    def download(self, order: int = L.LF_ORDER_ID) -> dict:
        n = self.n
        out = {k: np.empty((n, 3), np.float32) for k in ("pos", "colour", "vel", "acc")}
        ids = np.empty(n, np.uint32)
        check(lib.lf_download(self._h, int(order), fptr(out["pos"]), fptr(out["colour"]), fptr(out["vel"]),
                              fptr(out["acc"]), uptr(ids)), "lf_download")
        out["ids"] = ids
        return out

    # -- stepping -----------------------------------------------------------
    def step(self, dt: float, substeps: int = 1) -> None:
        check(lib.lf_step(self._h, C.c_float(dt), int(substeps)), "lf_step")

    def run(self, dt: float, steps: int) -> None:
        """Enqueue `steps` substeps of length dt each (lf_step(dt, 1) per step)."""
        for _ in range(int(steps)):
            check(lib.lf_step(self._h, C.c_float(dt), 1), "lf_step")

    def sync(self) -> None:
        check(lib.lf_sync(self._h), "lf_sync")

    def step_count(self) -> int:
        c = C.c_uint64(0)
        check(lib.lf_step_count(self._h, C.byref(c)), "lf_step_count")
        return c.value

    # -- inspection ---------------------------------------------------------
    def set_debug(self, forces: bool = True, scan: bool = False, normals_pass: bool = False) -> None:
        """forces: keep the four force components; scan: test every stencil
        slot in the normals, forces and XSPH passes instead of walking the
        density pass's neighbour masks; normals_pass: run the normals pass
        every substep even without surface tension (otherwise they are
        computed when read)."""
        flags = (L.DEBUG_FORCES if forces else 0) | (L.DEBUG_SCAN if scan else 0)
        flags |= L.DEBUG_NORMALS_PASS if normals_pass else 0
        check(lib.lf_set_debug(self._h, flags), "lf_set_debug")

    def field(self, which: int) -> np.ndarray:
        n, nc = self.n, self.params.ncells
        shapes = {
            L.FIELD_KEYS: (np.uint32, (n,)),
            L.FIELD_START: (np.uint32, (nc + 2,)),
            L.FIELD_SORTED_IDS: (np.uint32, (n,)),
            L.FIELD_DENSITY: (np.float32, (n, 4)),
            L.FIELD_NORMAL: (np.float32, (n, 4)),
            L.FIELD_FORCES: (np.float32, (n, 4, 4)),
            L.FIELD_VEL_HALF: (np.float32, (n, 4)),
            L.FIELD_SORTED_POS: (np.float32, (n, 4)),
        }
        dt, shape = shapes[which]
        a = np.empty(shape, dt)
        check(lib.lf_get_field(self._h, int(which), a.ctypes.data_as(C.c_void_p), a.nbytes), "lf_get_field")
        return a

    def neighbours(self):
        """(counts by id, concatenated ascending neighbour ids) of the density pass."""
        counts = np.empty(self.n, np.uint32)
        total = C.c_uint64(0)
        check(lib.lf_neighbours(self._h, uptr(counts), None, 0, C.byref(total)), "lf_neighbours")
        lst = np.empty(max(total.value, 1), np.uint32)
        check(lib.lf_neighbours(self._h, uptr(counts), uptr(lst), lst.size, C.byref(total)), "lf_neighbours")
        return counts, lst[: total.value]

    # -- timing -------------------------------------------------------------
    def profile(self, kernels=()) -> None:
        mask = 0
        for k in kernels:
            mask |= 1 << L.K_INDEX[k]
        check(lib.lf_profile_enable(self._h, mask), "lf_profile_enable")

    def profile_read(self) -> dict:
        cnt = (C.c_uint64 * len(L.KERNELS))()
        ms = (C.c_double * len(L.KERNELS))()
        check(lib.lf_profile_read(self._h, cnt, ms), "lf_profile_read")
        return {k: (int(cnt[i]), float(ms[i])) for i, k in enumerate(L.KERNELS)}

    def copy_probe(self, nbytes: int, repeats: int = 10) -> float:
        ms = C.c_double(0.0)
        check(lib.lf_copy_probe(self._h, int(nbytes), int(repeats), C.byref(ms)), "lf_copy_probe")
        return ms.value

    # -- formats either side of the path -------------------------------------
    def checkpoint_save(self, path: str) -> None:
        check(lib.lf_checkpoint_save(self._h, str(path).encode()), "lf_checkpoint_save")

    @classmethod
    def checkpoint_load(cls, path: str) -> "Scene":
        self = cls.__new__(cls)
        self._h = C.c_void_p()
        check(lib.lf_checkpoint_load(str(path).encode(), C.byref(self._h)), "lf_checkpoint_load")
        n = C.c_uint32(0)
        check(lib.lf_num_particles(self._h, C.byref(n)), "lf_num_particles")
        self.n = n.value
        self.params = LfParams()
        check(lib.lf_get_params(self._h, C.byref(self.params)), "lf_get_params")
        return self

    def setup_jittered_lattice(self, nx: int, ny: int, nz: int, dx: float, origin, jitter: float, seed: int,
                               first: int = 0) -> None:
        o = np.ascontiguousarray(origin, dtype=np.float32)
        check(lib.lf_setup_jittered_lattice(self._h, int(first), int(nx), int(ny), int(nz), C.c_float(dx), fptr(o),
                                            C.c_float(jitter), C.c_uint64(seed)), "lf_setup_jittered_lattice")

    def setup_fluid_block(self, n: int, mn, mx, first: int = 0) -> tuple:
        """(points written, mass, particle size)"""
        a = np.ascontiguousarray(mn, dtype=np.float32)
        b = np.ascontiguousarray(mx, dtype=np.float32)
        w, m, s = C.c_uint32(0), C.c_float(0.0), C.c_float(0.0)
        check(lib.lf_setup_fluid_block(self._h, int(first), int(n), fptr(a), fptr(b), C.byref(w), C.byref(m),
                                       C.byref(s)), "lf_setup_fluid_block")
        return w.value, m.value, s.value

    def export_geo_async(self, filebase: str, cnt: int, order: int = L.LF_ORDER_CELL) -> None:
        check(lib.lf_export_geo_async(self._h, str(filebase).encode(), int(cnt), int(order)), "lf_export_geo_async")

    def export_wait(self) -> None:
        check(lib.lf_export_wait(self._h), "lf_export_wait")

    # -- launch shapes ------------------------------------------------------
    def set_tuning(self, block_sizes) -> None:
        """Workgroup sizes of the density, normals, forces and XSPH passes."""
        bs = (C.c_int * 4)(*[int(b) for b in block_sizes])
        check(lib.lf_set_tuning(self._h, bs), "lf_set_tuning")

    def tuning(self) -> tuple:
        bs = (C.c_int * 4)()
        check(lib.lf_get_tuning(self._h, bs), "lf_get_tuning")
        return tuple(bs)

    def set_walks(self, flat) -> None:
        """Walk modes of the normals, forces and XSPH passes (1 flattened,
        0 row by row, 2 flattened in waves with a row of more than 32
        slots)."""
        f = (C.c_int * 3)(*[int(x) for x in flat])
        check(lib.lf_set_walks(self._h, f), "lf_set_walks")

    def walks(self) -> tuple:
        f = (C.c_int * 3)()
        check(lib.lf_get_walks(self._h, f), "lf_get_walks")
        return tuple(f)

    def autotune(self, reps: int = 5) -> np.ndarray:
        """Keeps the fastest block size per pass, then the fastest walk mode per
        pass; returns the (4, 3) ms table of the block-size sweep (passes
        density/normals/forces/xsph x block sizes 128/256/512)."""
        ms = (C.c_double * 12)()
        check(lib.lf_autotune(self._h, int(reps), ms), "lf_autotune")
        return np.array(ms[:], dtype=np.float64).reshape(4, 3)

    def set_graphs(self, on: bool) -> None:
        check(lib.lf_set_graphs(self._h, 1 if on else 0), "lf_set_graphs")


class Batch:
    """Many independent scenes stepped together (lf_batch_*)."""

    def __init__(self, max_particles: int, max_cells: int, max_scenes: int):
        self._h = C.c_void_p()
        check(lib.lf_batch_create(int(max_particles), int(max_cells), int(max_scenes), C.byref(self._h)),
              "lf_batch_create")
        self._n = {}

    @staticmethod
    def footprint(max_particles: int, max_cells: int, max_scenes: int) -> int:
        b = C.c_uint64(0)
        check(lib.lf_batch_footprint(int(max_particles), int(max_cells), int(max_scenes), C.byref(b)),
              "lf_batch_footprint")
        return b.value

    def close(self) -> None:
        if self._h and self._h.value:
            lib.lf_batch_destroy(self._h)
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

    def add(self, params: LfParams, num_particles: int, dt: float) -> int:
        sid = C.c_uint32(0)
        check(lib.lf_batch_add(self._h, C.byref(params), int(num_particles), C.c_float(dt), C.byref(sid)),
              "lf_batch_add")
        self._n[sid.value] = int(num_particles)
        return sid.value

    def try_add(self, params: LfParams, num_particles: int, dt: float):
        """Like add, but returns None when the batch has no room."""
        sid = C.c_uint32(0)
        rc = lib.lf_batch_add(self._h, C.byref(params), int(num_particles), C.c_float(dt), C.byref(sid))
        if rc == L.LF_ERR_FULL:
            return None
        check(rc, "lf_batch_add")
        self._n[sid.value] = int(num_particles)
        return sid.value

    def remove(self, scene: int) -> None:
        check(lib.lf_batch_remove(self._h, int(scene)), "lf_batch_remove")
        self._n.pop(int(scene), None)

    def set_params(self, scene: int, params: LfParams) -> None:
        check(lib.lf_batch_set_params(self._h, int(scene), C.byref(params)), "lf_batch_set_params")

    def set_dt(self, scene: int, dt: float) -> None:
        check(lib.lf_batch_set_dt(self._h, int(scene), C.c_float(dt)), "lf_batch_set_dt")

    def set_active(self, scene: int, active: bool) -> None:
        check(lib.lf_batch_set_active(self._h, int(scene), 1 if active else 0), "lf_batch_set_active")

    def upload(self, scene: int, pos: np.ndarray, vel: np.ndarray | None = None, acc: np.ndarray | None = None,
               first: int = 0) -> None:
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float32).reshape(-1, 3)
        acc = None if acc is None else np.ascontiguousarray(acc, dtype=np.float32).reshape(-1, 3)
        check(lib.lf_batch_upload(self._h, int(scene), int(first), pos.shape[0], fptr(pos), fptr(vel), fptr(acc)),
              "lf_batch_upload")

    def download(self, scene: int, order: int = L.LF_ORDER_ID) -> dict:
        n = self._n[int(scene)]
        out = {k: np.empty((n, 3), np.float32) for k in ("pos", "colour", "vel", "acc")}
        ids = np.empty(n, np.uint32)
        check(lib.lf_batch_download(self._h, int(scene), int(order), fptr(out["pos"]), fptr(out["colour"]),
                                    fptr(out["vel"]), fptr(out["acc"]), uptr(ids)), "lf_batch_download")
        out["ids"] = ids
        return out

    def step(self, substeps: int = 1) -> None:
        check(lib.lf_batch_step(self._h, int(substeps)), "lf_batch_step")

    def sync(self) -> None:
        check(lib.lf_batch_sync(self._h), "lf_batch_sync")

    def scene_info(self, scene: int) -> dict:
        n, steps, act = C.c_uint32(0), C.c_uint64(0), C.c_int(0)
        check(lib.lf_batch_scene_info(self._h, int(scene), C.byref(n), C.byref(steps), C.byref(act)),
              "lf_batch_scene_info")
        return {"particles": n.value, "steps": steps.value, "active": bool(act.value)}

    def usage(self) -> dict:
        v = [C.c_uint32(0) for _ in range(5)]
        check(lib.lf_batch_usage(self._h, *[C.byref(x) for x in v]), "lf_batch_usage")
        return dict(zip(("scenes", "slots_used", "slot_cap", "buckets_used", "bucket_cap"), [x.value for x in v]))

    def checkpoint_save(self, scene: int, path: str) -> None:
        check(lib.lf_batch_checkpoint_save(self._h, int(scene), str(path).encode()), "lf_batch_checkpoint_save")

    def checkpoint_load(self, path: str) -> int:
        sid = C.c_uint32(0)
        check(lib.lf_batch_checkpoint_load(self._h, str(path).encode(), C.byref(sid)), "lf_batch_checkpoint_load")
        n = C.c_uint32(0)
        check(lib.lf_batch_scene_info(self._h, sid.value, C.byref(n), None, None), "lf_batch_scene_info")
        self._n[sid.value] = n.value
        return sid.value

    def setup_jittered_lattice(self, scene: int, nx: int, ny: int, nz: int, dx: float, origin, jitter: float,
                               seed: int, first: int = 0) -> None:
        o = np.ascontiguousarray(origin, dtype=np.float32)
        check(lib.lf_batch_setup_jittered_lattice(self._h, int(scene), int(first), int(nx), int(ny), int(nz),
                                                  C.c_float(dx), fptr(o), C.c_float(jitter), C.c_uint64(seed)),
              "lf_batch_setup_jittered_lattice")

    def export_geo_async(self, scene: int, filebase: str, cnt: int, order: int = L.LF_ORDER_CELL) -> None:
        check(lib.lf_batch_export_geo_async(self._h, int(scene), str(filebase).encode(), int(cnt), int(order)),
              "lf_batch_export_geo_async")

    def export_wait(self) -> None:
        check(lib.lf_batch_export_wait(self._h), "lf_batch_export_wait")

    def set_tuning(self, block_sizes) -> None:
        bs = (C.c_int * 4)(*[int(b) for b in block_sizes])
        check(lib.lf_batch_set_tuning(self._h, bs), "lf_batch_set_tuning")

    def tuning(self) -> tuple:
        bs = (C.c_int * 4)()
        check(lib.lf_batch_get_tuning(self._h, bs), "lf_batch_get_tuning")
        return tuple(bs)

    def set_walks(self, flat) -> None:
        f = (C.c_int * 3)(*[int(x) for x in flat])
        check(lib.lf_batch_set_walks(self._h, f), "lf_batch_set_walks")

    def walks(self) -> tuple:
        f = (C.c_int * 3)()
        check(lib.lf_batch_get_walks(self._h, f), "lf_batch_get_walks")
        return tuple(f)

    def autotune(self, reps: int = 5) -> np.ndarray:
        ms = (C.c_double * 12)()
        check(lib.lf_batch_autotune(self._h, int(reps), ms), "lf_batch_autotune")
        return np.array(ms[:], dtype=np.float64).reshape(4, 3)

    def set_graphs(self, on: bool) -> None:
        check(lib.lf_batch_set_graphs(self._h, 1 if on else 0), "lf_batch_set_graphs")

    def profile(self, kernels=()) -> None:
        mask = 0
        for k in kernels:
            mask |= 1 << L.K_INDEX[k]
        check(lib.lf_batch_profile_enable(self._h, mask), "lf_batch_profile_enable")

    def profile_read(self) -> dict:
        cnt = (C.c_uint64 * len(L.KERNELS))()
        ms = (C.c_double * len(L.KERNELS))()
        check(lib.lf_batch_profile_read(self._h, cnt, ms), "lf_batch_profile_read")
        return {k: (int(cnt[i]), float(ms[i])) for i, k in enumerate(L.KERNELS)}


# -- host-side generators (no device needed) ----------------------------------

def fluid_block(n: int, mn, mx, rest_density: float):
    """createFluidBlock: (points[written,3], mass, particle_size)."""
    pts = np.zeros((n, 3), np.float32)
    w = C.c_uint32(0)
    m = C.c_float(0)
    ps = C.c_float(0)
    a = np.asarray(mn, np.float32)
    b = np.asarray(mx, np.float32)
    check(lib.lf_fluid_block(int(n), fptr(a), fptr(b), C.c_float(rest_density), fptr(pts), C.byref(w),
                             C.byref(m), C.byref(ps)), "lf_fluid_block")
    return pts[: w.value].copy(), m.value, ps.value


def jittered_lattice(nx: int, ny: int, nz: int, dx: float, origin, jitter: float, seed: int) -> np.ndarray:
    out = np.empty((nx * ny * nz, 3), np.float32)
    o = np.asarray(origin, np.float32)
    check(lib.lf_jittered_lattice(nx, ny, nz, C.c_float(dx), fptr(o), C.c_float(jitter), C.c_uint64(seed),
                                  fptr(out)), "lf_jittered_lattice")
    return out


def uniform_points(n: int, lo, hi, seed: int) -> np.ndarray:
    out = np.empty((n, 3), np.float32)
    a = np.asarray(lo, np.float32)
    b = np.asarray(hi, np.float32)
    check(lib.lf_uniform_points(int(n), fptr(a), fptr(b), C.c_uint64(seed), fptr(out)), "lf_uniform_points")
    return out


def write_geo(filebase: str, cnt: int, pos, colour, vel, pscale: float) -> None:
    pos = np.ascontiguousarray(pos, np.float32)
    colour = np.ascontiguousarray(colour, np.float32)
    vel = np.ascontiguousarray(vel, np.float32)
    check(lib.lf_write_geo(filebase.encode(), int(cnt), fptr(pos), fptr(colour), fptr(vel), pos.shape[0],
                           C.c_float(pscale)), "lf_write_geo")
