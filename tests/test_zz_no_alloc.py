"""GPU: lf_step_noalloc, the entry point that allocates nothing (its working
buffers live in a fixed device array of the library), against the entry
points it stands for -- lf_create + lf_upload + lf_step + lf_download --
bitwise, on two calls in a row, the second with a smaller scene; and its
size limit.  Device arrays for the caller's state come from the HIP runtime
directly (hipMalloc in the test, not in the library)."""
import ctypes as C
import gc

import numpy as np
import pytest

from libfluid_amd import _lib as L
from libfluid_amd import configs

pytestmark = pytest.mark.gpu

DT, SUBSTEPS = 4e-4, 4


class Hip:
    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipDeviceSynchronize.argtypes = []

    def ok(self, rc, what):
        assert rc == 0, f"{what} failed ({rc})"

    def put(self, a):
        a = np.ascontiguousarray(a, np.float32)
        p = C.c_void_p()
        self.ok(self.rt.hipMalloc(C.byref(p), a.nbytes), "hipMalloc")
        self.ok(self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return p

    def get(self, p, shape):
        out = np.empty(shape, np.float32)
        self.ok(self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def free(self, p):
        self.ok(self.rt.hipFree(p), "hipFree")


def with_scene(gpu, cfg, acc):
    with gpu.Scene(cfg.params, cfg.n) as sc:
        sc.upload(cfg.pos, cfg.vel, acc)
        sc.step(DT, SUBSTEPS)
        sc.sync()
        return sc.download(L.LF_ORDER_ID)


def no_alloc(hip, cfg, acc):
    d = [hip.put(cfg.pos), hip.put(cfg.vel), hip.put(acc)]
    try:
        before = L.live_allocs()
        L.check(L.lib.lf_step_noalloc(C.byref(cfg.params), cfg.n, C.c_float(DT), SUBSTEPS, d[0], d[1], d[2], None),
                "lf_step_noalloc")
        hip.ok(hip.rt.hipDeviceSynchronize(), "hipDeviceSynchronize")
        assert L.live_allocs() == before, "the library allocated device memory"
        return {k: hip.get(p, (cfg.n, 3)) for k, p in zip(("pos", "vel", "acc"), d)}
    finally:
        for p in d:
            hip.free(p)


def test_no_alloc_step_equals_scene_step(gpu):
    gc.collect()
    hip = Hip()
    big = configs.dambreak_c1()          # 32,768 particles, 31,250 cells
    small = configs.scaled_c1(12, 10, 9)  # 1,080 particles in the same box
    small.params.viscosity, small.params.surface_tension, small.params.adhesion = 0.2, 1.0, 0.1
    rng = np.random.Generator(np.random.PCG64(5))
    for cfg in (big, small):
        cfg.vel = rng.normal(0.0, 0.05, cfg.pos.shape).astype(np.float32)
        acc = rng.normal(0.0, 1.0, cfg.pos.shape).astype(np.float32)
        want = with_scene(gpu, cfg, acc)
        got = no_alloc(hip, cfg, acc)
        print(f"\n{cfg.name}: {cfg.n} particles, {SUBSTEPS} substeps")
        for k in ("pos", "vel", "acc"):
            assert np.array_equal(got[k], want[k]), f"{cfg.name}: {k} differs from lf_create + lf_step"


def test_no_alloc_rejects_scenes_beyond_its_array(gpu):
    hip = Hip()
    cfg = configs.column_c2()  # 1,048,576 particles: far beyond the fixed array
    d = [hip.put(np.zeros((16, 3), np.float32)) for _ in range(3)]
    try:
        rc = L.lib.lf_step_noalloc(C.byref(cfg.params), cfg.n, C.c_float(DT), 1, d[0], d[1], d[2], None)
        assert rc == L.LF_ERR_INVALID
        assert b"limit" in L.lib.lf_last_error()
    finally:
        for p in d:
            hip.free(p)
