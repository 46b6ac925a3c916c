"""ctypes binding of the FP64 CPU oracle (oracle/lib/liblf_oracle*.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PIN_F32 = 1
XSPH_F32_POS = 2  # XSPH weights from FP32-rounded integrated positions (storage floor)


class OrcParams(C.Structure):
    """Mirror of orc_params (oracle/lf_oracle.h); same layout as lf_params."""

    _fields_ = [
        ("rest_density", C.c_float),
        ("gas_constant", C.c_float),
        ("gamma", C.c_float),
        ("mass", C.c_float),
        ("particle_size", C.c_float),
        ("xsph_scale", C.c_float),
        ("surface_tension", C.c_float),
        ("viscosity", C.c_float),
        ("adhesion", C.c_float),
        ("vel_limit", C.c_float),
        ("acc_limit", C.c_float),
        ("vdamp", C.c_float),
        ("gravity", C.c_float * 3),
        ("h", C.c_float),
        ("inv_cell", C.c_float * 3),
        ("grid", C.c_uint32 * 3),
        ("box", C.c_float * 3),
    ]

    @property
    def ncells(self) -> int:
        return int(self.grid[0]) * int(self.grid[1]) * int(self.grid[2])


class OrcOut(C.Structure):
    _fields_ = [
        ("keys", C.POINTER(C.c_uint32)),
        ("order", C.POINTER(C.c_uint32)),
        ("start", C.POINTER(C.c_uint32)),
        ("density", C.POINTER(C.c_double)),
        ("pressure", C.POINTER(C.c_double)),
        ("sound_speed", C.POINTER(C.c_double)),
        ("normal", C.POINTER(C.c_double)),
        ("f_pressure", C.POINTER(C.c_double)),
        ("f_viscosity", C.POINTER(C.c_double)),
        ("f_tension", C.POINTER(C.c_double)),
        ("f_adhesion", C.POINTER(C.c_double)),
        ("vel_half", C.POINTER(C.c_double)),
        ("sdf_after", C.POINTER(C.c_double)),
    ]


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def to_orc(params) -> OrcParams:
    """Copy any structure with the lf_params layout into OrcParams."""
    o = OrcParams()
    assert C.sizeof(params) == C.sizeof(o)
    C.memmove(C.byref(o), C.byref(params), C.sizeof(o))
    return o


def _u32(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def _f32(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _f64(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    def __init__(self, omp: bool = False, path: str | None = None):
        """The built oracle (single-threaded or OpenMP), or the library at
        `path` (a build of a modified copy, used to check that the tests pin
        the oracle)."""
        if path is None:
            name = "liblf_oracle_omp.so" if omp else "liblf_oracle.so"
            path = os.path.join(_HERE, "lib", name)
            if not os.path.exists(path):
                build()
        lib = C.CDLL(path)
        P = C.POINTER
        op = P(OrcParams)
        for fn, res, args in [
            ("orc_version", C.c_int, []),
            ("orc_num_threads", C.c_int, []),
            ("orc_num_cells", C.c_uint32, [op]),
            ("orc_hash", C.c_int, [op, P(C.c_float), C.c_uint32, P(C.c_uint32), P(C.c_uint32)]),
            ("orc_sort", C.c_int, [P(C.c_uint32), C.c_uint32, P(C.c_uint32)]),
            ("orc_cell_start", C.c_int, [op, P(C.c_uint32), C.c_uint32, P(C.c_uint32)]),
            ("orc_neighbours", C.c_int64, [op, P(C.c_float), C.c_uint32, P(C.c_uint32), P(C.c_uint32), C.c_uint64]),
            ("orc_neighbours_brute", C.c_int64,
             [op, P(C.c_float), C.c_uint32, C.c_int, P(C.c_uint32), P(C.c_uint32), C.c_uint64]),
            ("orc_step", C.c_int, [op, C.c_double, C.c_uint32, P(C.c_double), P(C.c_double), P(C.c_double),
                                   C.c_uint32, P(OrcOut)]),
            ("orc_poly6", C.c_double, [op, C.c_double]),
            ("orc_grad_poly6", C.c_double, [op, C.c_double]),
            ("orc_grad_spiky", C.c_double, [op, C.c_double]),
            ("orc_cohesion", C.c_double, [op, C.c_double]),
            ("orc_boundary_density", C.c_double, [op, C.c_double]),
            ("orc_sdf", C.c_double, [op, P(C.c_double)]),
            ("orc_sdf_normal", None, [op, P(C.c_double), P(C.c_double)]),
            ("orc_colour", None, [op, C.c_double, P(C.c_double)]),
            ("orc_colours", None, [op, C.c_uint32, P(C.c_double), P(C.c_double)]),
        ]:
            f = getattr(lib, fn)
            f.restype = res
            f.argtypes = args
        self.lib = lib
        self.threads = lib.orc_num_threads()

    # -- FP32 replay stages ---------------------------------------------------
    def hash(self, params, pos32: np.ndarray):
        p = to_orc(params)
        pos32 = np.ascontiguousarray(pos32, np.float32)
        n = pos32.shape[0]
        keys = np.empty(n, np.uint32)
        counts = np.empty(p.ncells + 1, np.uint32)
        self.lib.orc_hash(C.byref(p), _f32(pos32), n, _u32(keys), _u32(counts))
        return keys, counts

    def sort(self, keys: np.ndarray) -> np.ndarray:
        keys = np.ascontiguousarray(keys, np.uint32)
        order = np.empty_like(keys)
        self.lib.orc_sort(_u32(keys), keys.size, _u32(order))
        return order

    def cell_start(self, params, keys: np.ndarray) -> np.ndarray:
        p = to_orc(params)
        keys = np.ascontiguousarray(keys, np.uint32)
        start = np.empty(p.ncells + 2, np.uint32)
        self.lib.orc_cell_start(C.byref(p), _u32(keys), keys.size, _u32(start))
        return start

    def neighbours(self, params, pos32: np.ndarray, brute: bool = False, stencil: bool = True):
        p = to_orc(params)
        pos32 = np.ascontiguousarray(pos32, np.float32)
        n = pos32.shape[0]
        counts = np.empty(n, np.uint32)
        cap = 1 << 16
        while True:
            lst = np.empty(cap, np.uint32)
            if brute:
                tot = self.lib.orc_neighbours_brute(C.byref(p), _f32(pos32), n, 1 if stencil else 0, _u32(counts),
                                                    _u32(lst), cap)
            else:
                tot = self.lib.orc_neighbours(C.byref(p), _f32(pos32), n, _u32(counts), _u32(lst), cap)
            if tot >= 0:
                return counts, lst[:tot].copy()
            cap *= 4

    # -- FP64 step ------------------------------------------------------------
    def step(self, params, dt: float, pos: np.ndarray, vel: np.ndarray, acc: np.ndarray, flags: int = 0,
             want=("density", "pressure", "sound_speed", "normal", "f_pressure", "f_viscosity", "f_tension",
                   "f_adhesion", "keys", "order", "start", "vel_half", "sdf_after")) -> dict:
        """Advance float64 (n,3) pos/vel/acc in place by one substep; returns requested fields."""
        p = to_orc(params)
        for a in (pos, vel, acc):
            assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
        n = pos.shape[0]
        out = OrcOut()
        res = {}
        shapes = {"keys": (np.uint32, (n,)), "order": (np.uint32, (n,)), "start": (np.uint32, (p.ncells + 2,)),
                  "density": (np.float64, (n,)), "pressure": (np.float64, (n,)),
                  "sound_speed": (np.float64, (n,)), "normal": (np.float64, (n, 3)),
                  "f_pressure": (np.float64, (n, 3)), "f_viscosity": (np.float64, (n, 3)),
                  "f_tension": (np.float64, (n, 3)), "f_adhesion": (np.float64, (n, 3)),
                  "vel_half": (np.float64, (n, 3)), "sdf_after": (np.float64, (n,))}
        for k in want:
            dt_, shp = shapes[k]
            arr = np.empty(shp, dt_)
            res[k] = arr
            setattr(out, k, _u32(arr) if dt_ == np.uint32 else _f64(arr))
        rc = self.lib.orc_step(C.byref(p), float(dt), n, _f64(pos), _f64(vel), _f64(acc), flags, C.byref(out))
        assert rc == 0
        return res

    # -- kernels for known-answer tests -----------------------------------------
    def poly6(self, params, r):
        return self.lib.orc_poly6(C.byref(to_orc(params)), float(r))

    def grad_poly6(self, params, r):
        return self.lib.orc_grad_poly6(C.byref(to_orc(params)), float(r))

    def grad_spiky(self, params, r):
        return self.lib.orc_grad_spiky(C.byref(to_orc(params)), float(r))

    def cohesion(self, params, r):
        return self.lib.orc_cohesion(C.byref(to_orc(params)), float(r))

    def boundary_density(self, params, d):
        return self.lib.orc_boundary_density(C.byref(to_orc(params)), float(d))

    def sdf(self, params, x):
        x = np.ascontiguousarray(x, np.float64)
        return self.lib.orc_sdf(C.byref(to_orc(params)), _f64(x))

    def sdf_normal(self, params, x):
        x = np.ascontiguousarray(x, np.float64)
        n = np.empty(3, np.float64)
        self.lib.orc_sdf_normal(C.byref(to_orc(params)), _f64(x), _f64(n))
        return n

    def colour(self, params, density):
        rgb = np.empty(3, np.float64)
        self.lib.orc_colour(C.byref(to_orc(params)), float(density), _f64(rgb))
        return rgb

    def colours(self, params, density):
        """DensityColourOperator of every density: (n, 3)."""
        d = np.ascontiguousarray(density, np.float64)
        rgb = np.empty((d.size, 3), np.float64)
        self.lib.orc_colours(C.byref(to_orc(params)), d.size, _f64(d), _f64(rgb))
        return rgb
