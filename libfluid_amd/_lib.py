"""ctypes binding of libfluid_hip.so (the C ABI declared in include/lf_abi.h).

The shared library is built in-tree (``make -C libfluid_amd``) into
``libfluid_amd/lib/libfluid_hip.so``.  Importing this module fails loudly when
it is missing: there is no CPU fallback for the solver.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libfluid_hip.so")

LF_OK = 0
LF_ERR_INVALID = -1
LF_ERR_HIP = -2
LF_ERR_ALLOC = -3
LF_ERR_STATE = -4
LF_ERR_OVERFLOW = -5
LF_NULL_HASH = 0xFFFFFFFF

LF_ORDER_CELL = 0
LF_ORDER_ID = 1

FIELD_KEYS = 0
FIELD_START = 1
FIELD_SORTED_IDS = 2
FIELD_DENSITY = 3
FIELD_NORMAL = 4
FIELD_FORCES = 5
FIELD_VEL_HALF = 6
FIELD_SORTED_POS = 7

DEBUG_FORCES = 1
DEBUG_SCAN = 2
DEBUG_NORMALS_PASS = 4

LF_ERR_FULL = -6
SERVE_FIFO, SERVE_BACKFILL = 0, 1
JOB_QUEUED, JOB_RUNNING, JOB_DONE = 0, 1, 2

KERNELS = ("hash", "scan", "scatter", "gather", "density", "normals", "forces", "xsph")
K_INDEX = {k: i for i, k in enumerate(KERNELS)}


class LfParams(C.Structure):
    """Mirror of ``lf_params`` (include/lf_abi.h)."""

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

    def as_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else float(v)
        return out

    @property
    def ncells(self) -> int:
        return int(self.grid[0]) * int(self.grid[1]) * int(self.grid[2])


class LfServerStats(C.Structure):
    """Mirror of ``lf_server_stats`` (include/lf_abi.h)."""

    _fields_ = [(k, C.c_uint64) for k in ("submitted", "admitted", "finished", "queued", "running", "ticks",
                                           "step_ticks", "particle_steps")] + [
        ("resident_particles", C.c_uint32),
        ("slot_cap", C.c_uint32),
    ] + [(k, C.c_double) for k in ("busy_seconds", "step_p50_ms", "step_p99_ms", "job_p50_ms", "job_p99_ms",
                                   "wait_p50_ms", "wait_p99_ms", "mean_occupancy")]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class LfError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where} failed ({code}): {msg}")
        self.code = code


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `make -C libfluid_amd` (or __graft_entry__.build())"
        )
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER
    f32p = P(C.c_float)
    u32p = P(C.c_uint32)
    sig = {
        "lf_version": (C.c_int, []),
        "lf_last_error": (C.c_char_p, []),
        "lf_device_count": (C.c_int, [P(C.c_int)]),
        "lf_params_default": (C.c_int, [P(LfParams), C.c_uint32]),
        "lf_params_check": (C.c_int, [P(LfParams)]),
        "lf_create": (C.c_int, [P(LfParams), C.c_uint32, P(C.c_void_p)]),
        "lf_destroy": (C.c_int, [C.c_void_p]),
        "lf_num_particles": (C.c_int, [C.c_void_p, u32p]),
        "lf_set_params": (C.c_int, [C.c_void_p, P(LfParams)]),
        "lf_get_params": (C.c_int, [C.c_void_p, P(LfParams)]),
        "lf_upload": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, f32p, f32p, f32p]),
        "lf_download": (C.c_int, [C.c_void_p, C.c_int, f32p, f32p, f32p, f32p, u32p]),
        "lf_step": (C.c_int, [C.c_void_p, C.c_float, C.c_uint32]),
        "lf_sync": (C.c_int, [C.c_void_p]),
        "lf_step_count": (C.c_int, [C.c_void_p, P(C.c_uint64)]),
        "lf_get_field": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]),
        "lf_set_debug": (C.c_int, [C.c_void_p, C.c_int]),
        "lf_neighbours": (C.c_int, [C.c_void_p, u32p, u32p, C.c_uint64, P(C.c_uint64)]),
        "lf_profile_enable": (C.c_int, [C.c_void_p, C.c_int]),
        "lf_profile_read": (C.c_int, [C.c_void_p, P(C.c_uint64), P(C.c_double)]),
        "lf_copy_probe": (C.c_int, [C.c_void_p, C.c_size_t, C.c_int, P(C.c_double)]),
        "lf_fluid_block": (C.c_int, [C.c_uint32, f32p, f32p, C.c_float, f32p, u32p, f32p, f32p]),
        "lf_jittered_lattice": (
            C.c_int,
            [C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, f32p, C.c_float, C.c_uint64, f32p],
        ),
        "lf_uniform_points": (C.c_int, [C.c_uint32, f32p, f32p, C.c_uint64, f32p]),
        "lf_write_geo": (C.c_int, [C.c_char_p, C.c_uint32, f32p, f32p, f32p, C.c_uint32, C.c_float]),
        "lf_set_tuning": (C.c_int, [C.c_void_p, P(C.c_int)]),
        "lf_get_tuning": (C.c_int, [C.c_void_p, P(C.c_int)]),
        "lf_set_walks": (C.c_int, [C.c_void_p, P(C.c_int)]),
        "lf_get_walks": (C.c_int, [C.c_void_p, P(C.c_int)]),
        "lf_autotune": (C.c_int, [C.c_void_p, C.c_uint32, P(C.c_double)]),
        "lf_set_graphs": (C.c_int, [C.c_void_p, C.c_int]),
        "lf_batch_footprint": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, P(C.c_uint64)]),
        "lf_batch_create": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, P(C.c_void_p)]),
        "lf_batch_destroy": (C.c_int, [C.c_void_p]),
        "lf_batch_add": (C.c_int, [C.c_void_p, P(LfParams), C.c_uint32, C.c_float, u32p]),
        "lf_batch_remove": (C.c_int, [C.c_void_p, C.c_uint32]),
        "lf_batch_set_params": (C.c_int, [C.c_void_p, C.c_uint32, P(LfParams)]),
        "lf_batch_set_dt": (C.c_int, [C.c_void_p, C.c_uint32, C.c_float]),
        "lf_batch_set_active": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int]),
        "lf_batch_upload": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, f32p, f32p, f32p]),
        "lf_batch_download": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, f32p, f32p, f32p, f32p, u32p]),
        "lf_batch_step": (C.c_int, [C.c_void_p, C.c_uint32]),
        "lf_batch_sync": (C.c_int, [C.c_void_p]),
        "lf_batch_scene_info": (C.c_int, [C.c_void_p, C.c_uint32, u32p, P(C.c_uint64), P(C.c_int)]),
        "lf_batch_usage": (C.c_int, [C.c_void_p, u32p, u32p, u32p, u32p, u32p]),
        "lf_batch_set_tuning": (C.c_int, [C.c_void_p, P(C.c_int)]),
        "lf_batch_get_tuning": (C.c_int, [C.c_void_p, P(C.c_int)]),
        "lf_batch_set_walks": (C.c_int, [C.c_void_p, P(C.c_int)]),
        "lf_batch_get_walks": (C.c_int, [C.c_void_p, P(C.c_int)]),
        "lf_batch_autotune": (C.c_int, [C.c_void_p, C.c_uint32, P(C.c_double)]),
        "lf_batch_set_graphs": (C.c_int, [C.c_void_p, C.c_int]),
        "lf_batch_profile_enable": (C.c_int, [C.c_void_p, C.c_int]),
        "lf_batch_profile_read": (C.c_int, [C.c_void_p, P(C.c_uint64), P(C.c_double)]),
        "lf_checkpoint_save": (C.c_int, [C.c_void_p, C.c_char_p]),
        "lf_checkpoint_load": (C.c_int, [C.c_char_p, P(C.c_void_p)]),
        "lf_setup_jittered_lattice": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                                C.c_float, f32p, C.c_float, C.c_uint64]),
        "lf_setup_fluid_block": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, f32p, f32p, u32p, f32p, f32p]),
        "lf_export_geo_async": (C.c_int, [C.c_void_p, C.c_char_p, C.c_uint32, C.c_int]),
        "lf_export_wait": (C.c_int, [C.c_void_p]),
        "lf_batch_checkpoint_save": (C.c_int, [C.c_void_p, C.c_uint32, C.c_char_p]),
        "lf_batch_checkpoint_load": (C.c_int, [C.c_void_p, C.c_char_p, u32p]),
        "lf_batch_setup_jittered_lattice": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                                      C.c_uint32, C.c_float, f32p, C.c_float, C.c_uint64]),
        "lf_batch_setup_fluid_block": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, f32p, f32p, u32p,
                                                 f32p, f32p]),
        "lf_batch_export_geo_async": (C.c_int, [C.c_void_p, C.c_uint32, C.c_char_p, C.c_uint32, C.c_int]),
        "lf_batch_export_wait": (C.c_int, [C.c_void_p]),
        "lf_server_create": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, P(C.c_void_p)]),
        "lf_server_destroy": (C.c_int, [C.c_void_p]),
        "lf_server_submit": (C.c_int, [C.c_void_p, P(LfParams), C.c_uint32, f32p, f32p, C.c_float, C.c_uint32,
                                       P(C.c_uint64)]),
        "lf_server_tick": (C.c_int, [C.c_void_p, u32p]),
        "lf_server_configure": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int]),
        "lf_server_take_finished": (C.c_int, [C.c_void_p, P(C.c_uint64), C.c_uint32, u32p]),
        "lf_server_job_status": (C.c_int, [C.c_void_p, C.c_uint64, P(C.c_int), u32p]),
        "lf_server_job_result": (C.c_int, [C.c_void_p, C.c_uint64, f32p, f32p, f32p, f32p, P(C.c_double)]),
        "lf_server_get_stats": (C.c_int, [C.c_void_p, P(LfServerStats)]),
        "lf_debug_flat_table": (C.c_int, [P(C.c_uint64), u32p, u32p, u32p, u32p]),
        "lf_debug_fail_alloc": (C.c_int, [C.c_int64]),
        "lf_debug_live_allocs": (C.c_int, [P(C.c_int64)]),
        "lf_step_noalloc": (C.c_int, [P(LfParams), C.c_uint32, C.c_float, C.c_uint32, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()

# Every symbol include/lf_abi.h declares.
EXPORTED = (
    "lf_version lf_last_error lf_device_count lf_params_default lf_params_check lf_create "
    "lf_destroy lf_num_particles lf_set_params lf_get_params lf_upload lf_download lf_step "
    "lf_sync lf_step_count lf_get_field lf_set_debug lf_neighbours lf_profile_enable "
    "lf_profile_read lf_copy_probe lf_fluid_block lf_jittered_lattice lf_uniform_points "
    "lf_write_geo lf_set_tuning lf_get_tuning lf_set_walks lf_get_walks lf_autotune lf_set_graphs "
    "lf_batch_footprint lf_batch_create lf_batch_destroy lf_batch_add lf_batch_remove "
    "lf_batch_set_params lf_batch_set_dt lf_batch_set_active lf_batch_upload lf_batch_download "
    "lf_batch_step lf_batch_sync lf_batch_scene_info lf_batch_usage lf_batch_set_tuning "
    "lf_batch_get_tuning lf_batch_set_walks lf_batch_get_walks lf_batch_autotune lf_batch_set_graphs lf_batch_profile_enable "
    "lf_batch_profile_read lf_server_create lf_server_destroy lf_server_submit lf_server_tick "
    "lf_server_job_status lf_server_job_result lf_server_get_stats lf_server_take_finished lf_server_configure "
    "lf_checkpoint_save lf_checkpoint_load lf_setup_jittered_lattice lf_setup_fluid_block lf_export_geo_async "
    "lf_export_wait lf_batch_checkpoint_save lf_batch_checkpoint_load lf_batch_setup_jittered_lattice "
    "lf_batch_setup_fluid_block lf_batch_export_geo_async lf_batch_export_wait "
    "lf_debug_flat_table lf_debug_fail_alloc lf_debug_live_allocs lf_step_noalloc"
).split()


def check(rc: int, where: str) -> None:
    if rc != LF_OK:
        raise LfError(rc, where, lib.lf_last_error().decode(errors="replace"))


def fptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_float))


def uptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


FLAT_ENTRIES = 16  # entries of the flattened walk's table


def flat_table(masks, first_slots):
    """The flattened walk's table for one particle's nine rows, built by the
    kernels' own code on the host: (entry masks, entry first slots, count)."""
    m = np.ascontiguousarray(masks, np.uint64)
    b = np.ascontiguousarray(first_slots, np.uint32)
    assert m.shape == (9,) and b.shape == (9,)
    em = np.zeros(FLAT_ENTRIES, np.uint32)
    eb = np.zeros(FLAT_ENTRIES, np.uint32)
    cnt = C.c_uint32(0)
    check(lib.lf_debug_flat_table(m.ctypes.data_as(C.POINTER(C.c_uint64)), uptr(b), uptr(em), uptr(eb),
                                  C.byref(cnt)), "lf_debug_flat_table")
    return em, eb, int(cnt.value)


def live_allocs() -> int:
    c = C.c_int64(0)
    check(lib.lf_debug_live_allocs(C.byref(c)), "lf_debug_live_allocs")
    return c.value


def device_count() -> int:
    c = C.c_int(0)
    check(lib.lf_device_count(C.byref(c)), "lf_device_count")
    return c.value
