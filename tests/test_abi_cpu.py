"""CPU tests of the C ABI library (no GPU calls that compute).

* libfluid_hip.so loads and exports every function include/lf_abi.h declares;
* parameter defaults are those of FluidSystem::init (fluidsystem.cu:202-235);
* parameter validation, scene-input generators, the .geo writer's bytes
  (test/src/main.cpp:10-74) and the fluidtest driver's command line.
"""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import libfluid_amd as lf
from libfluid_amd import _lib as L
from libfluid_amd import configs
from libfluid_amd.scene import default_params, fluid_block, jittered_lattice, uniform_points, write_geo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "lf_abi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lf_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 20
    lib = C.CDLL(L.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"
    assert sorted(L.EXPORTED) == names  # the Python binding covers the whole ABI


@pytest.mark.parametrize("ctype,cname", [(L.LfParams, "lf_params"), (L.LfServerStats, "lf_server_stats")])
def test_abi_version_and_struct_layout(tmp_path, ctype, cname):
    """The ctypes mirrors of the ABI structs match the C header (size and offsets)."""
    assert L.lib.lf_version() == 1
    fields = [f for f, _ in ctype._fields_]
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "lf_abi.h"\nint main(void){'
                   + f'printf("%zu", sizeof({cname}));'
                   + "".join(f'printf(" %zu", offsetof({cname}, {f}));' for f in fields) + "return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert out[0] == C.sizeof(ctype)
    assert out[1:] == [getattr(ctype, f).offset for f in fields]


def test_params_default_match_reference_init():
    p = default_params(16)
    assert (p.rest_density, p.gas_constant, p.gamma, p.mass, p.particle_size) == (1000.0, 100.0, 7.0, 1.0, 1.0)
    assert (p.vel_limit, p.acc_limit, p.xsph_scale, p.vdamp) == (5.0, 150.0, 0.5, 0.0)
    # constructor defaults that init() leaves alone (fluidparams.cuh:81-97)
    assert (p.surface_tension, p.viscosity, p.adhesion) == (1.0, np.float32(0.1), 1.0)
    assert list(p.gravity) == [0.0, np.float32(-9.8), 0.0]
    assert p.h == np.float32(1 / 16) and list(p.grid) == [16] * 3 and list(p.inv_cell) == [16.0] * 3
    assert list(p.box) == [1.0] * 3


@pytest.mark.parametrize("field,value", [("h", 0.0), ("h", 0.2), ("particle_size", 0.0), ("mass", -1.0),
                                         ("gamma", float("nan"))])
def test_params_check_rejects(field, value):
    p = default_params(16)
    setattr(p, field, value)
    assert L.lib.lf_params_check(C.byref(p)) == L.LF_ERR_INVALID
    assert L.lib.lf_last_error()


def test_params_check_grid_must_cover_box():
    p = default_params(16)
    p.grid[1] = 8
    assert L.lib.lf_params_check(C.byref(p)) == L.LF_ERR_INVALID
    assert b"cover" in L.lib.lf_last_error()


def test_create_without_device_fails_cleanly():
    if lf.device_count() > 0:
        pytest.skip("a device is visible")
    h = C.c_void_p()
    rc = L.lib.lf_create(C.byref(default_params(8)), 100, C.byref(h))
    assert rc < 0 and not h.value and L.lib.lf_last_error()


def test_batch_create_without_device_fails_cleanly():
    if lf.device_count() > 0:
        pytest.skip("a device is visible")
    h = C.c_void_p()
    rc = L.lib.lf_batch_create(4096, 4096, 4, C.byref(h))
    assert rc < 0 and not h.value and L.lib.lf_last_error()


def test_server_create_without_device_fails_cleanly():
    if lf.device_count() > 0:
        pytest.skip("a device is visible")
    h = C.c_void_p()
    assert L.lib.lf_server_create(4096, 4096, 4, L.SERVE_BACKFILL, C.byref(h)) < 0 and not h.value
    assert L.lib.lf_server_create(4096, 4096, 4, 7, C.byref(h)) == L.LF_ERR_INVALID


def test_served_workload_is_seeded():
    from libfluid_amd.serve import served_jobs
    a = served_jobs(20, rate=10.0, seed=3)
    b = served_jobs(20, rate=10.0, seed=3)
    assert [j.cfg.n for j in a] == [j.cfg.n for j in b] and [j.steps for j in a] == [j.steps for j in b]
    t = np.array([j.arrival for j in a])
    assert np.all(np.diff(t) > 0)
    assert all(4096 <= j.cfg.n <= 262144 and 200 <= j.steps <= 2000 for j in a)


def _checkpoint_bytes(params, n, steps=7, dt=1e-4, corrupt=False):
    """A checkpoint file: 136-byte header (magic, ABI version, count, steps,
    dt, reserved, lf_params, 4 bytes of padding), 12 floats per particle and
    the FNV-1a 64 of everything before."""
    import struct
    head = struct.pack("<8sIIQfI", b"LFCKPT01", 1, n, steps, dt, 0) + bytes(params) + b"\0" * 4
    assert len(head) == 136
    body = np.arange(12 * n, dtype=np.float32).tobytes()
    h = 0xCBF29CE484222325
    for byte in head + body:
        h = ((h ^ byte) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return head + body + struct.pack("<Q", h ^ (1 if corrupt else 0))


@pytest.mark.parametrize("kind,needle", [("magic", b"not a checkpoint"), ("checksum", b"checksum"),
                                         ("truncated", b"truncated"), ("missing", b"cannot open")])
def test_checkpoint_file_validation(tmp_path, kind, needle):
    p = default_params(8)
    data = _checkpoint_bytes(p, 10, corrupt=(kind == "checksum"))
    if kind == "magic":
        data = b"XXXXXXXX" + data[8:]
    elif kind == "truncated":
        data = data[:-20]
    path = tmp_path / "c.lfck"
    if kind != "missing":
        path.write_bytes(data)
    h = C.c_void_p()
    assert L.lib.lf_checkpoint_load(str(path).encode(), C.byref(h)) == L.LF_ERR_INVALID and not h.value
    assert needle in L.lib.lf_last_error()


def test_checkpoint_valid_file_reaches_scene_creation(tmp_path):
    if lf.device_count() > 0:
        pytest.skip("a device is visible")
    path = tmp_path / "c.lfck"
    path.write_bytes(_checkpoint_bytes(default_params(8), 10))
    h = C.c_void_p()
    rc = L.lib.lf_checkpoint_load(str(path).encode(), C.byref(h))
    assert rc < 0 and rc != L.LF_ERR_INVALID and not h.value  # the file passed; no device to load it on


def test_batch_footprint():
    """Bytes a batch allocates: 340 per slot (slots rounded up to 512), 8 per
    bucket, per scene its table entry, flag word and 2 KB move table (16
    slots a cache line apart), per 512-slot chunk 4 for its map entry and 8
    near-isolated mark bytes (one per 64 slots), and 16 per 513 slots for
    the sort's list of large cells."""
    b1 = lf.Batch.footprint(1, 10, 1)
    b2 = lf.Batch.footprint(512, 10, 1)
    assert b1 == b2  # one chunk either way
    b3 = lf.Batch.footprint(1024, 10, 1)
    assert b3 - b2 == 512 * 340 + 4 + 8 + 16
    assert lf.Batch.footprint(512, 20, 1) - b2 == 80
    per_scene = lf.Batch.footprint(512, 10, 2) - b2
    assert 2048 + 200 < per_scene < 2048 + 1024


def test_batched_c4_workload_is_deterministic():
    """256 scenes, counts log-uniform in [2048, 65536], boxes covering their
    blocks, unit gravity directions scaled to 9.8."""
    a = configs.batched_c4()
    b = configs.batched_c4()
    assert len(a.scenes) == 256
    ns = np.array([s.n for s in a.scenes])
    assert ns.min() >= 2048 and ns.max() <= 65536
    assert 3.5e6 < ns.sum() < 6.5e6  # E[n] = (65536 - 2048) / ln 32 ~ 18.3k
    for s, t in zip(a.scenes, b.scenes):
        assert s.n == t.n and np.array_equal(s.pos, t.pos)
        g = np.array(s.params.gravity[:], np.float64)
        assert abs(np.linalg.norm(g) - 9.8) < 1e-4
        box = np.array(s.params.box[:]), np.array(s.params.grid[:]) * 0.02
        assert np.all(s.pos.max(0) < box[0]) and np.all(box[1] >= box[0] - 1e-6)
        assert s.pos.shape == (s.n, 3)


def test_fluid_block_lattice():
    """createFluidBlock: spacing d = min(extent/round(extent/cbrt(V/n))), points
    from min + d/2 in steps of d, x outermost, truncated at n."""
    pts, mass, psize = fluid_block(1000, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 1000.0)
    d = np.float32(1.0) / (np.float32(1.0) / np.cbrt(np.float32(1e-3)) + np.float32(0.5))
    assert psize == pytest.approx(0.5 * d, rel=1e-6)
    assert mass == pytest.approx(1.0, rel=1e-6)  # (V / n) * rho0
    assert pts.shape == (1000, 3)
    np.testing.assert_allclose(pts[0], [0.5 * d] * 3, rtol=1e-6)
    np.testing.assert_allclose(pts[1] - pts[0], [0, 0, d], rtol=1e-5)  # z innermost
    assert np.all(pts <= 1.0) and np.all(pts >= 0.0)
    # fewer lattice sites than requested: every site written once
    pts2, _, _ = fluid_block(50, (0.0, 0.0, 0.0), (0.1, 0.1, 0.1), 1000.0)
    assert len(np.unique(pts2, axis=0)) == pts2.shape[0]


def test_jittered_lattice_seeded():
    a = jittered_lattice(4, 5, 6, 0.01, (0.1, 0.2, 0.3), 0.0005, 42)
    b = jittered_lattice(4, 5, 6, 0.01, (0.1, 0.2, 0.3), 0.0005, 42)
    c = jittered_lattice(4, 5, 6, 0.01, (0.1, 0.2, 0.3), 0.0005, 43)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    i, j, k = np.meshgrid(np.arange(4), np.arange(5), np.arange(6), indexing="ij")
    base = np.stack([0.1 + (i + 0.5) * 0.01, 0.2 + (j + 0.5) * 0.01, 0.3 + (k + 0.5) * 0.01], -1).reshape(-1, 3)
    dev = np.abs(a - base)
    assert dev.max() <= 0.0005 + 1e-7 and dev.max() > 0.0004


def test_uniform_points_in_range():
    p = uniform_points(10000, (0.0, 1.5, 0.0), (2.56, 2.5, 2.56), 7)
    assert np.all(p >= [0.0, 1.5, 0.0]) and np.all(p < [2.56, 2.5, 2.56])
    assert abs(p[:, 1].mean() - 2.0) < 0.02


def _expected_geo(pos, col, vel, pscale):
    """The byte layout of dumpToGeo (test/src/main.cpp:29-70); C++ default
    float output is %g with 6 significant digits."""
    g = lambda v: "%g" % float(v)  # noqa: E731
    n = len(pos)
    out = ["PGEOMETRY V5\n", f"NPoints {n} NPrims 1\n", "NPointGroups 0 NPrimGroups 1\n",
           "NPointAttrib 3  NVertexAttrib 0 NPrimAttrib 2 NAttrib 0\n", "PointAttrib \n",
           "Cd 3 float 1 1 1\n", "v 3 float 0 0 0\n", "pscale 1 float 1\n"]
    for p, c, v in zip(pos, col, vel):
        out.append(f"{g(p[0])} {g(p[1])} {g(p[2])} 1 ({g(c[0])} {g(c[1])} {g(c[2])} "
                   f"{g(v[0])} {g(v[1])} {g(v[2])} {g(pscale)})\n")
    out += ["PrimitiveAttrib\n", "generator 1 index 1 location1\n", "dopobject 1 index 1 /obj/AutoDopNetwork:1\n",
            f"Part {n} " + "".join(f"{i} " for i in range(n)) + " [0\t0]\n", "box_object1 unordered\n", "1 1\n",
            "beginExtra\n", "endExtra\n"]
    return "".join(out)


def test_geo_writer_bytes(tmp_path):
    rng = np.random.default_rng(3)
    pos = rng.uniform(-1, 1, (7, 3)).astype(np.float32)
    col = rng.uniform(0, 1, (7, 3)).astype(np.float32)
    vel = (rng.normal(0, 1e-3, (7, 3))).astype(np.float32)
    vel[0] = [1e-7, -123456.0, 0.0]
    base = str(tmp_path / "sim")
    write_geo(base, 12, pos, col, vel, np.float32(0.0125))
    text = open(base + ".0012.geo").read()
    assert text == _expected_geo(pos, col, vel, np.float32(0.0125))


def test_fluidtest_driver_cli():
    exe = os.path.join(ROOT, "libfluid_amd", "bin", "fluidtest")
    r = subprocess.run([exe, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    for flag in ("-n", "-g", "-a", "-v", "-s", "-d", "-t", "-p", "-q", "-f"):
        assert f"  {flag}, --" in r.stdout
    r = subprocess.run([exe, "-n", "100"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "filebase" in r.stderr


def test_python_mirror_has_reference_api():
    public = ["setup", "setupFromData", "exportToData", "getPscale", "advance", "setRestDensity",
              "setParticleSize", "setGasConstant", "setLimits", "setXSPHScale", "setSurfaceTension",
              "setViscosity", "setAdhesion", "setVDamp", "setGamma", "setMass"]
    protected = ["init", "clear", "createFluidBlock"]
    for name in public + protected:
        assert callable(getattr(lf.FluidSystem, name)), name
    header = open(os.path.join(ROOT, "include", "fluidsystem.h")).read()
    for name in public + protected:
        assert re.search(rf"\b{name}\s*\(", header), name


def _flat_reference(masks, first):
    """The table the flattened walk must see: every non-empty 32-slot
    pseudo-row in row order (low word, then high word 32 slots later)."""
    out = []
    for r in range(9):
        for half in range(2):
            w = (int(masks[r]) >> (32 * half)) & 0xFFFFFFFF
            if w:
                out.append((w, int(first[r]) + 32 * half))
    return out


@pytest.mark.parametrize("case", ["sixteen_then_empty", "sixteen_then_short_rows", "seventeen", "mixed", "empty"])
def test_flat_walk_table(case):
    """The flattened walk's table (built by the kernels' own code) holds
    exactly the non-empty pseudo-rows.  Exactly 16 of them followed by
    empty ones (an empty row, or the zero high word of a row of at most 32
    slots) used to overwrite entry 0 with an empty mask, so the walk of
    such a particle visited no neighbour."""
    first = np.arange(9, dtype=np.uint32) * 1000 + 7
    masks = np.zeros(9, np.uint64)
    if case == "sixteen_then_empty":  # rows 0-7: 33..64 slots, both words set; row 8 empty
        masks[:8] = np.uint64((0x5 << 32) | 0x3)
    elif case == "sixteen_then_short_rows":  # 16 entries, then a row of <= 32 slots with no neighbour
        # 7 rows of 33..64 slots (14 entries), then two rows of <= 32 slots:
        # entries 15 and 16, each followed by its empty high word
        masks[:7] = np.uint64((0x80000000 << 32) | 0x1)
        masks[7] = np.uint64(0x10)
        masks[8] = np.uint64(0x20)
    elif case == "seventeen":
        masks[:8] = np.uint64((0x5 << 32) | 0x3)
        masks[8] = np.uint64(0x1)
    elif case == "mixed":
        masks[[0, 2, 5, 8]] = [np.uint64(0xF0F0), np.uint64(1 << 40), np.uint64((3 << 32) | 9), np.uint64(0x2)]
    em, eb, cnt = L.flat_table(masks, first)
    want = _flat_reference(masks, first)
    assert cnt == len(want)
    k = min(cnt, L.FLAT_ENTRIES)
    assert [(int(a), int(b)) for a, b in zip(em[:k], eb[:k])] == want[:k]
    assert all(int(a) != 0 for a in em[:k])
    if case.startswith("sixteen"):
        assert cnt == 16
    if case == "seventeen":
        assert cnt == 17  # above the table: the kernels walk such a wave row by row


def test_bench_measures_one_gpu_only():
    """bench.py --gpus N with N > 1 exits at once with a message (the path is
    measured on one GPU; nothing is sharded), before touching a device."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run(["python3", os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stdout + r.stderr
    assert "{" not in r.stdout  # no JSON line
