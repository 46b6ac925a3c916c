"""GPU: the reference's own caller path, end to end.

* bin/fluidtest (the driver, test/src/main.cpp:76-158): MyFluidSystem's
  double dam break (test/src/myfluidsystem.cpp:6-30) through the C++
  FluidSystem class and the C ABI; its last .geo frame is byte for byte the
  one the ctypes path writes for the same scene, and that state matches the
  FP64 oracle.
* FluidSystem::setup (fluidsystem.cu:132-147) and setupFromData
  (fluidsystem.cu:179-189) through the C++ class (bin/apicheck) and the
  Python mirror (libfluid_amd.FluidSystem): bitwise the same, and against
  the oracle; setupFromData keeps mass and particle size at init's 1 (Q27).
"""
import os
import subprocess

import numpy as np
import pytest

import libfluid_amd as lf
from libfluid_amd import _lib as L
from libfluid_amd import configs
from tests.helpers import normwise

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "libfluid_amd", "bin")
TOL = 1e-5


def oracle_run(orc, params, pos, vel, dt, steps):
    p = np.array(pos, np.float64)
    v = np.array(vel, np.float64)
    a = np.zeros_like(p)
    o = None
    for _ in range(steps):
        o = orc.step(params, float(np.float32(dt)), p, v, a, 0, want=("density",))
    return p, v, o


def read_apicheck(path):
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw[:4], np.uint32)[0])
    pscale, mass = np.frombuffer(raw[4:12], np.float32)
    arr = np.frombuffer(raw[12:], np.float32).reshape(4, n, 3)
    return {"n": n, "pscale": float(pscale), "mass": float(mass), "pos": arr[0], "colour": arr[1], "vel": arr[2],
            "pos_by_id": arr[3]}


def test_fluidtest_driver_double_dam(gpu, oracle_omp, tmp_path):
    n, res, steps, dt = 4096, 16, 5, 0.002
    base = str(tmp_path / "sim")
    r = subprocess.run([os.path.join(BIN, "fluidtest"), "-n", str(n), "-g", str(res), "-q", str(steps), "-f", base],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.count("---Step ") == steps and "TOTAL TIME" in r.stdout
    for k in range(1, steps + 1):
        assert os.path.exists(f"{base}.{k:04d}.geo")
    # The same scene through the ctypes path: the reference driver's
    # coefficients (a=0.1, v=0.2, s=1, d=0.5) and blocks.
    cfg = configs.double_dam(n=n, res=res, dt=dt)
    with gpu.Scene(cfg.params, cfg.n) as sc:
        sc.upload(cfg.pos, cfg.vel)
        for _ in range(steps):
            sc.step(dt, 1)
        sc.sync()
        cell = sc.download(L.LF_ORDER_CELL)
        by_id = sc.download(L.LF_ORDER_ID)
    mine = str(tmp_path / "ctypes")
    lf.write_geo(mine, steps, cell["pos"], cell["colour"], cell["vel"], float(np.float32(cfg.params.particle_size)))
    a = open(f"{base}.{steps:04d}.geo", "rb").read()
    b = open(f"{mine}.{steps:04d}.geo", "rb").read()
    assert a == b, "fluidtest's last frame differs from the ctypes path's"
    p64, v64, o = oracle_run(oracle_omp, cfg.params, cfg.pos, cfg.vel, dt, steps)
    err = normwise(by_id["pos"], p64)
    col = normwise(cell["colour"][np.argsort(cell["ids"])], oracle_omp.colours(cfg.params, o["density"]))
    print(f"\nfluidtest after {steps} steps: position {err:.2e}, colour {col:.2e} normwise vs the oracle")
    assert err <= TOL and col <= TOL


def test_setup_cpp_and_python_mirror(gpu, oracle_omp, tmp_path):
    n, res, steps, dt = 5000, 16, 5, 0.002
    out = str(tmp_path / "setup.bin")
    r = subprocess.run([os.path.join(BIN, "apicheck"), "setup", str(n), str(res), str(steps), str(dt), out],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    cpp = read_apicheck(out)
    fs = lf.FluidSystem()
    fs.setup(n, res)
    init_pos = fs.scene.download(L.LF_ORDER_ID)["pos"]
    params = L.LfParams.from_buffer_copy(fs.params)
    for _ in range(steps):
        fs.advance(dt, 1)
    pts, colour, vel = fs.exportToData()
    by_id = fs.scene.download(L.LF_ORDER_ID)
    assert cpp["pscale"] == fs.getPscale() and cpp["mass"] == float(np.float32(fs.params.mass))
    for k, v in (("pos", pts), ("colour", colour), ("vel", vel), ("pos_by_id", by_id["pos"])):
        assert np.array_equal(cpp[k], v), f"C++ FluidSystem and the Python mirror differ in {k}"
    # setup(): one block in [0.1, 0.9]^3, plain coefficients, mass from the block
    ref_pts, mass, psize = lf.fluid_block(n, (0.1, 0.1, 0.1), (0.9, 0.9, 0.9), 1000.0)
    assert np.array_equal(init_pos[: ref_pts.shape[0]], ref_pts)
    assert (params.adhesion, params.viscosity, params.surface_tension) == (0.0, 0.0, 0.0)
    assert params.mass == np.float32(mass) and params.particle_size == np.float32(psize)
    fs.clear()
    p64, _, _ = oracle_run(oracle_omp, params, init_pos, np.zeros_like(init_pos), dt, steps)
    err = normwise(by_id["pos"], p64)
    print(f"\nsetup() after {steps} steps: position {err:.2e} normwise vs the oracle")
    assert err <= TOL


def test_setup_from_data_cpp_and_python_mirror(gpu, oracle_omp, tmp_path):
    cfg = configs.scaled_c1(14, 12, 10)
    pos = (cfg.pos * np.float32(0.8) + np.float32(0.1)).astype(np.float32)  # inside the unit cube
    rng = np.random.Generator(np.random.PCG64(9))
    vel = rng.normal(0.0, 0.05, pos.shape).astype(np.float32)
    res, steps, dt = 32, 1, 1e-4  # psize = 1 projects every particle onto a wall each substep (Q27)
    inp = str(tmp_path / "in.bin")
    with open(inp, "wb") as f:
        f.write(np.uint32(pos.shape[0]).tobytes() + pos.tobytes() + vel.tobytes())
    out = str(tmp_path / "fromdata.bin")
    r = subprocess.run([os.path.join(BIN, "apicheck"), "fromdata", inp, str(res), str(steps), str(dt), out],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    cpp = read_apicheck(out)
    assert cpp["pscale"] == 1.0 and cpp["mass"] == 1.0  # Q27: init's defaults are kept
    fs = lf.FluidSystem()
    fs.setupFromData(pos, vel, res)
    params = L.LfParams.from_buffer_copy(fs.params)
    for _ in range(steps):
        fs.advance(dt, 1)
    pts, colour, v = fs.exportToData()
    by_id = fs.scene.download(L.LF_ORDER_ID)
    fs.clear()
    for k, a in (("pos", pts), ("colour", colour), ("vel", v), ("pos_by_id", by_id["pos"])):
        assert np.array_equal(cpp[k], a), f"C++ FluidSystem and the Python mirror differ in {k}"
    assert params.mass == 1.0 and params.particle_size == 1.0
    p64, _, _ = oracle_run(oracle_omp, params, pos, vel, dt, steps)
    err = normwise(by_id["pos"], p64)
    print(f"\nsetupFromData after {steps} steps: position {err:.2e} normwise vs the oracle")
    assert err <= TOL
