"""CPU tests that pin the FP64 oracle (oracle/lf_oracle.cpp) to mathematics.

The reference ships no tests, golden vectors or fixtures, and its CUDA
sources cannot be built here, so every check below comes from a closed form,
an integral identity, a brute-force enumeration or a library routine
(numpy / scipy / std::stable_sort), chosen so that a plausible slip in the
oracle (wrong coefficient, dropped term, self term, transposed sign, wrong
stencil clip, wrong tie order) fails at least one of them.
"""
import json
import math
import os

import numpy as np
import pytest
from scipy import integrate

from libfluid_amd import configs
from libfluid_amd.scene import make_params
from oracle.pyoracle import PIN_F32

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "kat_vectors.json")) as f:
    KAT = json.load(f)

NULL = 0xFFFFFFFF


def unit_params(res=16, **kw):
    p = make_params(**kw)
    p.h = np.float32(1.0) / np.float32(res)
    for a in range(3):
        p.inv_cell[a] = float(res)
        p.grid[a] = res
        p.box[a] = 1.0
    return p


def lone_params(**kw):
    base = dict(rest_density=1000.0, gas_constant=2500.0, gamma=7.0, mass=1e-3, particle_size=0.005,
                xsph_scale=0.5, surface_tension=0.0, viscosity=0.0, adhesion=0.0, vdamp=0.5)
    base.update(kw)
    return unit_params(50, **base)


# ---- smoothing kernels ---------------------------------------------------------

@pytest.mark.parametrize("res", [16, 50])
def test_poly6_is_normalised(oracle, res):
    p = unit_params(res)
    h = float(np.float32(p.h))
    mass, _ = integrate.quad(lambda r: 4 * math.pi * r * r * oracle.poly6(p, r), 0.0, h, epsabs=1e-13)
    # The reference's pi constant 3.141592654 differs from pi by 4e-10.
    assert abs(mass - 1.0) < 1e-8


def test_grad_poly6_is_derivative(oracle):
    p = unit_params(16)
    h = float(np.float32(p.h))
    for r in np.linspace(0.05 * h, 0.95 * h, 7):
        e = 1e-7 * h
        fd = (oracle.poly6(p, r + e) - oracle.poly6(p, r - e)) / (2 * e)
        assert oracle.grad_poly6(p, r) == pytest.approx(fd, rel=1e-6)
    assert oracle.grad_poly6(p, h) == 0.0 and oracle.poly6(p, h) == 0.0


def test_grad_spiky_integrates_to_normalised_spiky(oracle):
    p = unit_params(16)
    h = float(np.float32(p.h))

    def spiky(r):  # W(r) = -int_r^h gradSpiky
        v, _ = integrate.quad(lambda t: oracle.grad_spiky(p, t), r, h, epsabs=1e-14)
        return -v

    mass, _ = integrate.quad(lambda r: 4 * math.pi * r * r * spiky(r), 0.0, h, epsabs=1e-12)
    assert abs(mass - 1.0) < 1e-7
    assert oracle.grad_spiky(p, 0.5 * h) < 0.0  # repulsive gradient direction
    assert oracle.grad_spiky(p, h) == 0.0


def test_cohesion_kernel_shape(oracle):
    """Q8: the coefficient never scales the value (operator precedence)."""
    p = unit_params(16)
    h = float(np.float32(p.h))
    r = 0.75 * h
    assert oracle.cohesion(p, r) == pytest.approx((h - r) ** 3 * r ** 3, rel=1e-12)
    lo = oracle.cohesion(p, 0.5 * h)                      # inner branch at h/2
    hi = oracle.cohesion(p, 0.5 * h * (1 + 1e-12))        # outer branch just above
    assert abs(lo - hi) < 5e-5 * h ** 6                   # 0.0156 vs 1/64 offset
    assert oracle.cohesion(p, h) == 0.0 and oracle.cohesion(p, 2 * h) == 0.0
    assert oracle.cohesion(p, 1e-6 * h) == pytest.approx(-float(np.float32(0.0156)) * h ** 6, rel=1e-6)


@pytest.mark.parametrize("row", KAT["half_space_mass"])
def test_boundary_density_is_missing_kernel_mass(oracle, row):
    """1 - boundaryDensity(d) is the poly6 mass beyond a wall at distance d
    (up to 0.0039062 vs 1/256 in the reference's constant)."""
    p = unit_params(16)
    h = float(np.float32(p.h))
    d = row["t"] * h
    got = 1.0 - oracle.boundary_density(p, d)
    assert got == pytest.approx(row["mass"] * (256 * 0.0039062), abs=1e-6)
    # and the same mass by direct quadrature of the oracle's poly6
    quad, _ = integrate.quad(lambda z: math.pi / 4 * oracle.poly6(p, 0.0) / h ** 6 * (h * h - z * z) ** 4, d, h)
    assert got == pytest.approx(quad * 256 * 0.0039062, abs=1e-6)


# ---- boundary SDF ---------------------------------------------------------------

def test_box_sdf_and_normal(oracle):
    p = unit_params(16)
    p.box[0], p.box[1], p.box[2] = 1.0, 0.5, 0.5
    assert oracle.sdf(p, [0.3, 0.01, 0.2]) == pytest.approx(-0.01)       # nearest: floor
    assert oracle.sdf(p, [0.995, 0.2, 0.2]) == pytest.approx(-0.005)     # nearest: x = 1 wall
    assert oracle.sdf(p, [1.3, 0.6, 0.2]) == pytest.approx(math.hypot(0.3, 0.1))  # outside an edge
    assert oracle.sdf(p, [0.5, 0.25, 0.25]) == pytest.approx(-0.25)
    np.testing.assert_allclose(oracle.sdf_normal(p, [0.3, 0.01, 0.2]), [0, -1, 0], atol=1e-12)
    np.testing.assert_allclose(oracle.sdf_normal(p, [0.3, 0.2, 0.49]), [0, 0, 1], atol=1e-12)
    n = oracle.sdf_normal(p, [1.1, 0.6, 0.7])
    np.testing.assert_allclose(n, np.array([0.1, 0.1, 0.2]) / math.sqrt(0.06), atol=1e-6)


def test_colour_is_continuous(oracle):
    p = unit_params(16)
    rho0 = 1000.0
    for r in (1.0, 2.0):
        a = oracle.colour(p, rho0 * r * (1 - 1e-12))
        b = oracle.colour(p, rho0 * r)
        np.testing.assert_allclose(a, b, atol=1e-9)
    assert list(oracle.colour(p, 0.0)) == [0.0, 0.0, 1.0]


def test_colours_closed_form(oracle):
    """DensityColourOperator (fluidutil.cu:67-80): r = rho/rho0; r < 1 ->
    (0, r, 1-r); r < 2 -> (r-1, 2-r, 0); else (1, 0, 0); the array form
    equals the scalar one."""
    p = unit_params(16)
    rho = np.array([0.0, 250.0, 999.0, 1000.0, 1500.0, 1999.0, 2000.0, 3500.0])
    r = rho / 1000.0
    want = np.where((r < 1)[:, None], np.stack([0 * r, r, 1 - r], 1),
                    np.where((r < 2)[:, None], np.stack([r - 1, 2 - r, 0 * r], 1), np.array([1.0, 0.0, 0.0])))
    got = oracle.colours(p, rho)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-15)
    for k, d in enumerate(rho):
        assert np.array_equal(oracle.colour(p, d), got[k])


# ---- integer stages ----------------------------------------------------------------

def test_hash_matches_numpy_fp32(oracle):
    cfg = configs.dambreak_c1()
    rng = np.random.default_rng(5)
    pos = cfg.pos.copy()
    pos[:50] = rng.uniform(-0.1, 1.1, (50, 3)).astype(np.float32)  # some outside the grid
    pos[50] = [np.nan, 0.1, 0.1]
    keys, counts = oracle.hash(cfg.params, pos)
    inv = np.float32(cfg.params.inv_cell[0])
    f = np.floor(pos * inv)  # float32 multiply, as in the kernel
    g = np.array(cfg.params.grid)
    inside = np.all((f >= 0) & (f < g), axis=1)
    fi = np.where(inside[:, None], f, 0).astype(np.int64)
    exp = np.where(inside, fi[:, 0] + g[0] * (fi[:, 1] + g[1] * fi[:, 2]), NULL).astype(np.uint32)
    assert np.array_equal(keys, exp)
    b = np.where(keys == NULL, cfg.params.ncells, keys)
    assert np.array_equal(counts, np.bincount(b, minlength=cfg.params.ncells + 1))


def test_sort_is_key_then_id(oracle):
    rng = np.random.default_rng(1)
    keys = rng.integers(0, 50, 5000).astype(np.uint32)
    keys[::7] = NULL
    order = oracle.sort(keys)
    ids = np.arange(keys.size)
    assert np.array_equal(order, np.lexsort((ids, keys)))


def test_cell_start_is_exclusive_scan(oracle):
    cfg = configs.dambreak_c1()
    keys, counts = oracle.hash(cfg.params, cfg.pos)
    start = oracle.cell_start(cfg.params, keys)
    exp = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
    assert np.array_equal(start, exp)


def test_step_sort_equals_stable_sort(oracle):
    cfg = configs.scaled_c1(10, 9, 8)
    p = cfg.pos.astype(np.float64)
    v = np.zeros_like(p)
    a = np.zeros_like(p)
    out = oracle.step(cfg.params, 1e-4, p, v, a, PIN_F32, want=("keys", "order", "start"))
    assert np.array_equal(out["order"], oracle.sort(out["keys"]))
    assert np.array_equal(out["start"], oracle.cell_start(cfg.params, out["keys"]))


@pytest.mark.parametrize("seed", [0, 1])
def test_neighbours_cells_equal_brute_force(oracle, seed):
    cfg = configs.scaled_c1(12, 10, 9, seed=seed)
    cnt, lst = oracle.neighbours(cfg.params, cfg.pos)
    bcnt, blst = oracle.neighbours(cfg.params, cfg.pos, brute=True, stencil=True)
    assert np.array_equal(cnt, bcnt) and np.array_equal(lst, blst)
    # With cells no smaller than h the stencil holds every pair within h.
    ucnt, ulst = oracle.neighbours(cfg.params, cfg.pos, brute=True, stencil=False)
    assert np.array_equal(cnt, ucnt) and np.array_equal(lst, ulst)
    assert lst.size > 10 * cfg.n


def test_neighbour_stencil_is_clipped_at_grid_faces(oracle):
    """A pair across the grid's outer face is never in a stencil (positions
    outside the grid hash to NULL_HASH and have no neighbours)."""
    p = lone_params()
    pos = np.array([[0.001, 0.5, 0.5], [-0.001, 0.5, 0.5], [0.011, 0.5, 0.5]], np.float32)
    cnt, lst = oracle.neighbours(p, pos)
    assert list(cnt) == [1, 0, 1] and list(lst) == [2, 0]


# ---- density, pressure, forces ----------------------------------------------------------

def _step(oracle, p, pos, vel=None, acc=None, dt=1e-4, flags=0, want=None):
    pos = np.array(pos, np.float64).reshape(-1, 3)
    vel = np.zeros_like(pos) if vel is None else np.array(vel, np.float64).reshape(-1, 3)
    acc = np.zeros_like(pos) if acc is None else np.array(acc, np.float64).reshape(-1, 3)
    kw = {} if want is None else {"want": want}
    out = oracle.step(p, dt, pos, vel, acc, flags, **kw)
    return out, pos, vel, acc


def test_isolated_particle_density_pressure(oracle):
    """Q3: no neighbours -> density = rest density, pressure and 'sound speed' 0."""
    p = lone_params()
    out, *_ = _step(oracle, p, [[0.5, 0.25, 0.25]])
    assert out["density"][0] == 1000.0 and out["pressure"][0] == 0.0 and out["sound_speed"][0] == 0.0


def test_two_particle_density_has_no_self_term(oracle):
    p = lone_params()
    r = 0.013
    out, *_ = _step(oracle, p, [[0.5, 0.25, 0.25], [0.5 + r, 0.25, 0.25]])
    m = float(np.float32(1e-3))  # parameters are FP32 values
    w = oracle.poly6(p, (0.5 + r) - 0.5)
    assert out["density"] == pytest.approx([m * w, m * w], rel=1e-12)
    # Tait EOS on that density
    ratio = out["density"][0] / 1000.0
    assert out["pressure"][0] == pytest.approx(2500 * 1000 / 7 * (ratio ** 7 - 1), rel=1e-12)
    assert out["sound_speed"][0] == pytest.approx(2500 * ratio ** 6, rel=1e-12)


@pytest.mark.parametrize("row", KAT["lattice_density"])
def test_interior_lattice_density(oracle, row):
    dx, h, m = row["dx"], row["h"], row["m"]
    n = int(round(0.5 / dx))
    p = unit_params(1, mass=m)
    p.h = h
    inv = np.float32(1.0) / np.float32(h)
    g = int(math.floor(1.0 * inv))
    for a in range(3):
        p.inv_cell[a] = inv
        p.grid[a] = g
    ax = (np.arange(n) + 0.5) * dx + 0.25
    pos = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    out, *_ = _step(oracle, p, pos, want=("density",))
    centre = np.argmin(np.linalg.norm(pos - pos.mean(0), axis=1))
    assert out["density"][centre] == pytest.approx(row["rho"], rel=1e-6)


def test_pressure_forces_are_antisymmetric(oracle):
    p = lone_params()
    pos = [[0.5, 0.25, 0.25], [0.51, 0.257, 0.246]]
    out, *_ = _step(oracle, p, pos)
    f = out["f_pressure"]
    np.testing.assert_allclose(f[0], -f[1], rtol=1e-12, atol=1e-9 * np.abs(f).max())
    # repulsive or attractive along the pair axis only
    d = np.array(pos[1]) - np.array(pos[0])
    assert np.linalg.norm(np.cross(f[0], d)) < 1e-9 * np.linalg.norm(f[0]) * np.linalg.norm(d)


def test_viscosity_only_for_approaching_pairs(oracle):
    """Artificial viscosity acts only on approaching pairs (u . (x_i - x_j) < 0)."""
    cfg = configs.scaled_c1(5, 5, 5)
    p = cfg.params
    p.viscosity = 0.2
    nu = float(np.float32(0.2))
    centre = cfg.pos.astype(np.float64).mean(0)
    outward = 100.0 * (cfg.pos - centre)
    # separating everywhere: F_V = nu * F_P exactly
    out, *_ = _step(oracle, p, cfg.pos, vel=outward)
    np.testing.assert_allclose(out["f_viscosity"], nu * out["f_pressure"], rtol=1e-12, atol=1e-9)
    # converging: the Pi term adds to the viscosity force
    out2, *_ = _step(oracle, p, cfg.pos, vel=-outward)
    diff = np.abs(out2["f_viscosity"] - nu * out2["f_pressure"]).max()
    assert diff > 1e-3 * np.abs(out2["f_viscosity"]).max()


def test_adhesion_grows_toward_four_particle_sizes(oracle):
    """Q9: 1/(1-z)^3 with z = distance * 0.25 / psize, zero beyond 4 psize."""
    p = lone_params(adhesion=0.1)
    ps = float(np.float32(0.005))
    mags = []
    for dist in (0.001, 0.01, 0.019, 0.0201):
        out, *_ = _step(oracle, p, [[0.5, dist, 0.25]])
        mags.append(out["f_adhesion"][0])
    z = 0.001 * 0.25 / ps
    alpha = float(np.float32(0.1))
    assert mags[0] == pytest.approx([0, -alpha * 1.0 / (1 - z) ** 3, 0], rel=1e-9)  # isolated: rho = rho0
    assert abs(mags[2][1]) > abs(mags[1][1]) > abs(mags[0][1])
    assert list(mags[3]) == [0.0, 0.0, 0.0]


# ---- integrator ------------------------------------------------------------------------

@pytest.mark.parametrize("row", KAT["free_fall"])
def test_free_fall_leapfrog(oracle, row):
    p = lone_params()
    p.gravity[0], p.gravity[1], p.gravity[2] = 0.0, row["g"], 0.0
    assert float(np.float32(row["g"])) == row["g"]
    pos = np.array([[0.3, row["x0"], 0.25]])
    vel = np.array([[0.0, row["v0"], 0.0]])
    acc = np.array([[0.0, row["a0"], 0.0]])
    for _ in range(row["k"]):
        out, pos, vel, acc = _step(oracle, p, pos, vel, acc, dt=row["dt"], want=())
    assert pos[0, 1] == pytest.approx(row["x"], rel=1e-13, abs=1e-15)
    assert vel[0, 1] == pytest.approx(row["v"], rel=1e-12, abs=1e-15)
    assert acc[0, 1] == row["g"]


@pytest.mark.parametrize("vdamp", [0.0, 0.5, 1.0])
def test_wall_reflection(oracle, vdamp):
    p = lone_params(vdamp=vdamp)
    p.gravity[1] = 0.0
    ps = float(np.float32(0.005))
    out, pos, vel, acc = _step(oracle, p, [[0.3, 0.0051, 0.25]], vel=[[0.2, -1.0, 0.0]], dt=1e-3)
    assert pos[0, 1] == pytest.approx(ps, abs=1e-15)          # projected to psize from the floor
    assert vel[0, 1] == pytest.approx(vdamp * 1.0, rel=1e-12)  # v_n -> -vdamp v_n
    assert vel[0, 0] == pytest.approx(0.2, rel=1e-12)


def test_velocity_and_acceleration_limits(oracle):
    p = lone_params(vel_limit=5.0, acc_limit=150.0)
    p.gravity[0], p.gravity[1], p.gravity[2] = 300.0, 400.0, 0.0  # |g| = 500 > 150
    out, pos, vel, acc = _step(oracle, p, [[0.5, 0.25, 0.25]], vel=[[0.0, 0.0, 10.0]])
    np.testing.assert_allclose(acc[0], [90.0, 120.0, 0.0], rtol=1e-12)
    assert np.linalg.norm(vel[0]) == pytest.approx(5.0, rel=1e-12)


def test_xsph_conserves_momentum(oracle):
    cfg = configs.scaled_c1(6, 6, 6)
    p = cfg.params
    p.gravity[1] = 0.0
    rng = np.random.default_rng(2)
    vel = rng.normal(0, 0.1, cfg.pos.shape)
    out, pos, v, acc = _step(oracle, p, cfg.pos, vel=vel, want=("vel_half",))
    np.testing.assert_allclose(v.sum(0), out["vel_half"].sum(0), atol=1e-12)
    assert not np.allclose(v, out["vel_half"])


def test_pinned_and_free_agree_on_fp32_state(oracle):
    cfg = configs.dambreak_c1()
    a, *_ = _step(oracle, cfg.params, cfg.pos, flags=PIN_F32, want=("keys", "order", "density"))
    b, *_ = _step(oracle, cfg.params, cfg.pos, flags=0, want=("keys", "order", "density"))
    assert np.array_equal(a["keys"], b["keys"]) and np.array_equal(a["order"], b["order"])
    np.testing.assert_allclose(a["density"], b["density"], rtol=1e-12)


def test_openmp_build_is_bit_identical(oracle, oracle_omp):
    cfg = configs.scaled_c1(14, 12, 10)
    outs = []
    for orc in (oracle, oracle_omp):
        p = cfg.pos.astype(np.float64)
        v = np.zeros_like(p)
        a = np.zeros_like(p)
        for _ in range(3):
            o = orc.step(cfg.params, 1e-4, p, v, a, 0, want=("density", "f_pressure"))
        outs.append((p, v, a, o["density"], o["f_pressure"]))
    for x, y in zip(*outs):
        assert np.array_equal(x, y)


def test_cloud_scene_reaches_the_full_flat_table(oracle):
    """The 'cloud' scene of the mask-transparency GPU test holds a lane whose
    flattened-walk table has exactly 16 entries followed by an empty
    pseudo-row, in a wave that walks it (the case that lost a lane's
    neighbours before the table stored only non-empty entries)."""
    from libfluid_amd import configs
    from tests.helpers import flat_table_hazards
    cfg = configs.cloud()
    keys, _ = oracle.hash(cfg.params, cfg.pos)
    order = oracle.sort(keys)
    start = oracle.cell_start(cfg.params, keys)
    assert flat_table_hazards(cfg.params, cfg.pos, order, start) >= 1


# ---- force side against closed-form vectors (tests/golden/make_kat.py) -------------

def _kat_rel(got, want):
    got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
    return float(np.max(np.abs(got - want)) / np.max(np.abs(want)))


@pytest.mark.parametrize("field", ["density", "pressure", "sound_speed", "normal", "f_pressure", "f_viscosity",
                                   "f_tension"])
def test_triple_against_closed_form(oracle, field):
    """Three particles with different densities and pressures, every force
    term on: the oracle's density, EOS, normal and force components equal
    the 50-digit evaluation of the reference's formulas.  Distinct densities
    make the swapped mid-term indices (Q5), the rho_i + rho_i tension
    denominator (Q7) and the 1/rho_j of the normal observable; the pairs
    both approach and separate (viscosity switch)."""
    k = KAT["triple"]
    p = make_params(**k["params"])
    out, *_ = _step(oracle, p, k["pos"], vel=k["vel"], dt=1e-6)
    assert _kat_rel(out[field], k["expect"][field]) < 1e-11, (out[field], k["expect"][field])


def test_triple_pairs_cover_both_viscosity_branches():
    """The triple's pairs include approaching and separating ones."""
    k = KAT["triple"]
    x, v = np.array(k["pos"]), np.array(k["vel"])
    ux = [-np.dot(v[i] - v[j], x[j] - x[i]) for i in range(3) for j in range(3) if i != j]
    assert min(ux) < 0.0 < max(ux)


def test_triple_xsph_against_closed_form(oracle):
    """XSPH (Jacobi) over the moved positions with the unmoved densities:
    zero stiffness, forces and gravity leave v unchanged by the leapfrog and
    move x by v dt; the final velocities equal the closed form, which
    carries the 2 m / (rho_i + rho_j) weight."""
    k = KAT["triple_xsph"]
    p = make_params(**k["params"])
    out, pos, vel, acc = _step(oracle, p, k["pos"], vel=k["vel"], dt=k["dt"], want=("vel_half",))
    np.testing.assert_allclose(out["vel_half"], k["vel"], rtol=0, atol=1e-15)
    np.testing.assert_allclose(pos, k["expect"]["pos"], rtol=1e-14, atol=0)
    assert _kat_rel(vel, k["expect"]["vel"]) < 1e-11
    assert _kat_rel(vel - np.array(k["vel"]), np.array(k["expect"]["vel"]) - np.array(k["vel"])) < 1e-10


def test_compressed_lattice_surface_is_pushed_outward(oracle):
    """Sign check from physics, not from the formula: a small lattice packed
    tighter than its rest spacing has positive pressures, and the pressure
    force on its surface particles points away from the lattice centre (the
    integrated acceleration too); with the lattice stretched, the pressures
    are negative and the surface is pulled inward."""
    h = 0.02
    spacing = 0.4 * h
    ax = 0.5 + spacing * (np.arange(5) - 2.0)
    pos = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    centre = pos.mean(0)
    surf = np.flatnonzero(np.max(np.abs(pos - centre), axis=1) > 1.5 * spacing)
    # heavy particles: every density above rest (compressed); light: below
    for mass, sign in ((4e-3, 1.0), (2e-4, -1.0)):
        p = lone_params(mass=mass)
        out, *_ = _step(oracle, p, pos, want=("density", "pressure", "f_pressure"))
        assert np.all(np.sign(out["pressure"]) == sign), (mass, out["density"].min(), out["density"].max())
        radial = np.einsum("ij,ij->i", out["f_pressure"][surf], pos[surf] - centre)
        assert np.all(np.sign(radial) == sign), f"mass {mass}: surface force not {'out' if sign > 0 else 'in'}ward"
