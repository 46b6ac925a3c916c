"""Writes kat_vectors.json: known answers derived from mathematics alone.

No part of the solver or the oracle is used here; every number follows from
the closed forms below, so the oracle tests that read these vectors check
the oracle against an independent derivation.

* lattice_density: density of a particle in an infinite simple-cubic lattice
  of spacing dx with smoothing length h, mass m, poly6 kernel
  W(r) = 315/(64 pi h^9) (h^2 - r^2)^3 and no self term:
  rho = m * sum_{o != 0, |o| < h} W(|o|) over the lattice offsets o.
* free_fall: leapfrog of a lone particle (no neighbours, away from walls)
  with a_0 given and a = g afterwards, closed form in free_fall().
* kernel_mass: 4 pi int_0^h W(r) r^2 dr = 1 for poly6 and spiky.
* half_space_mass: fraction of the poly6 mass beyond a plane at distance d,
  pi C6 / 4 int_d^h (h^2 - z^2)^4 dz, closed form in t = d/h.

* triple: three particles alone in the unit cube (h = 1/16, binary-exact
  positions, pairwise within h, different distances so that densities and
  pressures differ), every force term on.  Density, pressure, sound speed,
  the Akinci normal, and the pressure, viscosity and surface-tension forces
  of each particle, evaluated in 50-digit decimal arithmetic straight from
  the reference's formulas (fluidphysics.cu:8-84, 154-210, 215-352;
  fluidkernel.cu:5-38; fluidparams.cu:80-86): the mid term p_j/rho_i^2 +
  p_i/rho_j^2 with its swapped indices, viscosity re-adding the pressure
  term, the rho_i + rho_i tension denominator, the cohesion precedence
  quirk.  Pairs both approach and separate, so the artificial viscosity's
  switch is exercised.
* triple_xsph: the same three particles with zero stiffness, forces and
  gravity, so the leapfrog moves them by v dt exactly and leaves v; the
  XSPH update (fluidphysics.cu:355-409) over the moved positions with the
  densities of the unmoved ones gives the final velocities.

Run: python tests/golden/make_kat.py  (rewrites kat_vectors.json next to it)
"""
import json
import math
import os
import struct
from decimal import Decimal, getcontext
from fractions import Fraction

PI_REF = 3.141592654  # the constant the reference uses for pi


def lattice_density(dx, h, m):
    c6 = 315.0 / (64.0 * PI_REF * h ** 9)
    n = int(math.ceil(h / dx))
    s = 0.0
    for i in range(-n, n + 1):
        for j in range(-n, n + 1):
            for k in range(-n, n + 1):
                if i == j == k == 0:
                    continue
                r2 = (i * i + j * j + k * k) * dx * dx
                if r2 < h * h:
                    s += (h * h - r2) ** 3
    return m * c6 * s


def free_fall(x0, v0, a0, g, dt, k):
    """Leapfrog x_{n+1} = x_n + v_n dt + a_n dt^2/2, v_{n+1} = v_n + (a_n + a_{n+1}) dt/2
    with a_n = g for n >= 1, from (x0, v0, a0).  For k >= 1:
      v_1 = v_0 + (a_0 + g) dt/2,   v_k = v_1 + (k-1) g dt,
      x_k = x_0 + v_0 dt + a_0 dt^2/2 + (k-1) v_1 dt + g dt^2 (k-1)^2 / 2."""
    v1 = v0 + 0.5 * (a0 + g) * dt
    v = v1 + (k - 1) * g * dt
    x = x0 + v0 * dt + 0.5 * a0 * dt * dt + (k - 1) * v1 * dt + 0.5 * g * dt * dt * (k - 1) ** 2
    return x, v


def half_space_mass(t):
    """int_t^1 (1 - u^2)^4 du / int_0^1 (1 - u^2)^4 du * 1/2  (exact polynomial)."""
    # antiderivative of (1-u^2)^4 = 1 - 4u^2 + 6u^4 - 4u^6 + u^8
    def F(u):
        return u - Fraction(4, 3) * u ** 3 + Fraction(6, 5) * u ** 5 - Fraction(4, 7) * u ** 7 + Fraction(1, 9) * u ** 9
    tt = Fraction(t).limit_denominator(10 ** 12)
    full = F(Fraction(1))  # = 128/315
    return float((F(Fraction(1)) - F(tt)) / full / 2)


def f32(x):
    """x rounded to FP32 (the value a float parameter holds), as a float."""
    return struct.unpack("f", struct.pack("f", x))[0]


def D(x):
    return Decimal(repr(float(x)))


TRIPLE_PARAMS = {  # FP32-exact values; unit cube, res 16
    "rest_density": 1000.0, "gas_constant": 2500.0, "gamma": 7.0, "mass": 0.25, "particle_size": 0.0078125,
    "xsph_scale": 0.5, "surface_tension": 1.0, "viscosity": 0.25, "adhesion": 0.0, "vel_limit": 5.0,
    "acc_limit": 150.0, "vdamp": 0.5, "gravity": [0.0, 0.0, 0.0], "h": 0.0625, "inv_cell": [16.0, 16.0, 16.0],
    "grid": [16, 16, 16], "box": [1.0, 1.0, 1.0],
}
# Offsets in 1/128 from the centre of the cube: |AB|^2 = 10, |AC|^2 = 14,
# |BC|^2 = 30 (in 1/128^2), all within h^2 = 64.
TRIPLE_POS = [[0.5, 0.5, 0.5], [0.5 + 3 / 128, 0.5 + 1 / 128, 0.5], [0.5 - 2 / 128, 0.5 + 3 / 128, 0.5 + 1 / 128]]
TRIPLE_VEL = [[0.3, 0.1, 0.0], [-0.2, 0.05, 0.1], [-0.1, 0.3, 0.05]]  # A-B, B-C approach; A-C separate


def triple(params, pos, vel):
    getcontext().prec = 50
    pi = Decimal("3.141592654")
    h = D(params["h"])
    hh = h * h
    m = D(params["mass"])
    rho0, K, gamma = D(params["rest_density"]), D(params["gas_constant"]), D(params["gamma"])
    nu, sigma = D(params["viscosity"]), D(params["surface_tension"])
    c6 = Decimal(315) / (64 * pi * h ** 9)          # poly6
    cg6 = Decimal(-945) / (32 * pi * h ** 9)        # gradPoly6
    cgs = Decimal(-45) / (pi * h ** 6)              # gradSpiky
    coh_off = D(f32(0.0156)) * h ** 6
    x = [[D(c) for c in p] for p in pos]
    v = [[D(c) for c in p] for p in vel]
    n = len(x)

    def sub(a, b):
        return [a[k] - b[k] for k in range(3)]

    def dot(a, b):
        return sum(a[k] * b[k] for k in range(3))

    def W(r2):
        return c6 * (hh - r2) ** 3

    def cohesion(r):
        base = (h - r) ** 3 * r ** 3
        return base if r > h / 2 else 2 * base - coh_off

    rho = []
    for i in range(n):
        s = sum(m * W(dot(sub(x[j], x[i]), sub(x[j], x[i]))) for j in range(n) if j != i)
        rho.append(s)  # no walls within h: no boundary term; s > 0
    ratio = [r / rho0 for r in rho]
    prs = [(K * rho0 / gamma) * (q ** 7 - 1) for q in ratio]
    cs = [K * q ** 6 for q in ratio]
    nrm = []
    for i in range(n):
        acc = [Decimal(0)] * 3
        for j in range(n):
            if j == i:
                continue
            d = sub(x[i], x[j])  # x_i - x_j
            r = dot(d, d).sqrt()
            g = cg6 * r * (hh - r * r) ** 2
            acc = [acc[k] + m * g * d[k] / (rho[j] * r) for k in range(3)]
        nrm.append([h * a for a in acc])
    fp, fv, ft = [], [], []
    for i in range(n):
        P, V, T = [Decimal(0)] * 3, [Decimal(0)] * 3, [Decimal(0)] * 3
        for j in range(n):
            if j == i:
                continue
            rv = sub(x[j], x[i])  # x_j - x_i
            r2 = dot(rv, rv)
            r = r2.sqrt()
            u = sub(v[i], v[j])
            ux = -dot(u, rv)
            Pi = Decimal(0)
            if ux < 0:
                mu = h * ux / (r2 + Decimal("0.01") * hh)
                Pi = -nu * (cs[i] + cs[j]) * mu / (rho[i] + rho[j])
            mid = prs[j] / rho[i] ** 2 + prs[i] / rho[j] ** 2
            grad = cgs * (h - r) ** 2
            common = [m * grad * c / r for c in rv]
            coh = [-m * m * cohesion(r) * c / r for c in rv]
            curv = [-sigma * m * (nrm[i][k] - nrm[j][k]) for k in range(3)]
            P = [P[k] + common[k] * mid for k in range(3)]
            V = [V[k] + common[k] * (mid + Pi) for k in range(3)]
            T = [T[k] + 2 * rho0 * (coh[k] + curv[k]) / (rho[i] + rho[i]) for k in range(3)]
        fp.append(P)
        fv.append([nu * c for c in V])
        ft.append([sigma * c for c in T])

    def fl(a):
        return [[float(c) for c in row] for row in a] if isinstance(a[0], list) else [float(c) for c in a]

    return {"density": fl(rho), "pressure": fl(prs), "sound_speed": fl(cs), "normal": fl(nrm),
            "f_pressure": fl(fp), "f_viscosity": fl(fv), "f_tension": fl(ft)}


def triple_xsph(params, pos, vel, dt):
    """Zero stiffness, forces and gravity: x' = x + v dt, v unchanged by the
    leapfrog; XSPH over x' with the densities of x."""
    getcontext().prec = 50
    pi = Decimal("3.141592654")
    h = D(params["h"])
    hh = h * h
    m = D(params["mass"])
    eps_x = D(params["xsph_scale"])
    c6 = Decimal(315) / (64 * pi * h ** 9)
    x = [[D(c) for c in p] for p in pos]
    v = [[D(c) for c in p] for p in vel]
    n = len(x)

    def r2(a, b):
        return sum((a[k] - b[k]) ** 2 for k in range(3))

    rho = [sum(m * c6 * (hh - r2(x[i], x[j])) ** 3 for j in range(n) if j != i) for i in range(n)]
    xm = [[x[i][k] + v[i][k] * D(dt) for k in range(3)] for i in range(n)]
    out = []
    for i in range(n):
        s = [Decimal(0)] * 3
        for j in range(n):
            if j == i:
                continue
            d2 = r2(xm[i], xm[j])
            assert d2 <= hh
            c = c6 * (hh - d2) ** 3 * (2 * m) / (rho[i] + rho[j])
            s = [s[k] + c * (v[j][k] - v[i][k]) for k in range(3)]
        out.append([float(v[i][k] + eps_x * s[k]) for k in range(3)])
    return {"pos": [[float(c) for c in p] for p in xm], "vel": out}


def main():
    out = {
        "lattice_density": [
            {"dx": dx, "h": h, "m": m, "rho": lattice_density(dx, h, m)}
            for dx, h, m in [(0.01, 0.02, 1e-3), (0.01, 0.025, 1e-3), (0.05, 0.1, 0.125)]
        ],
        "free_fall": [],
        "half_space_mass": [{"t": t, "mass": half_space_mass(t)} for t in (0.0, 0.1, 0.25, 0.5, 0.75, 0.9, 1.0)],
    }
    g = -9.800000190734863  # -9.8 rounded to FP32, the value the solver's float parameter holds
    for k in (1, 2, 5, 20):
        x, v = free_fall(0.5, 0.3, 2.0, g, 1e-3, k)
        out["free_fall"].append({"x0": 0.5, "v0": 0.3, "a0": 2.0, "g": g, "dt": 1e-3, "k": k, "x": x, "v": v})
    out["triple"] = {"params": TRIPLE_PARAMS, "pos": TRIPLE_POS, "vel": TRIPLE_VEL,
                     "expect": triple(TRIPLE_PARAMS, TRIPLE_POS, TRIPLE_VEL)}
    xp = dict(TRIPLE_PARAMS, gas_constant=0.0, viscosity=0.0, surface_tension=0.0)
    out["triple_xsph"] = {"params": xp, "pos": TRIPLE_POS, "vel": TRIPLE_VEL, "dt": 1e-3,
                          "expect": triple_xsph(xp, TRIPLE_POS, TRIPLE_VEL, 1e-3)}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kat_vectors.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
