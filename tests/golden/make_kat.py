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

Run: python tests/golden/make_kat.py  (rewrites kat_vectors.json next to it)
"""
import json
import math
import os
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
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kat_vectors.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
