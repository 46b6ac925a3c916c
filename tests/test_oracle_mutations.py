"""The oracle tests pin the oracle: plausible slips in its force, normal,
EOS and XSPH code each fail at least one of them.

Each case copies oracle/lf_oracle.cpp to a temporary directory, applies one
edit (a flipped sign, a dropped factor, swapped indices, a "corrected"
reference quirk), builds it with g++ and runs the pinning checks of
tests/test_oracle.py against that build.  A mutant that passes every check
would mean the checks cannot tell it from the oracle."""
import os
import shutil
import subprocess

import pytest

from oracle.pyoracle import Oracle
from tests import test_oracle as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "lf_oracle.cpp")

MUTANTS = {
    "normal sign": ("nrm[i] = Vec3{d.h * s[0], d.h * s[1], d.h * s[2]};",
                    "nrm[i] = Vec3{-d.h * s[0], -d.h * s[1], -d.h * s[2]};"),
    "normal without h": ("nrm[i] = Vec3{d.h * s[0], d.h * s[1], d.h * s[2]};", "nrm[i] = Vec3{s[0], s[1], s[2]};"),
    "normal over rho_i": ("double w = d.m * grad_poly6(d, r) / (rho[j] * r);",
                          "double w = d.m * grad_poly6(d, r) / (rho[i] * r);"),
    "pressure and viscosity term negated": ("double common = d.m * grad_spiky(d, r) * inv_r;",
                                            "double common = -d.m * grad_spiky(d, r) * inv_r;"),
    "pressure sign": ("P[a] += common * rv[a] * mid;", "P[a] -= common * rv[a] * mid;"),
    "mid term indices un-swapped (Q5)": (
        "double mid = prs[j] / (rho[i] * rho[i]) + prs[i] / (rho[j] * rho[j]);",
        "double mid = prs[i] / (rho[i] * rho[i]) + prs[j] / (rho[j] * rho[j]);"),
    "tension over rho_i + rho_j (Q7)": ("double tscale = 2.0 * d.rho0 / (rho[i] + rho[i]);",
                                        "double tscale = 2.0 * d.rho0 / (rho[i] + rho[j]);"),
    "tension without rho0": ("double tscale = 2.0 * d.rho0 / (rho[i] + rho[i]);",
                             "double tscale = 2.0 / (rho[i] + rho[i]);"),
    "no curvature term": ("double curv = -d.sigma * d.m * (ni[a] - nj[a]);", "double curv = 0.0 * (ni[a] - nj[a]);"),
    "cohesion sign": ("double coh = -d.m * d.m * cohesion(d, r) * inv_r;",
                      "double coh = d.m * d.m * cohesion(d, r) * inv_r;"),
    "Pi doubled": ("Pi = (-d.nu * (cs[i] + cs[j]) * mu) / (rho[i] + rho[j]);",
                   "Pi = 2.0 * (-d.nu * (cs[i] + cs[j]) * mu) / (rho[i] + rho[j]);"),
    "Pi sign": ("Pi = (-d.nu * (cs[i] + cs[j]) * mu) / (rho[i] + rho[j]);",
                "Pi = (d.nu * (cs[i] + cs[j]) * mu) / (rho[i] + rho[j]);"),
    "Pi for separating pairs": ("if (ux < 0.0) {", "if (ux > 0.0) {"),
    "mu without h": ("double mu = (d.h * ux) / (rr + 0.01 * d.h2_pred);",
                     "double mu = (ux) / (rr + 0.01 * d.h2_pred);"),
    "viscosity without the pressure term (Q6)": ("V[a] += common * rv[a] * (mid + Pi);",
                                                 "V[a] += common * rv[a] * (Pi);"),
    "sound speed exponent gamma": ("cs[i] = d.K * std::pow(ratio, d.gamma - 1.0);",
                                   "cs[i] = d.K * std::pow(ratio, d.gamma);"),
    "XSPH with m instead of 2m": ("double c = poly6(d, std::sqrt(rr)) * ((2.0 * d.m) / (rho[i] + rho[j]));",
                                  "double c = poly6(d, std::sqrt(rr)) * ((d.m) / (rho[i] + rho[j]));"),
    "XSPH sign": ("s[a] += c * (vh[3 * j + a] - vh[3 * i + a]);", "s[a] += c * (vh[3 * i + a] - vh[3 * j + a]);"),
}


def pinning_checks(orc):
    """The checks of tests/test_oracle.py that a force-side slip should trip."""
    for field in ("density", "pressure", "sound_speed", "normal", "f_pressure", "f_viscosity", "f_tension"):
        yield f"triple {field}", lambda f=field: T.test_triple_against_closed_form(orc, f)
    yield "triple xsph", lambda: T.test_triple_xsph_against_closed_form(orc)
    yield "compressed lattice", lambda: T.test_compressed_lattice_surface_is_pushed_outward(orc)
    yield "pressure antisymmetry", lambda: T.test_pressure_forces_are_antisymmetric(orc)
    yield "viscosity switch", lambda: T.test_viscosity_only_for_approaching_pairs(orc)
    yield "xsph momentum", lambda: T.test_xsph_conserves_momentum(orc)


def build_mutant(tmp, edit):
    old, new = edit
    text = open(SRC).read()
    assert text.count(old) == 1, f"mutation anchor not found once: {old}"
    d = str(tmp)
    shutil.copy(os.path.join(ROOT, "oracle", "lf_oracle.h"), d)
    with open(os.path.join(d, "lf_oracle.cpp"), "w") as f:
        f.write(text.replace(old, new))
    so = os.path.join(d, "liblf_oracle_mut.so")
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wno-unknown-pragmas", "-shared",
                    "-o", so, os.path.join(d, "lf_oracle.cpp")], check=True)
    return Oracle(path=so)


def test_unmutated_oracle_passes_the_pinning_checks(oracle):
    for name, check in pinning_checks(oracle):
        check()


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_mutant_fails_a_pinning_check(tmp_path, name):
    orc = build_mutant(tmp_path, MUTANTS[name])
    caught = []
    for check_name, check in pinning_checks(orc):
        try:
            check()
        except AssertionError:
            caught.append(check_name)
    print(f"\n{name}: caught by {caught}")
    assert caught, f"mutant '{name}' passes every pinning check"
