"""The pins of tests/test_cpu_switch_pins.py and the flux pins of
tests/test_cpu_oracle_pins.py catch plausible slips in the oracle.  Each
case copies oracle/afar_oracle.cpp, plants one mistake in the copy, builds
it on the CPU into a temporary directory and runs every pin against it: at
least one pin must fail.  (The golden fixture is not used:
it was made by the oracle itself.)  The unmodified oracle passes every pin
(the tests of test_cpu_switch_pins.py)."""
import concurrent.futures
import os
import subprocess

import pytest

from oracle import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "afar_oracle.cpp")

# (name, text in the oracle, replacement)
MUTATIONS = [
    ("bvd_factor_0.6", "BVD_FACTOR = 0.5", "BVD_FACTOR = 0.6"),
    ("smoothness_threshold_2", "(smoothnessValue > 1.0)", "(smoothnessValue > 2.0)"),
    ("valid_lim_factor_0.5", "VALID_LIM_FACTOR = 0.9", "VALID_LIM_FACTOR = 0.5"),
    ("limiter_cap_1", "yy = fmin(2.0, 0.5 * d1", "yy = fmin(1.0, 0.5 * d1"),
    ("transform_selection_inverted", "const bool isShock = finalType < 0;", "const bool isShock = finalType > 0;"),
    ("cfl_weights_swapped",
     "const double fraction = afar_hypot(F[0] - f->CELL[2 * own], F[1] - f->CELL[2 * own + 1]) / deltaR;",
     "const double fraction = 1.0 - afar_hypot(F[0] - f->CELL[2 * own], F[1] - f->CELL[2 * own + 1]) / deltaR;"),
    ("slip_wall_reflects",
     "g[1] = w[1] - (w[1] * nfx + w[2] * nfy) * nfx;\n                g[2] = w[2] - (w[1] * nfx + w[2] * nfy) * nfy;",
     "g[1] = w[1] - 2.0 * (w[1] * nfx + w[2] * nfy) * nfx;\n"
     "                g[2] = w[2] - 2.0 * (w[1] * nfx + w[2] * nfy) * nfy;"),
    # further slips of the same kind
    ("d1_branches_swapped", "d1 = (phif < phic) * (phiMin - phic) + (phif >= phic) * (phiMax - phic);",
     "d1 = (phif < phic) * (phiMax - phic) + (phif >= phic) * (phiMin - phic);"),
    ("small_guard_dropped", "const bool valid = (fabs(d2) >= SMALL);", "const bool valid = (fabs(d2) >= VSMALL);"),
    ("shock_needs_one_extended", "(shockType_extend > 1) * (-2)", "(shockType_extend > 0) * (-2)"),
    ("corner_rule_three_faces", "if (bfacesNum >= 2) {", "if (bfacesNum >= 3) {"),
    ("isothermal_wall_keeps_T", "g[0] = w[0]; g[1] = 0.0; g[2] = 0.0; g[3] = 1.0;",
     "g[0] = w[0]; g[1] = 0.0; g[2] = 0.0; g[3] = w[3];"),
    ("bvd_uses_raw_limiter", "const double l = to_bj(lim[3]);", "const double l = lim[3];"),
    ("cfl_boundary_velocity_dropped", "fraction * fabs(Wn[1] * Sfx + Wn[2] * Sfy)", "fraction * fabs(Wn[1] * Sfx)"),
    # slips in the face fluxes (caught by the flux pins of test_cpu_oracle_pins.py)
    ("roe_r2_sign", "const double r2 = deltaRho - deltaP / (cTilde * cTilde);",
     "const double r2 = deltaRho + deltaP / (cTilde * cTilde);"),
    ("roe_acoustic_entropy_fix_dropped",
     "lambda1 = (lambda1 >= eps) * lambda1 + (lambda1 < eps) * ((lambda1 * lambda1 + eps * eps) / (2.0 * eps + 1e-30));",
     "lambda1 = lambda1;"),
    ("kfvs_internal_energy_dof", "const double K = 2.0 * Cv / R - 2.0;", "const double K = 2.0 * Cv / R;"),
    ("central_energy_pressure_dropped",
     "flux4[3] = 0.5 * rhoFlux * (Cv * TLeft + 0.5 * (UxLeft * UxLeft + UyLeft * UyLeft) + pLeft / rhoLeft) +",
     "flux4[3] = 0.5 * rhoFlux * (Cv * TLeft + 0.5 * (UxLeft * UxLeft + UyLeft * UyLeft)) +"),
    ("roe_sound_speed_gamma", "const double gamma = (Cv + R) / Cv;", "const double gamma = Cv / R;"),
]


def _flux_pins():
    from tests import test_cpu_oracle_pins as F
    normals = [(1.0, 0.0), (0.6, -0.8), (-0.3, 0.95)]

    def consistency():
        for st in F.STATES:
            for n in normals:
                F.test_flux_consistency(st, n)

    return [consistency, F.test_roe_upwind_supersonic, F.test_kfvs_hypersonic_upwind,
            F.test_roe_entropy_fix_transonic_rarefaction, F.test_rk3_amplification_polynomial]


def _build(tmp, name, old, new):
    src = open(SRC).read()
    assert src.count(old) == 1, name
    path = os.path.join(tmp, name + ".cpp")
    with open(path, "w") as f:
        f.write(src.replace(old, new))
    so = os.path.join(tmp, "lib" + name + ".so")
    # -iquote: the oracle includes the library's math header by a path relative to oracle/
    cmd = ["g++", "-O1", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared", "-iquote",
           os.path.join(ROOT, "oracle"), "-o", so, path]
    subprocess.run(cmd, check=True, capture_output=True)
    return so


@pytest.fixture(scope="module")
def mutated_libs(tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("oracle_mutations"))
    with concurrent.futures.ThreadPoolExecutor(max_workers=4) as ex:
        futs = {m[0]: ex.submit(_build, tmp, *m) for m in MUTATIONS}
        return {k: f.result() for k, f in futs.items()}


@pytest.mark.parametrize("name", [m[0] for m in MUTATIONS])
def test_pins_catch_mutation(afar, mutated_libs, monkeypatch, name):
    from tests import test_cpu_switch_pins as P
    monkeypatch.setattr(O, "_lib", O.load(mutated_libs[name]))
    failed = []
    for pin in P.PINS + _flux_pins():
        try:
            pin()
        except AssertionError:
            failed.append(pin.__name__)
    assert failed, "no pin catches the mutation %s" % name
