"""Planted mistakes in the oracle are caught by its pins.

Each case copies oracle/pq_oracle.c, plants one plausible mistake (a wrong
index, sign, coefficient, branch or loop range), builds the copy with the
oracle's own flags and runs the pin that must catch it from
tests/test_oracle_pins.py against that build (PQ_ORACLE_LIB).  The pin
passes on the real oracle (it runs in the same suite) and must fail on the
copy.  CPU only; the copies live in a temporary directory.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "pq_oracle.c")
PINS = "tests/test_oracle_pins.py"
CFLAGS = ["-O1", "-march=x86-64-v2", "-fPIC", "-std=c11", "-ffp-contract=off", "-shared"]

# (name, original text, planted text, occurrence (0-based), pin test)
MUTANTS = [
    ("beta_face_offset", "FI(ui, i + 2, j + 1, l) + FI(ui, i + 1, j + 1, l)",
     "FI(ui, i + 1, j + 1, l) + FI(ui, i + 1, j + 1, l)", 0, "test_beta_cell_centre_velocity"),
    ("beta_time_weight", "(1.0 + tweight_a) * rb -", "(1.0 - tweight_a) * rb -", 0,
     "test_beta_cell_centre_velocity"),
    ("drag_sign", "rv = dt * (-visc_bot * ND(rel, i, j, l) + viscosity * LAP(l));",
     "rv = dt * (visc_bot * ND(rel, i, j, l) + viscosity * LAP(l));", 0, "test_forcing_closed_forms"),
    ("wind_index", "rv = dt * (FW(i - 1, j - 1) + viscosity * LAP(0));",
     "rv = dt * (FW(i, j - 1) + viscosity * LAP(0));", 0, "test_forcing_closed_forms"),
    ("laplacian_neighbour", "ND(rel, i, j - 1, l) + ND(rel, i, j + 1, l))",
     "ND(rel, i, j - 1, l) + ND(rel, i + 1, j + 1, l))", 0, "test_forcing_closed_forms"),
    ("psi_edge_corner", "ND(psi, i, j - 1, l) + ND(psi, i - 1, j, l)",
     "ND(psi, i, j + 1, l) + ND(psi, i - 1, j, l)", 0, "test_cell_boundaries_quadratic_psi"),
    ("cell_time_weight", "FI(ui, i, j, l) = (1.0 + tw) * uim - tw * FI(uimo, i, j, l);",
     "FI(ui, i, j, l) = (1.0 + tw) * uim + tw * FI(uimo, i, j, l);", 0,
     "test_cell_boundaries_quadratic_psi"),
    ("partial_slip_stencil",
     "ND(rel, 2, j, l) = cff * (-3.5 * ND(psi, 2, j, l) + 4.0 * ND(psi, 3, j, l) -",
     "ND(rel, 2, j, l) = cff * (-3.5 * ND(psi, 2, j, l) + 3.0 * ND(psi, 3, j, l) -", 0,
     "test_apply_rel_bc_wall_profiles"),
    ("no_slip_sign", "2.0 * (ND(psi, g, j, l) - ND(psi, g + 1, j, l))",
     "2.0 * (ND(psi, g + 1, j, l) - ND(psi, g, j, l))", 0, "test_apply_rel_bc_wall_profiles"),
    ("rel_coupling_s1_s2", "s1[l] * (ND(psi, i, j, l) - ND(psi, i, j, l - 1)) +",
     "s2[l] * (ND(psi, i, j, l) - ND(psi, i, j, l - 1)) +", 0,
     "test_rel_from_zeta_inverts_pv_definition"),
    ("rel_bottom_layer", "ND(rel, i, j, L - 1) = ND(zeta, i, j, L - 1) +",
     "ND(rel, i, j, L - 1) = ND(zeta, i, j, L - 2) +", 0, "test_rel_from_zeta_inverts_pv_definition"),
    ("zeta_bc_sign", "- s2[0] * (ND(psi, i, j, 0) - ND(psi, i, j, 1)); \\",
     "- s2[0] * (ND(psi, i, j, 1) - ND(psi, i, j, 0)); \\", 0,
     "test_update_zeta_bc_is_pv_definition_on_walls"),
    ("wind_asymmetry", "double a_wind = -1.0 * wind_asym, b_wind = 1.0 / wind_asym;",
     "double a_wind = -1.0 / wind_asym, b_wind = 1.0 * wind_asym;", 0,
     "test_wind_tilted_asymmetric_known_answers"),
    ("dt_growth", "d = fmin(d, fmin(old_dt * max_dt_increase, max_dt));",
     "d = fmin(d, fmin(old_dt, max_dt));", 0, "test_adapt_dt_closed_form"),
    ("tweight_u", "*tweight_u = *dt / (old_dt + *dt);", "*tweight_u = *dt / (2.0 * old_dt);", 0,
     "test_adapt_dt_closed_form"),
    ("cfl_range", "for (int i = 3; i <= g + 1; ++i) mx", "for (int i = 2; i <= g + 1; ++i) mx", 0,
     "test_adapt_dt_closed_form"),
    ("limiter_flux_term", "d_f = -(psi_n - psi_o) * 2.0 - dt * v0 * (fi1 - fi2);",
     "d_f = -(psi_n - psi_o) * 2.0 - dt * v0 * (fi2 - fi1);", 0, "test_limiter_known_answers"),
    ("limiter_upwind_face", "fi2 = FI(zeta_fi, i + 1, j, l);", "fi2 = FI(zeta_fi, i - 1, j, l);", 0,
     "test_limiter_mirror_symmetry"),
    ("limiter_j_cell", "psi_n = ND(zeta, i, j - 1, l);", "psi_n = ND(zeta, i - 1, j, l);", 0,
     "test_limiter_transpose_symmetry"),
    ("limiter_wall_tendency", "d_f = -(psi_n - psi_o);", "d_f = -(psi_n - psi_o) * 2.0;", 0,
     "test_limiter_wall_faces"),
]


def _plant(text, old, new, nth):
    if old is None:  # the control: the oracle as it is
        return text
    pos = -1
    for _ in range(nth + 1):
        pos = text.index(old, pos + 1)
    return text[:pos] + new + text[pos + len(old):]


def _run(case, tmp):
    name, old, new, nth, pin = case
    src = open(SRC).read()
    mutated = _plant(src, old, new, nth)
    assert (mutated != src) == (old is not None)
    c = os.path.join(tmp, f"{name}.c")
    so = os.path.join(tmp, f"lib{name}.so")
    with open(c, "w") as f:
        f.write(mutated)
    subprocess.run(["gcc", *CFLAGS, "-I", os.path.dirname(SRC), "-o", so, c, "-lm"], check=True,
                   capture_output=True)
    env = dict(os.environ, PQ_ORACLE_LIB=so)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        f"{PINS}::{pin}"], cwd=ROOT, env=env, capture_output=True, text=True)
    return name, r.returncode, r.stdout[-400:]


# every pin through the same harness on an unchanged copy: must pass
CONTROLS = [(f"control_{pin}", None, None, 0, pin) for pin in sorted({m[4] for m in MUTANTS})]


@pytest.fixture(scope="module")
def outcomes(tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("mutants"))
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        res = list(ex.map(lambda c: _run(c, tmp), MUTANTS + CONTROLS))
    return {n: (rc, out) for n, rc, out in res}


@pytest.mark.parametrize("name", [c[0] for c in CONTROLS])
def test_pin_passes_through_the_harness(outcomes, name):
    rc, out = outcomes[name]
    assert rc == 0, f"{name} (rc {rc}):\n{out}"


@pytest.mark.parametrize("name", [m[0] for m in MUTANTS])
def test_planted_mistake_is_caught(outcomes, name):
    rc, out = outcomes[name]
    # pytest exits 1 when a test failed (0: the mistake went unnoticed)
    assert rc == 1, f"pin did not catch {name} (rc {rc}):\n{out}"
