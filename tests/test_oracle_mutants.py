"""The oracle's pinning checks must catch plausible mistakes.

Each case copies oracle/pvo.c to a temporary directory, plants one mistake by
a textual substitution (a wrong weight, a dropped term, a flipped sign, a
wrong neighbour index, a transposed coefficient, a wrong table entry), builds
it with gcc and runs one of the closed-form checks of test_oracle.py against
it: the check must fail.  Everything here runs on the CPU.
"""
import os
import shutil
import subprocess

import pytest

import oracle_lib
import test_oracle as T

ROOT = oracle_lib.ROOT

# (id, text in oracle/pvo.c, planted replacement, check, check kwargs)
MUTANTS = [
    ("weno_linear_weight", "double a1 = 0.6 /", "double a1 = 0.5 /",
     T.test_weno_fifth_order_on_smooth_data, {}),
    ("weno_dropped_beta_term", "+ 0.25 * pow(v1 - v3, 2.0);", ";",
     T.test_weno_fifth_order_on_smooth_data, {}),
    ("weno_candidate_coefficient", "double q1 = (-v1 + 5.0 * v2 + 2.0 * v3) / 6.0;",
     "double q1 = (-v1 + 5.0 * v2 + 2.0 * v3) / 5.0;", T.test_weno_exact_for_quadratic_cell_averages, {}),
    ("laplacian_wrong_neighbour", "double dfdyN = (AT(f, i, j + 1) - AT(f, i, j))",
     "double dfdyN = (AT(f, i, j + 1) - AT(f, i, j - 1))", T.test_laplacian_neumann_y_eigenpair, {"k": 3}),
    ("psi_sign", "+ 0.5 * c - Cn * Cn * AT(L, i, j);", "+ 0.5 * c + Cn * Cn * AT(L, i, j);",
     T.test_bulk_S_and_psi_pointwise, {}),
    ("biharmonic_sign", "AT(Ax, i, j) = a_dt * AT(x, i, j) + bih * AT(bih_x, i, j);",
     "AT(Ax, i, j) = a_dt * AT(x, i, j) - bih * AT(bih_x, i, j);",
     T.test_biharmonic_cg_matches_spectral_solve, {}),
    ("clamp_delta_sign", "delta += (new_cl - old_cl) * measure;", "delta += (old_cl - new_cl) * measure;",
     T.test_clamp_bounds_idempotence_and_delta, {}),
    ("pressure_transposed_coefficient", "double cxm = (2.0 / (r0 + AT(rho, i - 1, j))) * idx2;",
     "double cxm = (2.0 / (r0 + AT(rho, i + 1, j))) * idx2;",
     T.test_pressure_operator_symmetric_nsd_nullspace, {}),
    ("momentum_wrong_face_viscosity", "double mu_xp = AT(mu, i, j), mu_xm = AT(mu, i - 1, j);",
     "double mu_xp = AT(mu, i, j), mu_xm = AT(mu, i, j);", T.test_momentum_operator_spd, {}),
    ("fft_eigenvalue_factor", "double mw = 2.0 * (1 - cos(", "double mw = 1.0 * (1 - cos(",
     T.test_fft_solve_residual_and_pin, {}),
    ("rk3_table_entry", "{ 4.0 / 15.0, 1.0 / 15.0, 1.0 / 6.0 }", "{ 4.0 / 15.0, 1.0 / 16.0, 1.0 / 6.0 }",
     T.test_rk3_identities, {}),
    ("dt_limiter_else", "    else dt = MIND(dt, s->prm.max_dt);", "    dt = MIND(dt, s->prm.max_dt);",
     T.test_dt_limiter_truth_table, {}),
]


def build_mutant(tmp, old, new):
    src = open(os.path.join(ROOT, "oracle", "pvo.c")).read()
    assert src.count(old) == 1, f"mutation site not unique or missing: {old!r}"
    os.makedirs(os.path.join(tmp, "oracle"))
    os.makedirs(os.path.join(tmp, "include"))
    shutil.copy(os.path.join(ROOT, "oracle", "pvo.h"), os.path.join(tmp, "oracle"))
    shutil.copy(os.path.join(ROOT, "include", "parties_vof.h"), os.path.join(tmp, "include"))
    with open(os.path.join(tmp, "oracle", "pvo.c"), "w") as f:
        f.write(src.replace(old, new))
    so = os.path.join(tmp, "libpvo_mutant.so")
    subprocess.run(["gcc", "-O1", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-std=c99", "-shared", "-o", so,
                    os.path.join(tmp, "oracle", "pvo.c"), "-lm"], check=True)
    return oracle_lib.load_oracle(so)


@pytest.mark.parametrize("name,old,new,check,kw", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_check_catches_planted_mistake(tmp_path, name, old, new, check, kw):
    lib = build_mutant(str(tmp_path), old, new)
    with pytest.raises(AssertionError):
        check(lib, **kw)


def test_checks_pass_on_the_oracle():
    """The same checks pass on the unmodified oracle (guards the harness)."""
    lib = oracle_lib.load_oracle()
    for _, _, _, check, kw in MUTANTS:
        check(lib, **kw)
