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
import test_oracle_pins as Q

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
    # momentum, surface tension, density and time step (test_oracle_pins.py)
    ("f_sigma_scale", "2.0 * 6.0 * sqrt(2.0)", "2.0 * 5.0 * sqrt(2.0)",
     Q.test_f_sigma_balances_its_pressure_at_constant_potential, {"direction": "x"}),
    ("f_sigma_band", "if (cl < 0.005 || cl > 0.995 || cg < 0.005 || cg > 0.995 || cs > 0.05) continue;\n"
                     "            psi_f = 0.5 * (AT(psi, i, j) + AT(psi, i - 1, j));",
     "if (cl < 0.05 || cl > 0.995 || cg < 0.005 || cg > 0.995 || cs > 0.05) continue;\n"
     "            psi_f = 0.5 * (AT(psi, i, j) + AT(psi, i - 1, j));",
     Q.test_f_sigma_on_a_tanh_profile, {"direction": "x"}),
    ("f_sigma_y_gradient_index", "gradC = (AT(C, i, j) - AT(C, i, j - 1)) * MT(s->m_idy_c, j - 1);",
     "gradC = (AT(C, i, j + 1) - AT(C, i, j - 1)) * MT(s->m_idy_c, j - 1);",
     Q.test_f_sigma_on_a_tanh_profile, {"direction": "y"}),
    ("gravity_factor", "AT(ru, i, j) += 2.0 * rho_f * Ri * gx;", "AT(ru, i, j) += 1.0 * rho_f * Ri * gx;",
     Q.test_gravity_hydrostatic_balance, {"grav": (3.0, -4.0, 0.0), "rho_kind": "uniform"}),
    ("gravity_v_face_density", "double rho_f = 0.5 * (AT(rho, i, j) + AT(rho, i, j - 1));\n"
                               "            AT(rv, i, j) += 2.0 * rho_f * Ri * gy;",
     "double rho_f = 0.5 * (AT(rho, i, j) + AT(rho, i, j + 1));\n"
     "            AT(rv, i, j) += 2.0 * rho_f * Ri * gy;",
     Q.test_gravity_hydrostatic_balance, {"grav": (0.0, -1.0, 0.0), "rho_kind": "linear_y"}),
    ("convective_uv_sign", "AT(ex, i, j) = -(duudx + duvdy + duwdz);", "AT(ex, i, j) = -(duudx - duvdy + duwdz);",
     Q.test_convective_terms_exact_for_linear_velocity, {}),
    ("convective_v_face_average", "double vuE = 0.5 * (AT(v, i, j) + AT(v, i + 1, j)) * uCE;",
     "double vuE = 0.5 * (AT(v, i, j) + AT(v, i + 1, j - 1)) * uCE;",
     Q.test_convective_terms_exact_for_linear_velocity, {}),
    ("u_rhs_zeta_sign_convective", "double rho_face = 0.5 * (AT(rho, i, j) + AT(rho, i - 1, j));\n"
                                   "            double conv = GAMB[rk] * AT(ex, i, j) + ZETB[rk] * AT(exo, i, j);",
     "double rho_face = 0.5 * (AT(rho, i, j) + AT(rho, i - 1, j));\n"
     "            double conv = GAMB[rk] * AT(ex, i, j) - ZETB[rk] * AT(exo, i, j);",
     Q.test_momentum_rhs_low_storage_rk3, {"stage": 1}),
    ("u_rhs_zeta_sign_viscous", "(GAMB[rk] * AT(ve, i, j) + ZETB[rk] * AT(veo, i, j)) - 2.0 * dpdx;",
     "(GAMB[rk] * AT(ve, i, j) - ZETB[rk] * AT(veo, i, j)) - 2.0 * dpdx;",
     Q.test_momentum_rhs_low_storage_rk3, {"stage": 2}),
    ("v_rhs_pressure_factor", "(GAMB[rk] * AT(ve, i, j) + ZETB[rk] * AT(veo, i, j)) - 2.0 * dpdy;",
     "(GAMB[rk] * AT(ve, i, j) + ZETB[rk] * AT(veo, i, j)) - 1.0 * dpdy;",
     Q.test_momentum_rhs_low_storage_rk3, {"stage": 0}),
    ("rho_from_liquid_fraction", "AT(rho, i, j) = AT(C, i, j) + rho_s * cs + rho_g * AT(G, i, j);",
     "AT(rho, i, j) = AT(C, i, j) + rho_s * cs + rho_g * AT(C, i, j);",
     Q.test_density_viscosity_mixture_rule, {}),
    ("mu_gas_ratio", "AT(mu, i, j) = AT(C, i, j) + mu_s * cs + mu_g * AT(G, i, j);",
     "AT(mu, i, j) = AT(C, i, j) + mu_s * cs + rho_g * AT(G, i, j);",
     Q.test_density_viscosity_mixture_rule, {}),
    ("ch_zeta_on_current_rhs", "AT(rv, i, j) = a_dt * AT(C, i, j) + GAMB[st] * AT(rc, i, j) + ZETB[st] * AT(rn, i, j);",
     "AT(rv, i, j) = a_dt * AT(C, i, j) + GAMB[st] * AT(rc, i, j) + ZETB[st] * AT(rc, i, j);",
     Q.test_ch_step_is_low_storage_rk3, {"stage": 2}),
    ("ch_explicit_biharmonic_sign", "AT(b, i, j) -= bih * AT(psiLG, i, j);", "AT(b, i, j) += bih * AT(psiLG, i, j);",
     Q.test_biharmonic_step_crank_nicolson_factor, {"k": 1, "m": 1}),
    ("explicit_diffusion_peclet", "AT(rhs, i, j) = AT(aux, i, j) / s->prm.Pe_CH;",
     "AT(rhs, i, j) = AT(aux, i, j) * s->prm.Cn;", Q.test_explicit_diffusion_on_a_cosine, {"k": 1}),
    ("jacobi_diagonal_y_dropped", "return d_time + diag_x + diag_y + diag_z + diag_axisym;",
     "return d_time + diag_x + diag_z + diag_axisym;", Q.test_jacobi_diagonal_is_operator_diagonal, {"comp": "u"}),
    ("jacobi_diagonal_v_time_factor", "d_time = rho_face / (alpha_k * s->dt);\n        mu_yp = AT(mu, i, j);",
     "d_time = 0.5 * rho_face / (alpha_k * s->dt);\n        mu_yp = AT(mu, i, j);",
     Q.test_jacobi_diagonal_is_operator_diagonal, {"comp": "v"}),
    ("implicit_viscous_harmonic_mean", "double mu_yp = 2.0 * (0.5 * (AT(mu, i - 1, j + 1) + AT(mu, i, j + 1))) *",
     "double mu_yp = 1.0 * (0.5 * (AT(mu, i - 1, j + 1) + AT(mu, i, j + 1))) *",
     Q.test_implicit_viscous_term_is_the_cg_operator, {"comp": "u"}),
    ("implicit_viscous_v_y_metric", "double d2vdy2 = iRe * (mu_yp * dvdyN - mu_ym * dvdyS) * MT(s->m_idy_c, j - 1);",
     "double d2vdy2 = iRe * (mu_yp * dvdyN - mu_ym * dvdyS) * MT(s->m_idy_c, j - 1) * 0.5;",
     Q.test_implicit_viscous_uniform_mu_is_laplacian_eigenpair, {"k": 3, "m": 2}),
    ("transposed_viscous_dropped_cross_term", "AT(ve, i, j) = ddxdudx + ddydvdx + ddzdwdx;",
     "AT(ve, i, j) = ddxdudx + ddzdwdx;", Q.test_transposed_viscous_uniform_mu_is_grad_div, {}),
    ("cell_center_wrong_neighbour", "AT(ub, i, j) = 0.5 * (AT(u, i, j) + AT(u, i + 1, j));",
     "AT(ub, i, j) = 0.5 * (AT(u, i, j) + AT(u, i - 1, j));", Q.test_cell_center_exact_for_linear_velocity, {}),
    ("capillary_dt_four_pi", "(4.0 * PI_REF));", "(2.0 * PI_REF));", Q.test_dt_capillary_limit_at_rest, {}),
    ("convective_dt_metric", "double convective = u_cfl * MT(s->m_idx_u, i) + v_cfl * MT(s->m_idy_v, j);",
     "double convective = u_cfl * MT(s->m_idx_u, i) + v_cfl * MT(s->m_idx_u, j);", Q.test_dt_convective_limit, {}),
    ("velocity_cg_beta", "beta = lv0 / (rz + EPS);", "beta = lv1 / (rz + EPS);",
     Q.test_velocity_cg_solves_its_system, {"comp": "u"}),
    ("pressure_rhs_substep_factor", "double a_dt = 1.0 / (s->dt * 2.0 * BET[s->stage]);",
     "double a_dt = 1.0 / (s->dt * BET[s->stage]);", Q.test_pressure_rhs_is_scaled_divergence, {}),
    ("divergence_wrong_neighbour", "double dvdy = (AT(v, i, j + 1) - AT(v, i, j)) * MT(s->m_idy_v, j);",
     "double dvdy = (AT(v, i, j + 1) - AT(v, i, j - 1)) * MT(s->m_idy_v, j);",
     Q.test_pressure_rhs_is_scaled_divergence, {}),
    ("jacobi_pcg_stop_test", "if (sqrt(gamma_new) < tol) break;", "if (sqrt(gamma_new) < 1e3 * tol) break;",
     Q.test_pressure_solve_satisfies_poisson, {"solver": 0}),
    ("init_clamp", "C[n] = (F[n] < 0.0) ? 0.0 : ((F[n] > 1.0) ? 1.0 : F[n]);",
     "C[n] = (F[n] < 0.0) ? 0.0 : ((F[n] > 1.0) ? 0.99 : F[n]);", Q.test_init_from_F_bootstrap, {}),
    ("init_clamp_range", "C[n] = (F[n] < 0.0) ? 0.0 : ((F[n] > 1.0) ? 1.0 : F[n]);",
     "C[n] = (F[n] < 0.0) ? 0.0 : ((F[n] > 0.9) ? 0.9 : F[n]);",
     Q.test_vof_diffuse_init_clamps_and_completes_the_phases, {}),
    # ghost exchange: U_VELOCITY ('d') taken for a cell-type component, and
    # one layer too few received
    ("comm_u_type", "(component == 'u' || component == 'd' ||", "(component == 'u' ||",
     Q.test_ghost_exchange_is_the_periodic_extension, {"component": "d", "pnodes": 3}),
    ("comm_layers", "const int precv = ulike ? pnodes + 1 : pnodes", "const int precv = ulike ? pnodes : pnodes",
     Q.test_ghost_exchange_is_the_periodic_extension, {"component": "u", "pnodes": 2}),
    # the momentum CG's exact-step variant: steepest descent instead of CG,
    # and the reference's EPS left in the step length
    ("velocity_cg_exact_beta", "if (s->mom_exact) beta = (rz != 0.0) ? lv0 / rz : 0.0;",
     "if (s->mom_exact) beta = 0.0;", Q.test_velocity_cg_exact_steps, {"comp": "u"}),
    ("velocity_cg_exact_alpha", "if (s->mom_exact) alpha = (dAd != 0.0) ? rz / dAd : 0.0;",
     "if (s->mom_exact) alpha = rz / (dAd + EPS);", Q.test_velocity_cg_exact_steps, {"comp": "v"}),
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
