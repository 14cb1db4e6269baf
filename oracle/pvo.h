/*
 * pvo.h -- CPU oracle for the PARTIES two-phase RK3 substep.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a serial, clean-room C restatement of
 * the reference algorithm (PARTIES/src of m-jalabert/PARTIES-VoF-IBM), used
 * by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker.  The product (libparties_vof.so) never links or calls it.
 *
 * API mirrors include/parties_vof.h one to one (pvo_ prefix, host arrays in
 * the same canonical layout a[NYM+7][NXM+7], (i,j) -> a[j+3][i+3]).
 */
#ifndef PVO_H
#define PVO_H
#include "../include/parties_vof.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pvo_state pvo_state;

pvo_state *pvo_create(const pv_params *prm);
void   pvo_destroy(pvo_state *s);
void   pvo_get_params(const pvo_state *s, pv_params *out);
double *pvo_field(pvo_state *s, int field);        /* canonical layout, owned */
int    pvo_nxt(const pvo_state *s);                /* NXM + 7 */
int    pvo_nyt(const pvo_state *s);                /* NYM + 7 */
int    pvo_threads(void);                            /* host threads of the build (1 without OpenMP) */
int    pvo_init_from_F(pvo_state *s, const double *F);
void   pvo_set_time_state(pvo_state *s, double time, double dt, double dt_old, int stage);
void   pvo_set_momentum_cg_steps(pvo_state *s, int exact);
void   pvo_get_stats(const pvo_state *s, pv_stats *out);

int    pvo_Temporal_int_rk3(pvo_state *s, int nsteps);
int    pvo_Temporal_int_all_the_equations(pvo_state *s);
double pvo_Dtime_cfl(pvo_state *s);

void   pvo_VOF_DIFFUSE_set_boundary_values(pvo_state *s, int field);
void   pvo_VOF_DIFFUSE_compute_laplacian(pvo_state *s, int in, int out);
void   pvo_VOF_DIFFUSE_compute_bulk_S(pvo_state *s);
void   pvo_VOF_DIFFUSE_compute_psi_LG(pvo_state *s);
void   pvo_VOF_DIFFUSE_advect_WENO5(pvo_state *s, int rhs_out);
void   pvo_VOF_DIFFUSE_explicit_diffusion(pvo_state *s, int rhs_out);
int    pvo_VOF_DIFFUSE_solve_implicit_biharmonic(pvo_state *s, int rhs, double *rel_res);
double pvo_VOF_DIFFUSE_bound_liquid_fraction(pvo_state *s);
double pvo_clamp_interior(const pvo_state *s);
void   pvo_VOF_DIFFUSE_update_phase_cache(pvo_state *s);
void   pvo_VOF_DIFFUSE_init(pvo_state *s);
void   pvo_Communication_update_ghost_nodes_flow_variable(pvo_state *s, int field, int component, int pnodes);
void   pvo_VOF_DIFFUSE_step(pvo_state *s);
void   pvo_VOF_DIFFUSE_compute_f_sigma(pvo_state *s);
void   pvo_VOF_update_density_viscosity(pvo_state *s);
void   pvo_VOF_apply_f_sigma_old(pvo_state *s);
void   pvo_VOF_swap_f_sigma(pvo_state *s);
void   pvo_Velocity_add_gravity_2_RHS(pvo_state *s);

void   pvo_Velocity_cell_center(pvo_state *s);
void   pvo_Velocity_u_set_implicit_explicit(pvo_state *s);
void   pvo_Velocity_v_set_implicit_explicit(pvo_state *s);
void   pvo_Velocity_u_set_RHS(pvo_state *s);
void   pvo_Velocity_v_set_RHS(pvo_state *s);
void   pvo_Velocity_update_boundaries(pvo_state *s, int field, int type);
int    pvo_Velocity_solve_cg(pvo_state *s, int component);
void   pvo_matVec(pvo_state *s, int out, int in, int component);

void   pvo_Pressure_set_RHS(pvo_state *s);
void   pvo_Pressure_compute_preconditioner(pvo_state *s);
int    pvo_Pressure_solve_cg(pvo_state *s);
void   pvo_Pressure_operator_variableCoeff(pvo_state *s, int out, int in);
void   pvo_Pressure_apply_BCs(pvo_state *s, int field);
void   pvo_Pressure_project_velocity(pvo_state *s);
double pvo_Pressure_compute_velocity_divergence(pvo_state *s);
void   pvo_Pressure_solve_fft(pvo_state *s, int rhs_field, int out_field);
void   pvo_rk3_tables(double *out);

#ifdef __cplusplus
}
#endif
#endif
