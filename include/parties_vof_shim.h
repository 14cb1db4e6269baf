/*
 * parties_vof_shim.h -- the reference's own routine signatures over the
 * device path of parties_vof.h.
 *
 * A PARTIES build links libparties_vof_shim.so in place of the object files
 * that define these routines (the substep's callers stay as they are).  The
 * shim never looks inside the reference's structs: it keys everything on
 * pointer values the caller registers once at start-up.
 *
 *   pvs_bind(db, handle, grid, params, p, u, v)   which pv_bag stands for db
 *   pvs_bind_field(db, arr, PV_F_*, range)         which device field arr is
 *
 * Host double*** arrays are the reference's (Memory.c:36-96 for the L range,
 * with ghosts; :270-330 for the no-ghost G range): a[k][j][i] with the
 * pointer tables offset so that a[L_Ks][L_Js][L_Is] is the first element.
 * Their values live on the device between pvs_push (host -> device, the
 * k = 0 slab) and pvs_pull (device -> host, every k slab of the array, as
 * the 2-D z self-wrap leaves them); the routines below only pass the
 * arrays' identities to the device.
 *
 * Errors: the reference has no status codes; its fatal path is
 * Display_throw_error (IO/Display.c:910-937: message, MPI_Finalize, exit).
 * The shim calls the handler installed with pvs_set_fatal (default: print
 * the message to stderr and exit(EXIT_FAILURE)); if a handler returns, the
 * routine returns without doing more (a value-returning routine returns -1).
 *
 * Threading: the registry is locked; calls on one bag must not overlap.
 */
#ifndef PARTIES_VOF_SHIM_H
#define PARTIES_VOF_SHIM_H

#include "parties_vof.h"

#ifdef __cplusplus
extern "C" {
#endif

/* the reference's types, by their tags (Include/DataTypes.h:105, 297, 641,
 * 756, 986, 1636, 1663); a translation unit that has included DataTypes.h
 * defines PVS_HAVE_REFERENCE_TYPEDEFS first */
#ifndef PVS_HAVE_REFERENCE_TYPEDEFS
typedef struct cart3d_bag Cart3d_bag;
typedef struct debug_trace Debug_trace;
typedef struct mac_grid MAC_grid;
typedef struct parameters Parameters;
typedef struct velocity Velocity;
typedef struct pressure Pressure;
#endif

/* ---- binding (no counterpart in the reference) -------------------------- */
#define PVS_L_RANGE 0   /* flow variable with ghosts: k -3..4, j -3..NYM+3, i -3..NXM+3 */
#define PVS_G_RANGE 1   /* no-ghost variable: k 0..1, j 0..NYM, i 0..NXM */

/* db stands for the device subdomain h; grid and params identify db in the
 * routines that do not take it (Pressure_apply_BCs, Pressure_solve_x_period),
 * p is db->p and u, v are db->u, db->v (Velocity_solve_cg's component). */
int pvs_bind(Cart3d_bag *db, pv_bag *h, MAC_grid *grid, Parameters *params, Pressure *p, Velocity *u,
             Velocity *v);
/* arr (a reference double*** of the given range) is device field `field` of db */
int pvs_bind_field(Cart3d_bag *db, double ***arr, int field, int range);
/* forget db and every array bound to it (the handle is the caller's) */
int pvs_unbind(Cart3d_bag *db);
pv_bag *pvs_handle(Cart3d_bag *db);
/* host <-> device copies of a bound array (init / resume / output) */
int pvs_push(Cart3d_bag *db, double ***arr);
int pvs_pull(Cart3d_bag *db, double ***arr);
/* end time of Temporal_int_rk3's loop (params->time_max in the reference) */
int pvs_set_time_max(Cart3d_bag *db, double time_max);
/* the pvs_ functions return PV_OK or a PV_E* code; message of the last
 * failure on this thread */
const char *pvs_last_error(void);
typedef void (*pvs_fatal_fn)(const char *message, void *ctx);
void pvs_set_fatal(pvs_fatal_fn fn, void *ctx);

/* ---- Include/Temporal.h:4-5, Include/Dtime.h:3 -------------------------- */
/* the time-stepping loop of Temporal_int.c:270-448 (RK3 substeps and the
 * CFL step; output, statistics and particles are not part of this path):
 * steps while time - time_max < dt, returns 0 */
int Temporal_int_rk3(Cart3d_bag *data_bag, Debug_trace *dtrace);
void Temporal_int_all_the_equations(Cart3d_bag *data_bag, Debug_trace *dtrace);
double Dtime_cfl(Cart3d_bag *data_bag);

/* ---- Include/Communication.h:8 ------------------------------------------ */
void Communication_update_ghost_nodes_flow_variable(double ***data, char component, int pnodes,
                                                    Cart3d_bag *data_bag);

/* ---- Eulerian/Include/VOF_DIFFUSE.h:7-31 -------------------------------- */
void VOF_DIFFUSE_init(Cart3d_bag *db);
void VOF_DIFFUSE_set_boundary_values(double ***f, Cart3d_bag *db);
void VOF_DIFFUSE_update_phase_cache(Cart3d_bag *db);
void VOF_DIFFUSE_compute_laplacian(double ***in, double ***lap, Cart3d_bag *db);
void VOF_DIFFUSE_compute_bulk_S(Cart3d_bag *db);
void VOF_DIFFUSE_compute_psi_LG(Cart3d_bag *db);
void VOF_DIFFUSE_advect_WENO5(Cart3d_bag *db, double ***rhs_out);
void VOF_DIFFUSE_explicit_diffusion(Cart3d_bag *db, double ***rhs_out);
void VOF_DIFFUSE_solve_implicit_biharmonic(Cart3d_bag *db, double ***rhs_explicit);
void VOF_DIFFUSE_step(Cart3d_bag *db);
void VOF_DIFFUSE_compute_f_sigma(Cart3d_bag *db);
void VOF_DIFFUSE_apply_f_sigma_old(Cart3d_bag *db);
void VOF_DIFFUSE_update_density_viscosity(Cart3d_bag *db);

/* ---- Eulerian/Include/VolumeFraction.h:52, 75-76, 84 -------------------- */
void VOF_update_density_viscosity(Cart3d_bag *data_bag);
void VOF_apply_f_sigma_old(Cart3d_bag *data_bag);
void VOF_swap_f_sigma(Cart3d_bag *data_bag);
void Velocity_add_gravity_2_RHS(Cart3d_bag *data_bag);

/* ---- Eulerian/Include/Velocity.h:10-18, 28 ------------------------------ */
void Velocity_cell_center(Cart3d_bag *data_bag);
void Velocity_u_set_implicit_explicit(Cart3d_bag *data_bag);
void Velocity_v_set_implicit_explicit(Cart3d_bag *data_bag);
void Velocity_w_set_implicit_explicit(Cart3d_bag *data_bag);   /* 2-D: w is inactive, nothing to do */
void Velocity_u_set_RHS(Cart3d_bag *data_bag);
void Velocity_v_set_RHS(Cart3d_bag *data_bag);
void Velocity_w_set_RHS(Cart3d_bag *data_bag);                 /* 2-D: nothing to do */
void Velocity_update_boundaries(double ***data, char component, int type, Cart3d_bag *data_bag);
/* Velocity_solve under CG_SOLVE (Include/definitions.h:263-272); returns the iterations */
int Velocity_solve_cg(Velocity *vel, Cart3d_bag *data_bag);

/* ---- Eulerian/Include/Pressure.h:8-13, 32, 43, 51, 65 ------------------- */
void Pressure_set_RHS(Cart3d_bag *data_bag);
void Pressure_compute_preconditioner(Cart3d_bag *data_bag);
int Pressure_solve_cg(Cart3d_bag *data_bag);
/* Aphi = div(grad(phi) / rho); rho must be the bag's density field */
void Pressure_operator_variableCoeff(double ***Aphi, double ***phi, double ***rho, MAC_grid *grid,
                                     Parameters *params, Cart3d_bag *data_bag);
void Pressure_apply_BCs(double ***phi, MAC_grid *grid, Parameters *params);
void Pressure_project_velocity(Cart3d_bag *data_bag);
double Pressure_compute_velocity_divergence(Cart3d_bag *data_bag);
/* Pressure_solve for XPERIODIC without ZPERIODIC (Include/definitions.h:275-283,
 * lsolver/psolve_fft.c:1536-1802): p->rhs -> p->deltap; returns 0 */
int Pressure_solve_x_period(Pressure *p, MAC_grid *grid, Parameters *params);

#ifdef __cplusplus
}
#endif
#endif /* PARTIES_VOF_SHIM_H */
