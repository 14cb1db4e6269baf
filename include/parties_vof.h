/*
 * parties_vof.h -- C ABI of the MI355X (gfx950) implementation of PARTIES'
 * per-Runge-Kutta-substep Eulerian two-phase path.
 *
 * One handle (pv_bag) owns one MPI rank's whole 2-D subdomain on one GPU:
 * every field lives in HBM, halos and boundary conditions are applied on the
 * device, and the host only enqueues kernels.  Entry points keep the names of
 * the reference routines they replace (prefixed pv_) and take the handle in
 * place of the reference's Cart3d_bag*.  Reference anchors are relative to
 * PARTIES/src/ of m-jalabert/PARTIES-VoF-IBM.
 *
 * Boundary set: x periodic, no-slip walls in y, z collapsed (TWOD_CARTESIAN,
 * PERIODIC_NOSLIP_BOX, Include/Boundary.h:35,88-92), GRID_UNIFORM,
 * VOF_DIFFUSE + SURFACE_TENSION + VOF_GRAVITY, CG_SOLVE (FULLY_IMPLICIT),
 * no IBM (C_S == 0).
 *
 * Field layout seen through pv_get_field / pv_set_field ("canonical"):
 *   double a[NYM+7][NXM+7]; cell/face (i,j) -> a[j+3][i+3], i,j in [-3, N+3].
 * It is one k-slab of the reference's [k][j][i] block (Memory.c:36-96) with
 * the same ghost depth (ghost_nodes = 3).  Index conventions follow
 * Grid.c:62-64,187-202: NX = NXM+1, the "G" range is 0..NX-1 and includes
 * the periodic image column i = NX-1 and the wall ghost row j = NY-1.
 *
 * Errors: every entry point returns 0 on success and a negative PV_E* code
 * otherwise; pv_last_error() gives a thread-local message.  The reference
 * has no status codes (it calls Display_throw_error, IO/Display.c:910-937);
 * a C wrapper maps non-zero returns onto that fatal path.
 *
 * Threading: one host thread per handle; calls on one handle are stream
 * ordered and must not overlap.  Handles on different GPUs are independent.
 */
#ifndef PARTIES_VOF_H
#define PARTIES_VOF_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------ */
#define PV_OK             0
#define PV_EINVAL        -1   /* bad argument / size */
#define PV_EHIP          -2   /* HIP runtime error */
#define PV_ENOMEM        -3   /* device allocation failed */
#define PV_ENOCONV       -4   /* pressure outer loop exceeded 10 passes */
#define PV_ENODEV        -5   /* no usable gfx950 device */

/* ---- pressure solver selection ---------------------------------------- */
#define PV_PSOLVE_JACOBI_PCG 0   /* psolve_cg.c:130-293, the reference VOF path */
#define PV_PSOLVE_LEARNED    2   /* FFT-PCG from the learned network's warm start (no
                                    counterpart in the reference; needs a trained network) */
#define PV_PSOLVE_LEARNED_FCG 3  /* learned warm start + learned-preconditioned flexible
                                    CG, falling back to FFT-PCG when the residual stalls */
#define PV_PSOLVE_FFT_PCG    1   /* PCG preconditioned by the FFT-x/Thomas-y solve
                                    of psolve_fft.c:1536-1802, relative-residual stop */
#define PV_PSOLVE_MG_PCG     4   /* the same CG, start, gauge and stop test, preconditioned
                                    by one geometric multigrid V-cycle on the variable face
                                    coefficients (no counterpart in the reference): NXM and
                                    NYM must halve together down to <= 1024 cells; halving
                                    goes on while both are even, to a last level of at most
                                    64 cells */
#define PV_PSOLVE_LEARNED_MG 5   /* MG-PCG from the learned network's warm start, kept
                                    only when its energy is below the cold start's */

/* ---- arithmetic selection ---------------------------------------------- */
#define PV_PREC_FP64  0          /* all FP64 (the plain path) */
#define PV_PREC_MIXED 1          /* FP32 WENO / chemical potential / momentum
                                    stencils, FP64 phi-flux accumulation, Poisson
                                    and PCG inner products */

/* Parameters: names and defaults follow IO/default.inp; derived values as in
 * IO/Input.c:20-127 (Cn <= 0 -> 0.75 h, Pe_CH <= 0 -> 0.9/Cn, slab <= 0 -> h). */
typedef struct pv_params {
    int    NXM, NYM;                 /* [grid]      default 64, 64 */
    double xmin, xmax, ymin, ymax;   /* [geometry]  default 0,1,0,1 */
    double twod_slab_thickness;      /* [twod]      default -1 */
    double Re;                       /* [flow]      default 100 */
    double dp_dx;                    /* [flow]      default 0 */
    double We, rho1, rho2, mu1, mu2; /* [vof]       default 1,1,1,1,1 */
    double Cn, Pe_CH;                /* [vof]       default -1,-1 */
    int    weno_order;               /* [vof]       default 5 */
    int    ch_iter_max;              /* [vof]       default 300 */
    double ch_tol;                   /* [vof]       default 1e-9 */
    double cfl, default_dt, max_dt;  /* [simulation] 0.5, 1e-3, 0.05 */
    int    constant_dt;              /* [simulation] 0 */
    double grav[3];                  /* [particle]  {0,-1,0} */
    double richardson;               /* [conc] richardson[0], default 1 */
    int    P_CG_MAXIT;               /* [lsolve]    300 */
    double P_CG_ETOL;                /* [lsolve]    1e-7 (absolute sqrt(r.z)) */
    int    CG_MAXIT;                 /* [lsolve]    120 */
    double CG_ETOL;                  /* [lsolve]    1e-9 (RMS residual) */
    int    pressure_solver;          /* PV_PSOLVE_*  (not in the reference) */
    double p_rel_tol;                /* FFT-PCG stop ||r||/||b|| (not in the reference) */
    int    p_max_iter;               /* FFT-PCG iteration cap (not in the reference) */
    int    precision;                /* PV_PREC_*   (not in the reference) */
} pv_params;

/* Fill *p with the reference defaults of IO/default.inp (2-D, no IBM). */
void pv_params_default(pv_params *p);

/* ---- fields (canonical layout above) ---------------------------------- */
enum pv_field {
    PV_F_C_L = 0, PV_F_C_G, PV_F_RHO, PV_F_MU,              /* VOF_InterfaceReconstruction.c:39-153 */
    PV_F_PSI_LG, PV_F_LAP_C, PV_F_BULK_S, PV_F_PSI,
    PV_F_CH_AUX1, PV_F_CH_AUX2, PV_F_CH_RHS_N, PV_F_CH_RHS_NM1,
    PV_F_CH_RES, PV_F_CH_DIR, PV_F_FLUX_X, PV_F_FLUX_Y,
    PV_F_FSX_NEW, PV_F_FSY_NEW, PV_F_FSX_OLD, PV_F_FSY_OLD,
    PV_F_U, PV_F_V, PV_F_U_BC, PV_F_V_BC,                   /* Velocity.c */
    PV_F_U_EXPL, PV_F_U_EXPL_OLD, PV_F_U_IMPL, PV_F_U_VISC, PV_F_U_VISC_OLD, PV_F_U_RHS,
    PV_F_V_EXPL, PV_F_V_EXPL_OLD, PV_F_V_IMPL, PV_F_V_VISC, PV_F_V_VISC_OLD, PV_F_V_RHS,
    PV_F_U_R, PV_F_U_D, PV_F_U_AD, PV_F_V_R, PV_F_V_D, PV_F_V_AD,
    PV_F_P, PV_F_PHI, PV_F_P_RHS, PV_F_P_RES, PV_F_P_D, PV_F_P_AD, PV_F_M_INV, /* Pressure.c, psolve_cg.c */
    PV_F_SCRATCH0, PV_F_SCRATCH1, PV_F_SCRATCH2,
    PV_F_F,                                                 /* volume fraction F (VolumeFraction.F),
                                                               the input of VOF_DIFFUSE_init */
    PV_NFIELDS
};

/* Velocity_update_boundaries types (Include/definitions.h VEL_TYPE_*) */
#define PV_VEL_TYPE_NORMAL 0
#define PV_VEL_TYPE_CG     1

/* Per-substep solver accounting (the Display_progress lines of the reference) */
typedef struct pv_stats {
    double time, dt, dt_old;
    int    ntime, which_stage;
    int    n_ch[3], n_u[3], n_v[3], n_p[3], n_pass[3]; /* iterations per stage */
    double ch_res[3];         /* final relative residual of the biharmonic CG  */
    double p_res[3];          /* final pressure residual (sqrt(r.z) or ||r||/||b||) */
    double div_max[3];        /* Pressure_compute_velocity_divergence per stage */
    double clamp_delta[3];    /* diffuse_bound_liquid_fraction return value     */
    int    p_fallback[3];     /* learned preconditioner fell back to FFT-PCG    */
    double clamp_delta_interior[3]; /* the same correction summed over interior cells
                                       only (the reference's sum also counts the periodic
                                       image column i = NXM and the ghost row j = NYM);
                                       no counterpart in the reference */
} pv_stats;

typedef struct pv_bag pv_bag;

const char *pv_last_error(void);

/* Grid_create + Memory/VoF/Pressure/Velocity *_create on the device
 * (Cart3d.c:318-390; Grid.c:30-686, :797-925). */
int pv_create(const pv_params *prm, int device, pv_bag **out);
int pv_destroy(pv_bag *bag);
int pv_get_params(const pv_bag *bag, pv_params *out);   /* derived values filled in */

/* H2D / D2H of one field in canonical layout ((NYM+7)*(NXM+7) doubles).
 * Stand-in for Resume_h5_data / Output_h5_data (IO/Resume.c:341, IO/Output.c:277). */
int pv_set_field(pv_bag *bag, int field, const double *host);
int pv_get_field(pv_bag *bag, int field, double *host);
int pv_fill_field(pv_bag *bag, int field, double value);

/* Initial state from a volume-fraction field F (values on the G range):
 * VOF_DIFFUSE_set_boundary_values(F); VOF_DIFFUSE_init; density/viscosity;
 * psi_LG + f_sigma bootstrap, new -> old (Cart3d.c:1094-1130,
 * VOF_DIFFUSE.c:524-568); u = v = p = 0; t = 0, dt = default_dt and the
 * cell-centre velocity (Temporal_int.c:163-199). */
int pv_init_from_F(pv_bag *bag, const double *F_host);

/* Time and stage state normally held in Parameters (time, dt, dt_old, which_stage). */
int pv_set_time_state(pv_bag *bag, double time, double dt, double dt_old, int which_stage);
/* ntime of Parameters (the step counter a restart file carries, IO/Resume.c:90) */
int pv_set_step_count(pv_bag *bag, int ntime);
int pv_get_stats(pv_bag *bag, pv_stats *out);

/* ---- time integration (Include/Temporal.h:4-5) ------------------------- */
/* nsteps iterations of the RK3 loop body (Temporal_int.c:399-448): CFL on
 * stage 0 when time != 0, three substeps, time += dt * BETA2[rk]. */
int pv_Temporal_int_rk3(pv_bag *bag, int nsteps);
/* One substep at the current which_stage (Temporal_int.c:563-1112, no IBM). */
int pv_Temporal_int_all_the_equations(pv_bag *bag);
/* Dtime.c:24-249; writes the new dt (does not store it). */
int pv_Dtime_cfl(pv_bag *bag, double *dt_out);

/* ---- ghost exchange (Include/Communication.h:8) ------------------------ */
/* Communication.c:128-295 for one rank that is its own x neighbour
 * (XPERIODIC) between y walls: the x exchange of pnodes ghost layers
 * (1..3) on rows -pnodes .. NYM+pnodes.  'u'-type components ('u',
 * U_VELOCITY 'd', U_VELOCITY_PERTURBATION 'e', FLUX_X 'p') receive pnodes+1
 * layers below, the image column i = 0 included (:162-169); every other
 * component wraps the cells 0 .. NXM-1.  In y the rank has no neighbour
 * (MPI_PROC_NULL: ghost rows are left as they are) and the 2-D z self-wrap
 * is the identity on the k = 0 slab. */
int pv_Communication_update_ghost_nodes_flow_variable(pv_bag *bag, int field, int component, int pnodes);

/* ---- Cahn-Hilliard (Eulerian/Include/VOF_DIFFUSE.h:7-31) --------------- */
/* VOF_DIFFUSE.c:524-568 without IBM: C_L = clamp01(F) over the whole
 * L range (ghosts included), then boundary values, phase cache, clamp,
 * boundary values and phase cache of C_L twice, and boundary values of C_G.
 * Reads PV_F_F; writes C_L and C_G (rho and mu are not touched). */
int pv_VOF_DIFFUSE_init(pv_bag *bag);
int pv_VOF_DIFFUSE_set_boundary_values(pv_bag *bag, int field);              /* VOF_DIFFUSE.c:571-683 */
int pv_VOF_DIFFUSE_compute_laplacian(pv_bag *bag, int in, int out);          /* VOF_DIFFUSE.c:838-863 */
int pv_VOF_DIFFUSE_compute_bulk_S(pv_bag *bag);                              /* VOF_DIFFUSE.c:865-885 */
int pv_VOF_DIFFUSE_compute_psi_LG(pv_bag *bag);                              /* VOF_DIFFUSE.c:907-929 */
int pv_VOF_DIFFUSE_advect_WENO5(pv_bag *bag, int rhs_out);                   /* VOF_DIFFUSE.c:931-996 */
int pv_VOF_DIFFUSE_explicit_diffusion(pv_bag *bag, int rhs_out);             /* VOF_DIFFUSE.c:998-1016 */
int pv_VOF_DIFFUSE_solve_implicit_biharmonic(pv_bag *bag, int rhs_explicit,
                                             int *iters, double *rel_res);   /* VOF_DIFFUSE.c:1018-1069 */
int pv_VOF_DIFFUSE_bound_liquid_fraction(pv_bag *bag, double *mass_delta);   /* VOF_DIFFUSE.c:1093-1127 */
int pv_VOF_DIFFUSE_update_phase_cache(pv_bag *bag);                          /* VOF_DIFFUSE.c:108-126 */
int pv_VOF_DIFFUSE_step(pv_bag *bag);                                        /* VOF_DIFFUSE.c:1262-1359 */
int pv_VOF_DIFFUSE_compute_f_sigma(pv_bag *bag);                             /* VOF_DIFFUSE.c:1363-1445 */
int pv_VOF_update_density_viscosity(pv_bag *bag);                            /* VOF_Advection.c:743-748 -> VOF_DIFFUSE.c:1453-1488 */
int pv_VOF_apply_f_sigma_old(pv_bag *bag);                                   /* VOF_SurfaceTension.c:1983-2041 */
int pv_VOF_swap_f_sigma(pv_bag *bag);                                        /* VOF_SurfaceTension.c:2055-2071 */
int pv_Velocity_add_gravity_2_RHS(pv_bag *bag);                              /* VOF_SurfaceTension.c:2075-2154 */

/* ---- momentum (Eulerian/Include/Velocity.h:10-18,28) -------------------- */
int pv_Velocity_cell_center(pv_bag *bag);                                    /* Velocity.c:278-357 */
int pv_Velocity_u_set_implicit_explicit(pv_bag *bag);                        /* Velocity.c:366-762 */
int pv_Velocity_v_set_implicit_explicit(pv_bag *bag);                        /* Velocity.c:771-1144 */
int pv_Velocity_u_set_RHS(pv_bag *bag);                                      /* Velocity.c:1494-1728 */
int pv_Velocity_v_set_RHS(pv_bag *bag);                                      /* Velocity.c:1739-1947 */
int pv_Velocity_update_boundaries(pv_bag *bag, int field, int type);         /* Velocity.c:2427-2805 */
int pv_Velocity_solve_cg(pv_bag *bag, int component, int *iters);            /* lsolver/msolve_cg.c:525-722; component 'u' or 'v' */
int pv_matVec(pv_bag *bag, int out, int in, int component);                  /* lsolver/msolve_cg.c:255-493 */

/* ---- pressure (Eulerian/Include/Pressure.h:8-13,32,43,51,65) ------------ */
int pv_Pressure_set_RHS(pv_bag *bag);                                        /* Pressure.c:100-196 */
int pv_Pressure_compute_preconditioner(pv_bag *bag);                         /* lsolver/psolve_cg.c:63-101 */
int pv_Pressure_solve_cg(pv_bag *bag, int *iters);                           /* lsolver/psolve_cg.c:130-293 (or FFT-PCG) */
int pv_Pressure_operator_variableCoeff(pv_bag *bag, int out, int in);        /* lsolver/psolve_cg.c:302-344 */
int pv_Pressure_apply_BCs(pv_bag *bag, int field);                           /* lsolver/psolve_cg.c:358-529 + cell halo */
int pv_Pressure_project_velocity(pv_bag *bag);                               /* Pressure.c:218-464 */
int pv_Pressure_compute_velocity_divergence(pv_bag *bag, double *div_max);   /* Pressure.c:475-597 */
/* No counterpart in the reference: SCRATCH0 = V(P_RES), one V-cycle of the
 * MG-PCG preconditioner on the face coefficients of the current RHO field
 * (for tests of the preconditioner as a linear operator). */
int pv_mg_vcycle_apply(pv_bag *bag);
/* Constant-coefficient Poisson solve, FFT in x (R2HC, 1/NXM scale, modified
 * wavenumbers) and Thomas in y with the (0,0) mode pinned at j = 0:
 * lsolver/psolve_fft.c:17-207, :467-591, :1536-1802 (z of length 1). */
int pv_Pressure_solve_fft(pv_bag *bag, int rhs_field, int out_field);

/* ---- kernel timers (per-phase tracing; replaces the MPI_Wtime buckets of
 * Timer.c:9-63 / Display_time_results, IO/Display.c:573-655) ----
 * When enabled, launches of the listed kernel classes are bracketed by HIP
 * events on the handle's stream; totals are resolved at the next stream
 * synchronisation.  Off by default (no events recorded). */
enum pv_timer {
    PV_T_CH_APPLY = 0,     /* biharmonic operator + <p,Ap>       */
    PV_T_CH_UPDATE,        /* x, r update + <r,r>                */
    PV_T_MOM_APPLY,        /* momentum operator + <d,Ad>         */
    PV_T_MOM_UPDATE,       /* x, r update + two sums             */
    PV_T_P_APPLY,          /* pressure operator + <d,Ad>         */
    PV_T_P_UPDATE,         /* phi, r update + sums               */
    PV_T_P_DUPDATE,        /* d = z + beta d                     */
    PV_T_FFT_ROWS_FWD,     /* forward real FFT of rows           */
    PV_T_FFT_COLS,         /* DCT-II / scale / DCT-III of columns */
    PV_T_FFT_ROWS_INV,     /* inverse real FFT + <r,z>           */
    PV_T_CH_RHS,           /* fused explicit CH right-hand side  */
    PV_T_MOM_TERMS,        /* fused momentum right-hand side     */
    PV_T_P_TRANSPOSE,      /* spectrum transposes of the FFT-PCG */
    PV_T_MG_PRE0,          /* multigrid pre-smoothing + restriction, level 0 */
    PV_T_MG_COARSE,        /* multigrid levels 1 .. (both ways)  */
    PV_T_MG_POST0,         /* correction + post-smoothing + <r,z>, level 0 */
    PV_NTIMERS
};
/* on = 0 disables; on = k >= 1 times every k-th launch of each timed kernel
 * with a pair of HIP events on the library's stream (k > 1 keeps the event
 * records from perturbing the timed work). */
int pv_timers_enable(pv_bag *bag, int on);
int pv_timers_reset(pv_bag *bag);
/* total milliseconds and launch counts per timer (arrays of PV_NTIMERS) */
int pv_timers_read(pv_bag *bag, double *ms, long long *count);
/* the same over the working launches: of the sampled launches (the first 8192
 * per timer since the last reset), those that ran at least a quarter of the
 * timer's longest one (CG launches queued past convergence exit at once) */
int pv_timers_read_working(pv_bag *bag, double *ms, long long *count);
/* Block until all work queued on the handle's stream has finished. */
int pv_synchronize(pv_bag *bag);

/* ---- learned pressure preconditioner network --------------------------
 * No counterpart in the reference: the 6-layer 3x3 BF16-MFMA network of the
 * learned FFT-FCG pressure solve.  Parameters are PV_CNN_NPARAM floats in the
 * layout documented in csrc/pv_cnn.hip.  Batches are B images of H x W
 * pixels (multiples of 16): in3[B][H][W][3] -> out[B][H][W]. */
#define PV_CNN_NPARAM 38177
int pv_cnn_init(pv_bag *bag, unsigned long long seed);
int pv_cnn_set_params(pv_bag *bag, const float *host);
int pv_cnn_get_params(pv_bag *bag, float *host);
/* forward; with tgt != NULL also the loss mean((out - tgt)^2) over pixels at
 * least `border` from every edge and its gradient (to grad, if not NULL, and
 * kept on the device for pv_cnn_adam_step) */
int pv_cnn_eval(pv_bag *bag, const float *in3, const float *tgt, float *out, float *grad, double *loss,
                int B, int H, int W, int periodic_x, int border);
int pv_cnn_adam_step(pv_bag *bag, float lr, int t);

/* ---- launch graphs, autotuning and batches of cases ---------------------
 * No counterpart in the reference.  pv_set_graphs replays the pressure
 * FFT-PCG from hipGraphs of `chunk` iterations (captured on first use);
 * pv_autotune times eager launches and graph chunks 8, 16, 32 on the current
 * pressure solve and keeps the fastest (the trajectory is unchanged).  A
 * batch holds independent cases of any size (one handle and stream each);
 * pv_batch_rk3 advances all of them, one host thread per case, so their
 * kernels overlap on the GPU; each case computes exactly what it computes
 * alone.  pv_autotune_mg times the multigrid cycles (W on 0 .. 3 of the
 * in-workgroup levels) and sweep counts (1, 2) on cold solves of the
 * pressure right-hand side in place (call it after a step), keeps the
 * fastest and leaves the trajectory as those settings give it;
 * pv_batch_autotune_mg does it for every MG-PCG case of a batch (-1 for the
 * others). */
typedef struct pv_batch pv_batch;
int pv_set_graphs(pv_bag *bag, int on, int chunk);
int pv_autotune(pv_bag *bag, int *chunk);
int pv_batch_create(const pv_params *prms, int ncases, int device, pv_batch **out);
int pv_batch_destroy(pv_batch *bt);
int pv_batch_size(const pv_batch *bt);
pv_bag *pv_batch_case(pv_batch *bt, int k);
int pv_batch_rk3(pv_batch *bt, int nsteps);
/* nthreads host threads share the cases (0 = one thread per case, the
 * default); results are unchanged */
int pv_batch_set_threads(pv_batch *bt, int nthreads);
int pv_batch_autotune(pv_batch *bt, int *chunks);
int pv_autotune_mg(pv_bag *bag, int *wlevels, int *sweeps);
int pv_batch_autotune_mg(pv_batch *bt, int *wlevels, int *sweeps);
/* Tile shapes.  pv_set_momentum_rows: rows of each 256-thread block of the
 * momentum CG kernels, 1, 2 (default) or 4.  pv_set_mg_tile0: the tile of
 * the multigrid's level-0 smoothing kernels, 0 = 64 x 16, 1 = 32 x 16,
 * 2 = 32 x 8, -1 = by width (default: 64 x 16 above 512 columns).  Both
 * change only the order of the partial sums.  pv_autotune_tiles times the
 * candidates on the handle's own momentum system (16 iterations, u and v
 * restored) and cold pressure solves (PHI restored) and keeps the fastest
 * of each; the MG tile is left alone unless the handle runs MG-PCG with
 * tiled levels. */
int pv_set_momentum_rows(pv_bag *bag, int rows);
int pv_set_mg_tile0(pv_bag *bag, int shape);
/* Shape of the FFT preconditioner's y tridiagonal solves (partitioned
 * Thomas): rows per thread's segment, 1..16 (0 = 8), and columns per block,
 * a power of two <= 64 (0 = 2, or more where a block would hold less than a
 * wavefront; halved while a block would exceed 512 threads).  Changes the order of the arithmetic
 * (the separators' reduced system), so results differ at rounding level. */
int pv_set_fft_tri_shape(pv_bag *bag, int rows, int cols);
int pv_autotune_tiles(pv_bag *bag, int *mom_rows, int *mg_tile0);

/* ---- pressure and Cahn-Hilliard solver options ---------------------------
 * No counterpart in the reference. */
/* Pressure solver of the handle (PV_PSOLVE_*), as pv_params.pressure_solver. */
int pv_set_pressure_solver(pv_bag *bag, int solver);
/* Gauss-Seidel sweeps before and after the coarse correction of the
 * multigrid V-cycle (1 or 2, default 2). */
int pv_set_mg_sweeps(pv_bag *bag, int nu);
/* Cycle of the multigrid levels that run in one workgroup (64 x 64 cells and
 * below): the first wlevels of them take two coarse corrections with
 * smoothing between them (pre, correct, post, pre, correct, post), a
 * symmetric W there; 0 (default) = V-cycle throughout. */
int pv_set_mg_cycle(pv_bag *bag, int wlevels);
/* Start of the MG-PCG solves (PV_PSOLVE_MG_PCG, PV_PSOLVE_LEARNED_MG):
 * mode 0 = from phi = 0 (as the reference); mode 1 = also try the pressure
 * correction of the same RK3 stage of the previous step, kept only when the
 * energy |<r, V r>| of its residual is below the cold start's. */
int pv_set_pressure_warm_start(pv_bag *bag, int mode);
/* Arithmetic of the multigrid V-cycle inside the FP64 CG: 0 = FP64 (default
 * for PV_PREC_FP64 handles), 1 = FP32 (default for PV_PREC_MIXED handles).
 * The CG vectors, dots, stop test and solution stay FP64 either way. */
int pv_set_mg_precision(pv_bag *bag, int fp32);
/* Start of the momentum solves (Velocity_solve_cg, msolve_cg.c:525-722):
 * mode 0 = from u^{k-1}, as the reference; mode 1 = from u^{k-1} plus the
 * increment the same RK3 stage's solve made in the previous step, scaled by
 * the ratio of the time steps (the CG's stop test is unchanged). */
int pv_set_momentum_warm_start(pv_bag *bag, int mode);
/* Step lengths of the momentum CG (msolve_cg.c:674 alpha = rz / (dAd + EPS),
 * :705 beta = rz_new / (rz + EPS), EPS = 1e-14): 0 = as the reference
 * (default); 1 = without EPS (alpha = rz / dAd, beta = rz_new / rz, a zero
 * denominator giving a zero step).  Once rz and d.Ad fall to about EPS --
 * with this operator's diagonal of rho / (alpha_k dt), well before small
 * velocities meet CG_ETOL -- the added EPS shortens every step and the CG
 * creeps on until CG_MAXIT; without it the same stop test (RMS residual <
 * CG_ETOL) is met in a few iterations.  The diagonal keeps its + EPS. */
int pv_set_momentum_cg_steps(pv_bag *bag, int exact);
/* Preconditioner of the momentum CG: 0 = Jacobi, as the reference
 * (msolve_cg.c:627-645, default); 1 = the first-order Neumann polynomial of
 * the Jacobi splitting A = D - N, z = D^-1 (r + N D^-1 r), formed from the
 * FP32 operator coefficients (mixed-precision handles only).  Symmetric
 * positive definite for the diagonally dominant momentum operator on a
 * uniform grid; about half the iterations for one more stencil of r each.
 * The CG's stop test (the RMS residual against CG_ETOL) is unchanged. */
int pv_set_momentum_preconditioner(pv_bag *bag, int kind);
/* Cahn-Hilliard implicit solve (the reference runs the CG of
 * VOF_DIFFUSE.c:318-429 from C): 0 = that CG (default); 1 = the CG started
 * from the exact solution of a_dt x + (Cn^2/Pe) L^2 x = b, which the
 * constant-coefficient operator allows on the uniform grid through the
 * FFT-x / DCT-y transforms of the pressure preconditioner (power-of-two
 * sizes); the CG's own stop test accepts it, normally at iteration 0. */
int pv_set_ch_solver(pv_bag *bag, int spectral);

/* ---- learned pressure solve ----------------------------------------------
 * No counterpart in the reference.  Training data come from the plain
 * FFT-PCG solves of the handle's own time steps: while collecting, each
 * pressure solve stores at the given PCG iterations the features of
 * (r_k, M_fft r_k, rho) and the target of the network's correction, cut into
 * crop_size^2 crops (crops_per_snap per snapshot, at most max_samples). */
int pv_learn_collect_begin(pv_bag *bag, const int *snap_iters, int nsnap, int crops_per_snap, int crop_size,
                           int max_samples, unsigned long long seed);
int pv_learn_collect_end(pv_bag *bag, int *nsamples);
/* Adam on random batches of the collected crops; loss of the first and last batch */
int pv_learn_train(pv_bag *bag, int steps, int batch, float lr, unsigned long long seed, double *loss_first,
                   double *loss_last);
/* Energy of the last MG-PCG start relative to the cold start's (learned MG
 * solver), or ||b - A phi0|| / ||b|| of the last learned warm start (FFT
 * forms); >= 1: the candidate was discarded */
int pv_learn_start_residual(pv_bag *bag, double *res);

/* ---- no-allocation variant -----------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): max |div u| of Pressure_compute_velocity_divergence
 * (Pressure.c:475-597, as pv_Pressure_compute_velocity_divergence) on the
 * caller's device arrays u and v of a uniform grid (idx = 1/dx, idy = 1/dy),
 * laid out as this header's canonical fields with row pitch `pitch` (>= NXM+1)
 * and pointers at cell (0, 0).  The result goes to the device double
 * *div_max, in order on the HIP stream `stream` (NULL: the null stream); the
 * call does not synchronise.  It allocates nothing: its block partials live
 * in a fixed 8 KiB device array per GPU, which is why calls that overlap on
 * one GPU race.  NXM and NYM in 1 .. PV_NOALLOC_MAX_N; the value equals the
 * handle entry point's when the handle's metric arrays are idx and idy. */
#define PV_NOALLOC_MAX_N 16384
int pv_Pressure_compute_velocity_divergence_noalloc(const double *u, const double *v, int NXM, int NYM, int pitch,
                                                    double idx, double idy, double *div_max, void *stream);

/* ---- initial volume fraction F on the G range (host only, no device) ----
 * 6 x 6 sub-sampled tanh profile of width 2 sqrt(2) Cn, as in
 * IO/VoF_Init.c:877-953 / :1032-1100 (F = 1 liquid, 0 gas).  The seeded
 * generators use an in-library splitmix64 stream; they have no counterpart in
 * the reference (its RT is single-mode and it has no random bubble field). */
int pv_ic_rising_bubble(const pv_params *prm, double *F);      /* VoF_init_planar_rising_bubble_2d */
int pv_ic_rayleigh_taylor(const pv_params *prm, double *F);    /* VoF_init_rayleigh_taylor_2d */
int pv_ic_multimode_rt(const pv_params *prm, unsigned long long seed, int modes, double amp_max,
                       double y0, double *F);
int pv_ic_bubbles(const pv_params *prm, unsigned long long seed, int nmin, int nmax, double rmin,
                  double rmax, int gas_inside, double *F, int *nplaced);

#ifdef __cplusplus
}
#endif
#endif /* PARTIES_VOF_H */
