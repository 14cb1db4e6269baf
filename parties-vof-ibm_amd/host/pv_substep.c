/*
 * pv_substep.c -- sequencing of one RK3 substep on the device and the
 * per-routine entry points of include/parties_vof.h.
 *
 * Reference: PARTIES/src/Temporal_int.c:399-448 (RK loop) and :563-1112
 * (substep, no IBM); the routine each entry replaces is cited at its head.
 */
#include "pv_host.h"
#include <math.h>
#include <stdio.h>
#include <string.h>

#define HIPCHK(call)                                                                 \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess)                                                        \
            return pv_set_error(PV_EHIP, "%s failed: %s (%s:%d)", #call,             \
                                hipGetErrorString(e_), __FILE__, __LINE__);          \
    } while (0)
#define LAUNCHCHK() HIPCHK(hipGetLastError())
#define RC(x) do { int rc_ = (x); if (rc_) return rc_; } while (0)

#define G(b) (&(b)->g)
#define S(b) ((b)->stream)
#define FD(b, fid) ((b)->f[fid])

static int comp_of_field(int f) { return (f == PV_F_U || f == PV_F_U_D) ? 'u' : 'v'; }

static int begin(pv_bag *b)
{
    if (!b) return pv_set_error(PV_EINVAL, "null handle");
    HIPCHK(hipSetDevice(b->device));
    return PV_OK;
}

/* ------------------------------------------------------------------ */
/* Cahn-Hilliard                                                       */
/* ------------------------------------------------------------------ */

/* diffuse_solve_biharmonic_pcg loop (VOF_DIFFUSE.c:383-423) after its start
 * has been enqueued; x = C_L, r = CH_RES, p = CH_DIR, Ap = LAP_C. */
static int ch_cg_loop(pv_bag *b, int stage)
{
    double *x = FD(b, PV_F_C_L), *r = FD(b, PV_F_CH_RES), *Ap = FD(b, PV_F_LAP_C);
    int chunk = b->ch_chunk, n = 0, maxn = b->ph.ch_iter_max + 1;
    for (;;) {
        int k;
        for (k = 0; k < chunk && n <= maxn; k++, n++) {
            pv_tbegin(b, PV_T_CH_APPLY);
            pvk_ccg_iter(G(b), &b->ph, stage, b->d_sc, &b->ccg, 0, n, x, r, Ap, S(b));
            pv_tend(b, PV_T_CH_APPLY);
            pv_tbegin(b, PV_T_CH_UPDATE);
            pvk_ccg_iter(G(b), &b->ph, stage, b->d_sc, &b->ccg, 1, n, x, r, Ap, S(b));
            pv_tend(b, PV_T_CH_UPDATE);
        }
        LAUNCHCHK();
        RC(pv_read_scalars(b));
        if (b->h_sc->ch_done) break;
        if (n > maxn) return pv_set_error(PV_EHIP, "biharmonic CG did not stop (device flag)");
        chunk = 2;
    }
    b->ch_chunk = (b->h_sc->ch_iter > 1) ? b->h_sc->ch_iter + 1 : 2;
    b->st.n_ch[stage] = b->h_sc->ch_iter;
    b->st.ch_res[stage] = b->h_sc->ch_rel;
    pvk_bc_ch(G(b), x, S(b));
    LAUNCHCHK();
    return PV_OK;
}

static void ch_spectral_start(pv_bag *b, int st);

/* VOF_DIFFUSE_step (VOF_DIFFUSE.c:1262-1359) + VOF_update_density_viscosity.
 * With ch_spectral the CG starts from C + (a_dt + bih L^2)^-1 (b - A C),
 * the exact solution on the uniform grid (the CH operator is diagonal in the
 * FFT-x / DCT-y basis of the pressure preconditioner); the CG's stop test
 * then accepts it at iteration 0, or iterates on if it is not exact. */
static int ch_step(pv_bag *b)
{
    int st = b->stage, spec = b->ch_spectral && b->fft.ok && b->fft.dct_ok;
    /* VOF_DIFFUSE_advect_WENO5 starts from the boundary values of C_L
     * (VOF_DIFFUSE.c:945): they are current after every clamp of this
     * library, but not after a caller or a restart has written C_L */
    pvk_bc_ch(G(b), FD(b, PV_F_C_L), S(b));
    pv_tbegin(b, PV_T_CH_RHS);
    pvk_ccg_init(G(b), &b->ph, st, b->d_sc, &b->ccg, FD(b, PV_F_C_L), FD(b, PV_F_U), FD(b, PV_F_V),
                 FD(b, PV_F_CH_RHS_N), NULL, FD(b, PV_F_CH_RES), FD(b, PV_F_CH_DIR),
                 spec ? FD(b, PV_FI_TMP) : NULL, S(b));
    pv_tend(b, PV_T_CH_RHS);
    if (spec) ch_spectral_start(b, st);
    LAUNCHCHK();
    RC(ch_cg_loop(b, st));
    pvk_clamp_props(G(b), &b->ph, FD(b, PV_F_C_L), FD(b, PV_F_C_G), FD(b, PV_F_RHO), FD(b, PV_F_MU), &b->rd,
                    b->d_sc, S(b));
    LAUNCHCHK();
    return PV_OK;
}

int pv_VOF_DIFFUSE_set_boundary_values(pv_bag *b, int field)
{
    RC(begin(b));
    if (field < 0 || field >= PV_NFIELDS) return pv_set_error(PV_EINVAL, "bad field");
    pvk_bc_ch(G(b), FD(b, field), S(b));
    LAUNCHCHK();
    return PV_OK;
}

/* Communication.c:128-295, one rank: x exchange only (no y neighbour, z is
 * the k = 0 slab) */
int pv_Communication_update_ghost_nodes_flow_variable(pv_bag *b, int field, int component, int pnodes)
{
    int ulike;
    RC(begin(b));
    if (field < 0 || field >= PV_NFIELDS) return pv_set_error(PV_EINVAL, "bad field id %d", field);
    if (pnodes < 1 || pnodes > 3) return pv_set_error(PV_EINVAL, "pnodes must be 1..3 (got %d)", pnodes);
    ulike = component == 'u' || component == 'd' || component == 'e' || component == 'p';
    pvk_xwrap(G(b), FD(b, field), ulike, pnodes, S(b));
    LAUNCHCHK();
    return PV_OK;
}

/* VOF_DIFFUSE.c:524-568 (no IBM).  The phase-cache kernels also refresh rho
 * and mu from C_G, as they do everywhere in this library; F ends as the copy
 * of C_L that the reference's phase cache keeps (VOF_DIFFUSE.c:108-126).
 * The substep does not maintain that mirror: nothing on the path reads F
 * after start-up, and the output writers write C_L for it. */
int pv_VOF_DIFFUSE_init(pv_bag *b)
{
    RC(begin(b));
    pvk_clamp_slab(G(b), FD(b, PV_F_F), FD(b, PV_F_C_L), S(b));
    pvk_bc_ch(G(b), FD(b, PV_F_C_L), S(b));
    pvk_phase_props(G(b), &b->ph, FD(b, PV_F_C_L), FD(b, PV_F_C_G), FD(b, PV_F_RHO), FD(b, PV_F_MU), S(b));
    pvk_clamp_props(G(b), &b->ph, FD(b, PV_F_C_L), FD(b, PV_F_C_G), FD(b, PV_F_RHO), FD(b, PV_F_MU), &b->rd,
                    b->d_sc, S(b));
    pvk_bc_ch(G(b), FD(b, PV_F_C_L), S(b));
    pvk_phase_props(G(b), &b->ph, FD(b, PV_F_C_L), FD(b, PV_F_C_G), FD(b, PV_F_RHO), FD(b, PV_F_MU), S(b));
    pvk_bc_ch(G(b), FD(b, PV_F_C_L), S(b));
    pvk_phase_props(G(b), &b->ph, FD(b, PV_F_C_L), FD(b, PV_F_C_G), FD(b, PV_F_RHO), FD(b, PV_F_MU), S(b));
    pvk_bc_ch(G(b), FD(b, PV_F_C_G), S(b));
    pvk_copy(G(b), FD(b, PV_F_C_L), FD(b, PV_F_F), S(b));     /* the phase cache's F mirror (:108-126) */
    LAUNCHCHK();
    return PV_OK;
}

int pv_VOF_DIFFUSE_compute_laplacian(pv_bag *b, int in, int out)
{
    RC(begin(b));
    pvk_bc_ch(G(b), FD(b, in), S(b));
    pvk_laplacian(G(b), FD(b, in), FD(b, out), S(b));
    pvk_bc_ch(G(b), FD(b, out), S(b));
    LAUNCHCHK();
    return PV_OK;
}

int pv_VOF_DIFFUSE_compute_bulk_S(pv_bag *b)
{
    RC(begin(b));
    pvk_bulk_S(G(b), &b->ph, FD(b, PV_F_C_L), FD(b, PV_F_BULK_S), S(b));
    pvk_bc_ch(G(b), FD(b, PV_F_BULK_S), S(b));
    LAUNCHCHK();
    return PV_OK;
}

int pv_VOF_DIFFUSE_compute_psi_LG(pv_bag *b)
{
    RC(begin(b));
    pvk_bc_ch(G(b), FD(b, PV_F_C_L), S(b));
    pvk_laplacian(G(b), FD(b, PV_F_C_L), FD(b, PV_F_LAP_C), S(b));
    pvk_bc_ch(G(b), FD(b, PV_F_LAP_C), S(b));
    pvk_psi_only(G(b), &b->ph, FD(b, PV_F_C_L), FD(b, PV_F_PSI_LG), S(b));
    pvk_bc_ch(G(b), FD(b, PV_F_PSI_LG), S(b));
    LAUNCHCHK();
    return PV_OK;
}

int pv_VOF_DIFFUSE_advect_WENO5(pv_bag *b, int rhs_out)
{
    RC(begin(b));
    pvk_bc_ch(G(b), FD(b, PV_F_C_L), S(b));
    pvk_weno_advect(G(b), &b->ph, FD(b, PV_F_C_L), FD(b, PV_F_U), FD(b, PV_F_V), FD(b, rhs_out), S(b));
    pvk_bc_ch(G(b), FD(b, rhs_out), S(b));
    LAUNCHCHK();
    return PV_OK;
}

int pv_VOF_DIFFUSE_explicit_diffusion(pv_bag *b, int rhs_out)
{
    RC(begin(b));
    pvk_bc_ch(G(b), FD(b, PV_F_C_L), S(b));
    pvk_explicit_diffusion(G(b), &b->ph, FD(b, PV_F_C_L), FD(b, rhs_out), S(b));
    pvk_bc_ch(G(b), FD(b, rhs_out), S(b));
    LAUNCHCHK();
    return PV_OK;
}

int pv_VOF_DIFFUSE_solve_implicit_biharmonic(pv_bag *b, int rhs, int *iters, double *rel_res)
{
    RC(begin(b));
    pvk_bc_ch(G(b), FD(b, PV_F_C_L), S(b));
    pvk_ch_sub_bih(G(b), &b->ph, FD(b, PV_F_C_L), FD(b, rhs), S(b));
    pvk_ccg_init(G(b), &b->ph, b->stage, b->d_sc, &b->ccg, FD(b, PV_F_C_L), NULL, NULL, NULL, FD(b, rhs),
                 FD(b, PV_F_CH_RES), FD(b, PV_F_CH_DIR), NULL, S(b));
    LAUNCHCHK();
    RC(ch_cg_loop(b, b->stage));
    if (iters) *iters = b->h_sc->ch_iter;
    if (rel_res) *rel_res = b->h_sc->ch_rel;
    return PV_OK;
}

int pv_VOF_DIFFUSE_bound_liquid_fraction(pv_bag *b, double *mass_delta)
{
    RC(begin(b));
    pvk_clamp_props(G(b), &b->ph, FD(b, PV_F_C_L), FD(b, PV_F_C_G), FD(b, PV_F_RHO), FD(b, PV_F_MU), &b->rd,
                    b->d_sc, S(b));
    LAUNCHCHK();
    RC(pv_read_scalars(b));
    if (mass_delta) *mass_delta = b->h_sc->clamp_delta;
    return PV_OK;
}

int pv_VOF_DIFFUSE_update_phase_cache(pv_bag *b)
{
    RC(begin(b));
    pvk_phase_props(G(b), &b->ph, FD(b, PV_F_C_L), FD(b, PV_F_C_G), FD(b, PV_F_RHO), FD(b, PV_F_MU), S(b));
    LAUNCHCHK();
    return PV_OK;
}

int pv_VOF_update_density_viscosity(pv_bag *b) { return pv_VOF_DIFFUSE_update_phase_cache(b); }

int pv_VOF_DIFFUSE_step(pv_bag *b)
{
    RC(begin(b));
    RC(ch_step(b));
    RC(pv_read_scalars(b));
    b->st.clamp_delta[b->stage] = b->h_sc->clamp_delta;
    b->st.clamp_delta_interior[b->stage] = b->h_sc->clamp_delta_int;
    return PV_OK;
}

int pv_VOF_DIFFUSE_compute_f_sigma(pv_bag *b)
{
    RC(begin(b));
    pvk_psi_fsigma(G(b), &b->ph, FD(b, PV_F_C_L), FD(b, PV_F_C_G), NULL, FD(b, PV_F_PSI_LG), FD(b, PV_F_FSX_NEW),
                   FD(b, PV_F_FSY_NEW), S(b));
    LAUNCHCHK();
    return PV_OK;
}

int pv_VOF_apply_f_sigma_old(pv_bag *b)
{
    RC(begin(b));
    pvk_mom_force(G(b), &b->ph, 'u', FD(b, PV_F_RHO), FD(b, PV_F_FSX_OLD), FD(b, PV_F_U_RHS), 1, S(b));
    pvk_mom_force(G(b), &b->ph, 'v', FD(b, PV_F_RHO), FD(b, PV_F_FSY_OLD), FD(b, PV_F_V_RHS), 1, S(b));
    LAUNCHCHK();
    return PV_OK;
}

int pv_VOF_swap_f_sigma(pv_bag *b)
{
    double *t;
    if (!b) return pv_set_error(PV_EINVAL, "null handle");
    t = b->f[PV_F_FSX_OLD]; b->f[PV_F_FSX_OLD] = b->f[PV_F_FSX_NEW]; b->f[PV_F_FSX_NEW] = t;
    t = b->f[PV_F_FSY_OLD]; b->f[PV_F_FSY_OLD] = b->f[PV_F_FSY_NEW]; b->f[PV_F_FSY_NEW] = t;
    return PV_OK;
}

int pv_Velocity_add_gravity_2_RHS(pv_bag *b)
{
    RC(begin(b));
    pvk_mom_force(G(b), &b->ph, 'u', FD(b, PV_F_RHO), NULL, FD(b, PV_F_U_RHS), 2, S(b));
    pvk_mom_force(G(b), &b->ph, 'v', FD(b, PV_F_RHO), NULL, FD(b, PV_F_V_RHS), 2, S(b));
    LAUNCHCHK();
    return PV_OK;
}

/* ------------------------------------------------------------------ */
/* momentum                                                            */
/* ------------------------------------------------------------------ */

int pv_Velocity_cell_center(pv_bag *b)
{
    RC(begin(b));
    pvk_cell_center(G(b), FD(b, PV_F_U), FD(b, PV_F_V), FD(b, PV_F_U_BC), FD(b, PV_F_V_BC), S(b));
    pvk_halo_cell(G(b), FD(b, PV_F_U_BC), S(b));
    pvk_halo_cell(G(b), FD(b, PV_F_V_BC), S(b));
    LAUNCHCHK();
    return PV_OK;
}

static int mom_terms(pv_bag *b, int comp, int mode, int add_force)
{
    int u = (comp == 'u');
    pv_tbegin(b, PV_T_MOM_TERMS);
    pvk_mom_terms(G(b), &b->ph, b->stage, b->d_sc, comp, FD(b, PV_F_U), FD(b, PV_F_V), FD(b, PV_F_MU),
                  FD(b, PV_F_RHO), FD(b, PV_F_P), FD(b, u ? PV_F_U_EXPL_OLD : PV_F_V_EXPL_OLD),
                  FD(b, u ? PV_F_U_VISC_OLD : PV_F_V_VISC_OLD), FD(b, u ? PV_F_FSX_OLD : PV_F_FSY_OLD),
                  FD(b, u ? PV_F_U_RHS : PV_F_V_RHS), FD(b, u ? PV_F_U_EXPL : PV_F_V_EXPL),
                  FD(b, u ? PV_F_U_IMPL : PV_F_V_IMPL), FD(b, u ? PV_F_U_VISC : PV_F_V_VISC), mode, add_force,
                  S(b));
    pv_tend(b, PV_T_MOM_TERMS);
    LAUNCHCHK();
    return PV_OK;
}

int pv_Velocity_u_set_implicit_explicit(pv_bag *b) { RC(begin(b)); return mom_terms(b, 'u', 1, 0); }
int pv_Velocity_v_set_implicit_explicit(pv_bag *b) { RC(begin(b)); return mom_terms(b, 'v', 1, 0); }
int pv_Velocity_u_set_RHS(pv_bag *b) { RC(begin(b)); return mom_terms(b, 'u', 2, 0); }
int pv_Velocity_v_set_RHS(pv_bag *b) { RC(begin(b)); return mom_terms(b, 'v', 2, 0); }

int pv_Velocity_update_boundaries(pv_bag *b, int field, int type)
{
    RC(begin(b));
    if (field < 0 || field >= PV_NFIELDS) return pv_set_error(PV_EINVAL, "bad field");
    pvk_vel_bc(G(b), FD(b, field), comp_of_field(field), type, S(b));
    LAUNCHCHK();
    return PV_OK;
}

int pv_matVec(pv_bag *b, int out, int in, int component)
{
    RC(begin(b));
    if (component != 'u' && component != 'v') return pv_set_error(PV_EINVAL, "component must be 'u' or 'v'");
    pvk_mom_matvec(G(b), &b->ph, b->stage, b->d_sc, component, FD(b, in), FD(b, PV_F_RHO), FD(b, PV_F_MU),
                   FD(b, out), S(b));
    LAUNCHCHK();
    return PV_OK;
}

/* Velocity_solve_cg (lsolver/msolve_cg.c:525-722) */
/* Velocity_solve_cg (msolve_cg.c:525-722) for the components in mask
 * (1 = u, 2 = v): both run in the same launches and stop independently. */
static int mom_solve(pv_bag *b, int mask)
{
    int st = b->stage, n = 0, maxn = b->ph.CG_MAXIT + 1, c, chunk;
    pvk_mcg w = b->mcg;
    w.exact = b->m_exact;
    w.pc = b->m_pc;
    for (c = 0; c < 2; c++)
        if (!(mask & (1 << c))) w.nb[c] = 0;
    chunk = (mask & 1) ? b->mom_chunk[0] : b->mom_chunk[1];
    if ((mask & 2) && b->mom_chunk[1] > chunk) chunk = b->mom_chunk[1];
    pvk_mcg_init(G(b), &b->ph, st, b->d_sc, &w, S(b));
    if (w.pc) pvk_mcg_pc_start(G(b), &w, S(b));
    for (;;) {
        int k, done = 1;
        for (k = 0; k < chunk && n <= maxn; k++, n++) {
            pv_tbegin(b, PV_T_MOM_APPLY);
            pvk_mcg_iter(G(b), &b->ph, st, b->d_sc, &w, 0, n, S(b));
            pv_tend(b, PV_T_MOM_APPLY);
            pv_tbegin(b, PV_T_MOM_UPDATE);
            pvk_mcg_iter(G(b), &b->ph, st, b->d_sc, &w, 1, n, S(b));
            pv_tend(b, PV_T_MOM_UPDATE);
        }
        LAUNCHCHK();
        RC(pv_read_scalars(b));
        for (c = 0; c < 2; c++)
            if ((mask & (1 << c)) && !b->h_sc->m_done[c]) done = 0;
        if (done) break;
        if (n > maxn) return pv_set_error(PV_EHIP, "momentum CG did not stop (device flag)");
        chunk = 2;
    }
    if (w.pc) {
        pvk_mcg_pc_finish(G(b), b->d_sc, &w, S(b));
        LAUNCHCHK();
    }
    for (c = 0; c < 2; c++) {
        if (!(mask & (1 << c))) continue;
        b->mom_chunk[c] = (b->h_sc->m_iter[c] > 1) ? b->h_sc->m_iter[c] + 1 : 2;
        if (c == 0) b->st.n_u[st] = b->h_sc->m_iter[0]; else b->st.n_v[st] = b->h_sc->m_iter[1];
    }
    return PV_OK;
}

int pv_Velocity_solve_cg(pv_bag *b, int component, int *iters)
{
    RC(begin(b));
    if (component != 'u' && component != 'v') return pv_set_error(PV_EINVAL, "component must be 'u' or 'v'");
    RC(mom_solve(b, component == 'u' ? 1 : 2));
    if (iters) *iters = b->h_sc->m_iter[component == 'u' ? 0 : 1];
    return PV_OK;
}

/* ------------------------------------------------------------------ */
/* pressure                                                            */
/* ------------------------------------------------------------------ */

int pv_Pressure_set_RHS(pv_bag *b)
{
    RC(begin(b));
    pvk_p_rhs2(G(b), b->stage, b->d_sc, FD(b, PV_F_U), FD(b, PV_F_V), FD(b, PV_F_RHO), FD(b, PV_F_P_RHS), NULL,
               S(b));
    LAUNCHCHK();
    return PV_OK;
}

int pv_Pressure_compute_preconditioner(pv_bag *b)
{
    RC(begin(b));
    pvk_p_rhs2(G(b), b->stage, b->d_sc, FD(b, PV_F_U), FD(b, PV_F_V), FD(b, PV_F_RHO), FD(b, PV_FI_TMP),
               FD(b, PV_F_M_INV), S(b));
    LAUNCHCHK();
    return PV_OK;
}

int pv_Pressure_operator_variableCoeff(pv_bag *b, int out, int in)
{
    RC(begin(b));
    pvk_p_operator(G(b), FD(b, PV_F_RHO), FD(b, in), FD(b, out), S(b));
    LAUNCHCHK();
    return PV_OK;
}

int pv_Pressure_apply_BCs(pv_bag *b, int field)
{
    RC(begin(b));
    pvk_p_bc(G(b), FD(b, field), S(b));
    LAUNCHCHK();
    return PV_OK;
}

/* Jacobi-PCG (psolve_cg.c:130-293); phi starts at zero. */
static int p_solve_jacobi(pv_bag *b)
{
    double *phi = FD(b, PV_F_PHI), *rhs = FD(b, PV_F_P_RHS), *r = FD(b, PV_F_P_RES), *d = FD(b, PV_F_P_D);
    double *Ad = FD(b, PV_F_P_AD), *M = FD(b, PV_F_M_INV);
    int maxit = b->ph.P_CG_MAXIT, chunk = b->p_chunk, total = 0;
    double tol = b->ph.P_CG_ETOL;
    pvk_p_cg_init2(G(b), rhs, M, phi, r, d, 1, 0, b->d_sc, &b->rd, S(b));
    pvk_p_scalar2(G(b), b->d_sc, 0, 0, maxit, tol, &b->rd, 0, S(b));
    for (;;) {
        int k;
        for (k = 0; k < chunk; k++) {
            pvk_p_bc(G(b), d, S(b));
            pv_tbegin(b, PV_T_P_APPLY);
            pvk_p_cg_apply(G(b), b->d_sc, FD(b, PV_F_RHO), d, Ad, &b->rd, S(b));
            pv_tend(b, PV_T_P_APPLY);
            pvk_p_scalar2(G(b), b->d_sc, 1, 0, maxit, tol, &b->rd, 0, S(b));
            pv_tbegin(b, PV_T_P_UPDATE);
            pvk_p_cg_update(G(b), b->d_sc, phi, d, r, Ad, M, 1, &b->rd, S(b));
            pv_tend(b, PV_T_P_UPDATE);
            pvk_p_scalar2(G(b), b->d_sc, 2, 0, maxit, tol, &b->rd, 0, S(b));
            pv_tbegin(b, PV_T_P_DUPDATE);
            pvk_p_cg_dupdate(G(b), b->d_sc, r, M, 1, d, S(b));
            pv_tend(b, PV_T_P_DUPDATE);
        }
        LAUNCHCHK();
        total += chunk;
        RC(pv_read_scalars(b));
        if (b->h_sc->p_done) break;
        if (total > maxit + 16) return pv_set_error(PV_EHIP, "pressure CG did not stop (device flag)");
        chunk = 4;
    }
    b->p_chunk = (b->h_sc->p_iter > 4) ? b->h_sc->p_iter - 2 : 4;
    return PV_OK;
}

/* PCG with the FFT-x / tridiagonal-y preconditioner; relative-residual
 * stop.  Fused form (csrc/pv_pcg.hip): four launches per iteration (operator,
 * update + forward rows, stop test + y solves in place, inverse rows +
 * <r, z>), all scalars formed redundantly per block from the previous
 * kernel's partials.  The Cahn-Hilliard spectral start (chb > 0) needs the
 * y transforms instead of the tridiagonal solves: transpose, DCT-II /
 * divide / DCT-III, transpose back. */
static const int pcg_timer[7] = { PV_T_P_APPLY, PV_T_P_UPDATE, PV_T_P_TRANSPOSE, PV_T_FFT_COLS,
                                  PV_T_P_TRANSPOSE, PV_T_FFT_ROWS_INV, PV_T_FFT_COLS };
static const int pcg_tri_stages[4] = { 0, 1, 6, 5 };

static void pcg_iteration_w(pv_bag *b, const pvk_pcg *w, int n, double tol, int maxit)
{
    int k, st;
    if (w->chb > 0.0) {
        for (st = (n < 0) ? 2 : 0; st < 6; st++) {
            pv_tbegin(b, pcg_timer[st]);
            pvk_pcg_stage(G(b), &b->fft, b->d_sc, w, st, n, tol, maxit, S(b));
            pv_tend(b, pcg_timer[st]);
        }
        return;
    }
    for (k = (n < 0) ? 2 : 0; k < 4; k++) {
        st = pcg_tri_stages[k];
        pv_tbegin(b, pcg_timer[st]);
        pvk_pcg_stage(G(b), &b->fft, b->d_sc, w, st, n, tol, maxit, S(b));
        pv_tend(b, pcg_timer[st]);
    }
}

static void pcg_iteration(pv_bag *b, int n, double tol, int maxit) { pcg_iteration_w(b, &b->pcg, n, tol, maxit); }

/* start of the FFT-PCG: phi = 0, r = rhs - mean, z = M_fft r, ||b|| */
static void pcg_cold_start(pv_bag *b, double tol, int maxit)
{
    pv_tbegin(b, PV_T_P_UPDATE);
    pvk_pcg_start(G(b), &b->fft, b->d_sc, &b->pcg, tol, S(b));
    pv_tend(b, PV_T_P_UPDATE);
    pcg_iteration(b, -1, tol, maxit);
}

/* Iterations 0, 1, ... replayed from hipGraphs: graph kc holds iterations
 * [kc C, (kc + 1) C), captured on first use (kernel arguments are fixed per
 * handle, and the kernels stop themselves on the device flag). */
static int pcg_loop_graphs(pv_bag *b, double tol, int maxit)
{
    const int C = b->graph_chunk;
    int kc = 0, n = 0;
    for (;;) {
        if (kc >= b->npgraph) {
            int cap = (kc + 1) * 2, k;
            hipGraphExec_t *ng = (hipGraphExec_t *)realloc(b->pgraph, cap * sizeof(hipGraphExec_t));
            if (!ng) return pv_set_error(PV_ENOMEM, "graph table allocation failed");
            for (k = b->npgraph; k < cap; k++) ng[k] = NULL;
            b->pgraph = ng;
            b->npgraph = cap;
        }
        if (!b->pgraph[kc]) {
            hipGraph_t gr;
            int k;
            HIPCHK(hipStreamBeginCapture(S(b), hipStreamCaptureModeThreadLocal));
            for (k = 0; k < C && kc * C + k < maxit; k++) pcg_iteration(b, kc * C + k, tol, maxit);
            HIPCHK(hipStreamEndCapture(S(b), &gr));
            HIPCHK(hipGraphInstantiate(&b->pgraph[kc], gr, NULL, NULL, 0));
            HIPCHK(hipGraphDestroy(gr));
        }
        HIPCHK(hipGraphLaunch(b->pgraph[kc], S(b)));
        n = (kc + 1) * C;
        kc++;
        RC(pv_read_scalars(b));
        if (b->h_sc->p_done) break;
        if (n >= maxit) return pv_set_error(PV_EHIP, "pressure FFT-PCG did not stop (device flag)");
    }
    return PV_OK;
}

static void mg_graphs_free(pv_bag *b)
{
    int k;
    for (k = 0; k < b->nmggraph; k++)
        if (b->mggraph[k]) hipGraphExecDestroy(b->mggraph[k]);
    free(b->mggraph);
    b->mggraph = NULL;
    b->nmggraph = 0;
    b->mggraph_key = NULL;
}

void pv_graphs_free(pv_bag *b)
{
    int k;
    for (k = 0; k < b->npgraph; k++)
        if (b->pgraph[k]) hipGraphExecDestroy(b->pgraph[k]);
    free(b->pgraph);
    b->pgraph = NULL;
    b->npgraph = 0;
    mg_graphs_free(b);
}

/* iterations n0, n0 + 1, ... until the device stop flag */
static int pcg_loop(pv_bag *b, int n0, double tol, int maxit)
{
    int chunk = (b->p_chunk - n0 > 4) ? b->p_chunk - n0 : 4, n = n0;
    if (b->use_graphs && n0 == 0 && !b->timers_on) return pcg_loop_graphs(b, tol, maxit);
    for (;;) {
        int k;
        for (k = 0; k < chunk && n < maxit; k++, n++) pcg_iteration(b, n, tol, maxit);
        LAUNCHCHK();
        RC(pv_read_scalars(b));
        if (b->h_sc->p_done) break;
        if (n >= maxit) return pv_set_error(PV_EHIP, "pressure FFT-PCG did not stop (device flag)");
        chunk = 4;
    }
    b->p_chunk = (b->h_sc->p_iter > 4) ? b->h_sc->p_iter - 2 : 4;
    return PV_OK;
}

static int p_solve_fft_pcg(pv_bag *b)
{
    int maxit = b->ph.p_max_iter;
    double tol = b->ph.p_rel_tol;
    if (!b->fft.ok) return pv_set_error(PV_EINVAL, "FFT-PCG needs NXM = 2^a 3^b, a multiple of 16 and <= 4096, and 8 <= NYM <= 8192");
    pcg_cold_start(b, tol, maxit);
    return pcg_loop(b, 0, tol, maxit);
}

/* ---- MG-PCG: the CG above with one multigrid V-cycle as preconditioner ---- */
/* z = V(r_{n+1}) for CG iteration n (r_0 for n < 0); with tiled levels the
 * level-0 pre kernel first does the CG step of n and the next kernel its stop
 * test (n = -3: no test) */
static void mg_vcycle(pv_bag *b, int n, double tol, int maxit)
{
    const pvk_mg *m = b->mgp;
    int l;
    if (m->c0 > 0) {
        pv_tbegin(b, PV_T_MG_PRE0);
        pvk_mg_pre(m, 0, b->d_sc, &b->mgw, n, tol, maxit, S(b));
        pv_tend(b, PV_T_MG_PRE0);
    }
    pv_tbegin(b, PV_T_MG_COARSE);
    for (l = 1; l < m->c0; l++) pvk_mg_pre(m, l, b->d_sc, &b->mgw, n, tol, maxit, S(b));
    pvk_mg_coarse(m, G(b), b->d_sc, &b->mgw, n, tol, maxit, S(b));
    for (l = m->c0 - 1; l >= 1; l--) pvk_mg_post(m, l, b->d_sc, &b->mgw, n, S(b));
    pv_tend(b, PV_T_MG_COARSE);
    if (m->c0 > 0) {
        pv_tbegin(b, PV_T_MG_POST0);
        pvk_mg_post(m, 0, b->d_sc, &b->mgw, n, S(b));
        pv_tend(b, PV_T_MG_POST0);
    }
}

/* with tiled levels the level-0 pre-smoothing also does the CG step
 * (phi, r update); on one-workgroup grids pvk_mg_update does it */
static void mg_iteration(pv_bag *b, int n, double tol, int maxit)
{
    pv_tbegin(b, PV_T_P_APPLY);
    pvk_pcg_apply(G(b), b->d_sc, &b->mgw, n, S(b));
    pv_tend(b, PV_T_P_APPLY);
    if (b->mgp->c0 == 0) {
        pv_tbegin(b, PV_T_P_UPDATE);
        pvk_mg_update(G(b), b->d_sc, &b->mgw, n, S(b));
        pv_tend(b, PV_T_P_UPDATE);
    }
    mg_vcycle(b, n, tol, maxit);
}

/* Whether this solve builds the multigrid's bottom matrix: in the solve of
 * RK3 stage 0 and whenever the matrix is missing or two solves old.  The
 * bottom matrix is the exact inverse of the 16 x 16 (or smaller) last
 * level's operator, whose coefficients are 2 x 2 sums down from the
 * density: over the two later substeps of a step they change by a small
 * fraction, and a slightly stale exact coarse solve is still a symmetric
 * positive definite part of the preconditioner, so CG stays CG and its stop
 * test is unchanged.  The build (a band Cholesky and 256 column solves,
 * latency-bound) costs ~0.2 ms; once per step instead of per solve saves
 * two of three. */
static int mg_bottom_due(pv_bag *b)
{
    int *age = &b->mg_bottom_age[b->mgp == &b->mgf];
    if (*age < 0 || b->stage == 0 || *age >= 2) {
        *age = 0;
        return 1;
    }
    (*age)++;
    return 0;
}

/* faces, coarse levels, r = rhs - mean, phi = 0, z = V(r), stop test -1 */
static void mg_cold_start(pv_bag *b, double tol, int maxit)
{
    pvk_pcg_prep(G(b), b->d_sc, &b->mgw, S(b));
    pvk_mg_setup(b->mgp, b->mgw.bx, b->mgw.by, b->g.P, mg_bottom_due(b), S(b));
    pvk_mg_start(G(b), b->d_sc, &b->mgw, 0, S(b));
    mg_vcycle(b, -1, tol, maxit);
}

/* MG-PCG iterations 0, 1, ... replayed from hipGraphs: graph kc holds
 * iterations [kc C, (kc + 1) C), captured on first use for the hierarchy,
 * sweep count and W levels in use (the kernels stop themselves on the device
 * flag). */
static int mg_loop_graphs(pv_bag *b, double tol, int maxit)
{
    const int C = b->graph_chunk;
    int kc = 0, n = 0;
    if (b->mggraph_key != b->mgp || b->mggraph_nu != b->mgp->nu || b->mggraph_w != b->mgp->wmask ||
        b->mggraph_sh != b->mgp->sh0) {
        mg_graphs_free(b);
        b->mggraph_key = b->mgp;
        b->mggraph_nu = b->mgp->nu;
        b->mggraph_w = b->mgp->wmask;
        b->mggraph_sh = b->mgp->sh0;
    }
    for (;;) {
        if (kc >= b->nmggraph) {
            int cap = (kc + 1) * 2, k;
            hipGraphExec_t *ng = (hipGraphExec_t *)realloc(b->mggraph, cap * sizeof(hipGraphExec_t));
            if (!ng) return pv_set_error(PV_ENOMEM, "graph table allocation failed");
            for (k = b->nmggraph; k < cap; k++) ng[k] = NULL;
            b->mggraph = ng;
            b->nmggraph = cap;
        }
        if (!b->mggraph[kc]) {
            hipGraph_t gr;
            int k;
            HIPCHK(hipStreamBeginCapture(S(b), hipStreamCaptureModeThreadLocal));
            for (k = 0; k < C && kc * C + k < maxit; k++) mg_iteration(b, kc * C + k, tol, maxit);
            HIPCHK(hipStreamEndCapture(S(b), &gr));
            HIPCHK(hipGraphInstantiate(&b->mggraph[kc], gr, NULL, NULL, 0));
            HIPCHK(hipGraphDestroy(gr));
        }
        HIPCHK(hipGraphLaunch(b->mggraph[kc], S(b)));
        n = (kc + 1) * C;
        kc++;
        RC(pv_read_scalars(b));
        if (b->h_sc->p_done) break;
        if (n >= maxit) return pv_set_error(PV_EHIP, "pressure MG-PCG did not stop (device flag)");
    }
    return PV_OK;
}

/* iterations n0, n0 + 1, ... until the device stop flag; the first check
 * comes after the previous solve's iteration count plus one */
static int mg_loop(pv_bag *b, int n0, double tol, int maxit)
{
    int chunk = b->mg_chunk, n = n0;
    if (b->use_graphs && n0 == 0 && !b->timers_on) return mg_loop_graphs(b, tol, maxit);
    for (;;) {
        int k;
        for (k = 0; k < chunk && n < maxit; k++, n++) mg_iteration(b, n, tol, maxit);
        LAUNCHCHK();
        RC(pv_read_scalars(b));
        if (b->h_sc->p_done) break;
        if (n >= maxit) return pv_set_error(PV_EHIP, "pressure MG-PCG did not stop (device flag)");
        chunk = 2;
    }
    /* an iteration launched past the stop costs its early-exit launches, a
     * second round trip to the host costs more: aim one past the last count */
    b->mg_chunk = (b->h_sc->p_iter > 1) ? b->h_sc->p_iter + 1 : 2;
    return PV_OK;
}

/* energy |<r, V r>| of the current MG start (after its V-cycle) */
static int mg_energy(pv_bag *b, double *e)
{
    pvk_pcg_gamma(b->d_sc, &b->mgw, 0, 3, S(b));
    LAUNCHCHK();
    RC(pv_read_scalars(b));
    *e = fabs(b->h_sc->red[3]);
    return PV_OK;
}

/* restart the MG-PCG from the phi in PHI: r = rhs - mean - A phi, z = V r */
static void mg_warm_restart(pv_bag *b, double tol, int maxit)
{
    pvk_pcg_opx(G(b), &b->mgw, FD(b, PV_F_PHI), b->mgw.Ad, S(b));
    pvk_mg_start(G(b), b->d_sc, &b->mgw, 1, S(b));
    mg_vcycle(b, -2, tol, maxit);
}

static void fft_solve_field(pv_bag *b, double *src, double tol, int maxit);

/* MG-PCG.  The start is the best of up to three candidates by the energy
 * |<r, V r>| of their residuals (V the V-cycle, close to A^-1, so this is
 * close to the A-norm of the error): phi = 0 (always); with p_warm the last
 * correction of the same RK3 stage; with `learned` the better of those plus
 * the network's correction M_fft(r + s c(r, M_fft r, rho)) (the features it
 * was trained on, so the FFT tables are needed).  learn_start_res is the
 * chosen start's energy relative to the cold start's; a candidate that loses
 * counts as a fallback.  After 3 losses in a row the network's candidate is
 * not formed for the next 16 solves (then tried again); each further run of 3
 * losses doubles that skip, up to 256 solves, and a win resets it to 16, so
 * a network that does not help costs ever fewer tries.  Once the temporal
 * start has won 3 evaluations of its stage in a row (and no network
 * candidate is formed), it is taken without forming the cold candidate,
 * except in every 8th solve, which evaluates again; such a solve leaves
 * learn_start_res as it was. */
static int mg_solve(pv_bag *b, int learned)
{
    int maxit = b->ph.p_max_iter, st = b->stage, best = 0;
    double tol = b->ph.p_rel_tol, ec, eb, e;
    if (!b->mg_ok) return pv_set_error(PV_EINVAL, "MG-PCG is not available for this grid size");
    if (learned && !b->fft.ok) return pv_set_error(PV_EINVAL, "learned MG-PCG needs the FFT solve (NXM = 2^a 3^b, a multiple of 16)");
    if (learned && !b->cnn_ready)
        return pv_set_error(PV_EINVAL, "learned pressure solve needs a network (pv_cnn_init/set)");
    if (learned && b->net_skip > 0) {
        b->net_skip--;
        learned = 0;
    }
    if (b->p_warm && b->have_prev[st] && !learned && b->warm_wins[st] >= 3 && (++b->warm_solves[st] & 7) != 0) {
        /* trusted temporal start: ||b|| from the cold residual, then the start */
        pvk_pcg_prep(G(b), b->d_sc, &b->mgw, S(b));
        pvk_mg_setup(b->mgp, b->mgw.bx, b->mgw.by, b->g.P, mg_bottom_due(b), S(b));
        pvk_mg_start(G(b), b->d_sc, &b->mgw, 0, S(b));
        pvk_mg_check(b->d_sc, &b->mgw, -1, tol, maxit, S(b));
        pvk_mg_take(G(b), b->d_sc, FD(b, PV_FI_PHI0 + st), FD(b, PV_F_PHI), S(b));
        mg_warm_restart(b, tol, maxit);
        return mg_loop(b, 0, tol, maxit);
    }
    mg_cold_start(b, tol, maxit);
    b->learn_start_res = 1.0;
    if (!b->p_warm && !learned) return mg_loop(b, 0, tol, maxit);
    RC(mg_energy(b, &ec));
    eb = ec;
    if (b->h_sc->p_done) return PV_OK;
    if (b->p_warm && b->have_prev[st]) {
        pvk_learn_copy(G(b), FD(b, PV_FI_PHI0 + st), FD(b, PV_F_PHI), S(b));
        mg_warm_restart(b, tol, maxit);
        RC(mg_energy(b, &e));
        if (e < eb) {
            eb = e;
            best = 1;
            b->warm_wins[st]++;
        } else {
            b->st.p_fallback[st] += 1;
            b->warm_wins[st] = 0;
        }
        b->warm_solves[st] = 0;
    }
    if (learned && !b->h_sc->p_done) {
        if (best == 0 && b->p_warm && b->have_prev[st]) {   /* back to the cold state */
            pvk_mg_start(G(b), b->d_sc, &b->mgw, 0, S(b));
            mg_vcycle(b, -1, tol, maxit);
        }
        fft_solve_field(b, b->mgw.r, tol, maxit);            /* z = M_fft r: the second feature */
        RC(pv_learn_source(b, FD(b, PV_FI_LTMP)));
        fft_solve_field(b, FD(b, PV_FI_LTMP), tol, maxit);
        pvk_learn_axpy(G(b), b->pcg.z, FD(b, PV_F_PHI), S(b));
        mg_warm_restart(b, tol, maxit);
        RC(mg_energy(b, &e));
        if (e < eb) {
            eb = e;
            best = 2;
            b->net_losses = 0;
            b->net_backoff = 16;
        } else {
            b->st.p_fallback[st] += 1;
            if (++b->net_losses >= 3) {
                if (b->net_backoff < 16) b->net_backoff = 16;
                b->net_losses = 0;
                b->net_skip = b->net_backoff;
                if (b->net_backoff < 256) b->net_backoff *= 2;
            }
        }
    }
    if (!b->h_sc->p_done && ((learned && best != 2) || (!learned && best == 0 && b->p_warm && b->have_prev[st]))) {
        /* the winner is not the state in place: rebuild it */
        if (best == 0) {
            pvk_mg_start(G(b), b->d_sc, &b->mgw, 0, S(b));
            mg_vcycle(b, -1, tol, maxit);
        } else {
            pvk_learn_copy(G(b), FD(b, PV_FI_PHI0 + st), FD(b, PV_F_PHI), S(b));
            mg_warm_restart(b, tol, maxit);
        }
        LAUNCHCHK();
        RC(pv_read_scalars(b));
    }
    b->learn_start_res = (ec > 0.0) ? eb / ec : 0.0;
    if (b->h_sc->p_done) return PV_OK;
    return mg_loop(b, 0, tol, maxit);
}

static int p_solve_mg(pv_bag *b) { return mg_solve(b, 0); }
static int p_solve_learned_mg(pv_bag *b) { return mg_solve(b, 1); }

/* The plain FFT-PCG with training snapshots (phi_k, z_k, network features)
 * at the collector's iterations; crops are harvested after convergence. */
static int p_solve_collect(pv_bag *b)
{
    pv_learn *L = b->learn;
    int maxit = b->ph.p_max_iter, n = 0, si = 0;
    double tol = b->ph.p_rel_tol;
    if (!b->fft.ok) return pv_set_error(PV_EINVAL, "FFT-PCG needs NXM = 2^a 3^b, a multiple of 16 and <= 4096, and 8 <= NYM <= 8192");
    pcg_cold_start(b, tol, maxit);
    for (;;) {
        int next;
        LAUNCHCHK();
        RC(pv_read_scalars(b));
        if (b->h_sc->p_done) break;
        if (si < L->nsnap && n == L->snap_iter[si]) RC(pv_learn_snapshot(b, si++));
        if (n >= maxit) return pv_set_error(PV_EHIP, "pressure FFT-PCG did not stop (device flag)");
        if (si >= L->nsnap) {                  /* no snapshot left: the checked loop */
            RC(pcg_loop(b, n, tol, maxit));
            break;
        }
        next = L->snap_iter[si];
        for (; n < next && n < maxit; n++) pcg_iteration(b, n, tol, maxit);
    }
    return pv_learn_harvest(b, si);
}

/* z = M_fft src (the FFT stages on a given field; no PCG update, no stop test) */
static void fft_solve_field(pv_bag *b, double *src, double tol, int maxit)
{
    pvk_pcg w = b->pcg;
    w.r = src;
    pvk_pcg_stage(G(b), &b->fft, b->d_sc, &w, 1, -3, tol, maxit, S(b));
    pcg_iteration_w(b, &w, -3, tol, maxit);
}

/* CH spectral start: C += (a_dt + bih L^2)^-1 r with r = b - A C (the FFT
 * stages of the pressure PCG with the CH divide), then r = b - A C again from
 * the kept b (PV_FI_TMP) for the CG's stop test */
static void ch_spectral_start(pv_bag *b, int st)
{
    pvk_pcg w = b->pcg;
    w.r = FD(b, PV_F_CH_RES);
    w.chb = b->ph.bih;
    w.chstage = st;
    pvk_pcg_stage(G(b), &b->fft, b->d_sc, &w, 1, -3, 0.0, 1, S(b));
    pcg_iteration_w(b, &w, -3, 0.0, 1);
    pvk_learn_axpy(G(b), w.z, FD(b, PV_F_C_L), S(b));
    pvk_bc_ch(G(b), FD(b, PV_F_C_L), S(b));
    pvk_ccg_init(G(b), &b->ph, st, b->d_sc, &b->ccg, FD(b, PV_F_C_L), NULL, NULL, NULL, FD(b, PV_FI_TMP),
                 FD(b, PV_F_CH_RES), FD(b, PV_F_CH_DIR), NULL, S(b));
}

/* warm restart of the FFT-PCG from the current phi: r = rhs - mean - A phi */
static int pcg_warm_restart(pv_bag *b, double tol, int maxit)
{
    pvk_pcg_opx(G(b), &b->pcg, FD(b, PV_F_PHI), b->pcg.Ad, S(b));
    pvk_pcg_stage(G(b), &b->fft, b->d_sc, &b->pcg, 1, -2, tol, maxit, S(b));
    pcg_iteration(b, -2, tol, maxit);
    LAUNCHCHK();
    return pv_read_scalars(b);
}

/* The learned warm start phi0 = M_fft(b + s c(b, M_fft b, rho)).  It is
 * judged by the energy-norm proxy <r, M_fft r> of its residual against the
 * cold start's (CG converges in the A-norm): kept if smaller, else the cold
 * start is restored.  Leaves the solve ready for iteration 0. */
static int learned_warm_start(pv_bag *b, double tol, int maxit)
{
    double gw, gc;
    pcg_cold_start(b, tol, maxit);
    pvk_pcg_gamma(b->d_sc, &b->pcg, 0, 4, S(b));
    LAUNCHCHK();
    RC(pv_read_scalars(b));
    if (b->h_sc->p_done) return PV_OK;
    RC(pv_learn_source(b, FD(b, PV_FI_LTMP)));
    fft_solve_field(b, FD(b, PV_FI_LTMP), tol, maxit);
    pvk_learn_copy(G(b), b->pcg.z, FD(b, PV_F_PHI), S(b));
    RC(pcg_warm_restart(b, tol, maxit));
    pvk_pcg_gamma(b->d_sc, &b->pcg, 0, 3, S(b));
    LAUNCHCHK();
    RC(pv_read_scalars(b));
    /* M_fft is negative definite: the energies are -<r, M_fft r> */
    gw = fabs(b->h_sc->red[3]);
    gc = fabs(b->h_sc->red[4]);
    b->learn_start_res = (gc > 0.0) ? gw / gc : 0.0;
    if (!b->h_sc->p_done && gw >= gc) {
        b->st.p_fallback[b->stage] += 1;
        pcg_cold_start(b, tol, maxit);
        LAUNCHCHK();
        RC(pv_read_scalars(b));
    }
    return PV_OK;
}

/* Learned solve: the learned warm start, then the FFT-PCG iterations. */
static int p_solve_learned(pv_bag *b)
{
    int maxit = b->ph.p_max_iter;
    double tol = b->ph.p_rel_tol;
    if (!b->fft.ok) return pv_set_error(PV_EINVAL, "FFT-PCG needs NXM = 2^a 3^b, a multiple of 16 and <= 4096, and 8 <= NYM <= 8192");
    if (!b->cnn_ready) return pv_set_error(PV_EINVAL, "learned pressure solve needs a network (pv_cnn_init/set)");
    RC(learned_warm_start(b, tol, maxit));
    if (b->h_sc->p_done) return PV_OK;
    return pcg_loop(b, 0, tol, maxit);
}

/* Learned flexible CG: the learned warm start, then in every iteration the
 * learned preconditioner z = M_fft(r + s c(r, M_fft r, rho)) with beta =
 * -<z_{n+1}, A d_n> / <d_n, A d_n>.  If the residual stalls (less than a
 * factor 2 over 24 iterations) the solve restarts as the plain FFT-PCG from
 * the current iterate. */
static int p_solve_learned_fcg(pv_bag *b)
{
    int maxit = b->ph.p_max_iter, n = 0, since = 0, done_fcg;
    double tol = b->ph.p_rel_tol, best;
    pvk_pcg w;
    if (!b->fft.ok) return pv_set_error(PV_EINVAL, "FFT-PCG needs NXM = 2^a 3^b, a multiple of 16 and <= 4096, and 8 <= NYM <= 8192");
    if (!b->cnn_ready) return pv_set_error(PV_EINVAL, "learned pressure solve needs a network (pv_cnn_init/set)");
    RC(learned_warm_start(b, tol, maxit));
    if (b->h_sc->p_done) return PV_OK;
    w = b->pcg;
    w.flex = 1;
    w.nrz = b->NYM;
#define LEARNED_Z(slot)                                                                    \
    do {                                                                                   \
        RC(pv_learn_source(b, FD(b, PV_FI_LTMP)));                                         \
        fft_solve_field(b, FD(b, PV_FI_LTMP), tol, maxit);                                 \
        pvk_learn_dots(G(b), b->pcg.z, b->pcg.r, b->pcg.Ad, w.rz[slot], w.zad, S(b));      \
    } while (0)
    LEARNED_Z(0);
    best = b->h_sc->p_res;
    for (;;) {
        int k;
        for (k = 0; k < 8 && n < maxit; k++, n++) {
            pcg_iteration_w(b, &w, n, tol, maxit);
            LEARNED_Z((n + 1) & 1);
        }
        LAUNCHCHK();
        RC(pv_read_scalars(b));
        if (b->h_sc->p_done) return PV_OK;
        if (n >= maxit) return pv_set_error(PV_EHIP, "learned FCG did not stop (device flag)");
        since += 8;
        if (b->h_sc->p_res < 0.5 * best) { best = b->h_sc->p_res; since = 0; }
        if (since >= 24) break;                                 /* stalled */
    }
#undef LEARNED_Z
    done_fcg = b->h_sc->p_iter;
    b->st.p_fallback[b->stage] += 1;
    RC(pcg_warm_restart(b, tol, maxit));
    if (!b->h_sc->p_done) RC(pcg_loop(b, 0, tol, maxit));
    b->st.n_p[b->stage] += done_fcg;
    return PV_OK;
}

/* Pressure_solve_cg + the mean-removal gauge (psolve_cg.c:255-279) */
static int p_solve(pv_bag *b)
{
    if (b->prm.pressure_solver == PV_PSOLVE_FFT_PCG)
        RC((b->learn && b->learn->collecting) ? p_solve_collect(b) : p_solve_fft_pcg(b));
    else if (b->prm.pressure_solver == PV_PSOLVE_LEARNED) RC(p_solve_learned(b));
    else if (b->prm.pressure_solver == PV_PSOLVE_LEARNED_FCG) RC(p_solve_learned_fcg(b));
    else if (b->prm.pressure_solver == PV_PSOLVE_MG_PCG) RC(p_solve_mg(b));
    else if (b->prm.pressure_solver == PV_PSOLVE_LEARNED_MG) RC(p_solve_learned_mg(b));
    else RC(p_solve_jacobi(b));
    if (b->h_sc->p_done != 4) {
        pvk_p_mean(G(b), FD(b, PV_F_PHI), &b->rd, S(b));
        pvk_p_mean_finalize(G(b), b->d_sc, &b->rd, S(b));
        pvk_p_sub_mean(G(b), b->d_sc, FD(b, PV_F_PHI), 1, S(b));
        LAUNCHCHK();
    }
    b->st.n_p[b->stage] += b->h_sc->p_iter;
    b->st.p_res[b->stage] = b->h_sc->p_res;
    if (b->p_warm) {
        pvk_learn_copy(G(b), FD(b, PV_F_PHI), FD(b, PV_FI_PHI0 + b->stage), S(b));
        b->have_prev[b->stage] = 1;
    }
    return PV_OK;
}

int pv_Pressure_solve_cg(pv_bag *b, int *iters)
{
    RC(begin(b));
    RC(p_solve(b));
    if (iters) *iters = b->h_sc->p_iter;
    return PV_OK;
}

int pv_Pressure_project_velocity(pv_bag *b)
{
    RC(begin(b));
    pvk_p_bc(G(b), FD(b, PV_F_PHI), S(b));
    pvk_project(G(b), b->stage, b->d_sc, FD(b, PV_F_RHO), FD(b, PV_F_PHI), FD(b, PV_F_P), FD(b, PV_F_U),
                FD(b, PV_F_V), S(b));
    pvk_p_bc(G(b), FD(b, PV_F_P), S(b));
    LAUNCHCHK();
    return PV_OK;
}

int pv_mg_vcycle_apply(pv_bag *b)
{
    RC(begin(b));
    if (!b->mg_ok) return pv_set_error(PV_EINVAL, "MG-PCG is not available for this grid size");
    pvk_pcg_prep(G(b), b->d_sc, &b->mgw, S(b));
    pvk_mg_setup(b->mgp, b->mgw.bx, b->mgw.by, b->g.P, 1, S(b));
    b->mg_bottom_age[b->mgp == &b->mgf] = 0;
    mg_vcycle(b, -3, 0.0, 1);
    pvk_learn_copy(G(b), b->mgw.z, FD(b, PV_F_SCRATCH0), S(b));
    LAUNCHCHK();
    return pv_read_scalars(b);
}

int pv_Pressure_compute_velocity_divergence(pv_bag *b, double *div_max)
{
    RC(begin(b));
    pvk_divergence(G(b), FD(b, PV_F_U), FD(b, PV_F_V), &b->rd, S(b));
    pvk_div_finalize(G(b), b->d_sc, &b->rd, S(b));
    LAUNCHCHK();
    RC(pv_read_scalars(b));
    if (div_max) *div_max = b->h_sc->div_max;
    return PV_OK;
}

int pv_Pressure_solve_fft(pv_bag *b, int rhs_field, int out_field)
{
    RC(begin(b));
    if (!b->fft.ok) return pv_set_error(PV_EINVAL, "FFT solve needs NXM = 2^a 3^b, a multiple of 16 and <= 4096, and 8 <= NYM <= 8192");
    if (rhs_field < 0 || rhs_field >= PV_NFIELDS || out_field < 0 || out_field >= PV_NFIELDS)
        return pv_set_error(PV_EINVAL, "bad field id");
    {   /* the preconditioner's stages on the given fields, with the reference's pin */
        pvk_pcg w = b->pcg;
        w.r = FD(b, rhs_field);
        w.z = FD(b, out_field);
        w.keep_pin = 1;
        pvk_pcg_clear_done(b->d_sc, S(b));
        pvk_pcg_stage(G(b), &b->fft, b->d_sc, &w, 1, -3, 0.0, 1, S(b));
        pvk_pcg_stage(G(b), &b->fft, b->d_sc, &w, 6, -3, 0.0, 1, S(b));
        pvk_pcg_stage(G(b), &b->fft, b->d_sc, &w, 5, -3, 0.0, 1, S(b));
    }
    pvk_halo_cell(G(b), FD(b, out_field), S(b));
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(S(b)));
    return PV_OK;
}

/* ------------------------------------------------------------------ */
/* substep, time step, RK3                                             */
/* ------------------------------------------------------------------ */

/* Temporal_int.c:1051-1107 */
static int pressure_passes(pv_bag *b)
{
    int passes = 0, st = b->stage;
    double div;
    b->st.n_p[st] = 0;
    b->st.p_fallback[st] = 0;
    do {
        pvk_fill(FD(b, PV_F_PHI) - 3 * (size_t)b->P - 3, (size_t)b->P * b->NYT, 0.0, S(b));
        pvk_p_rhs2(G(b), st, b->d_sc, FD(b, PV_F_U), FD(b, PV_F_V), FD(b, PV_F_RHO), FD(b, PV_F_P_RHS),
                   FD(b, PV_F_M_INV), S(b));
        LAUNCHCHK();
        RC(p_solve(b));
        RC(pv_Pressure_project_velocity(b));
        pvk_vel_bc(G(b), FD(b, PV_F_U), 'u', PV_VEL_TYPE_NORMAL, S(b));
        pvk_vel_bc(G(b), FD(b, PV_F_V), 'v', PV_VEL_TYPE_NORMAL, S(b));
        RC(pv_Pressure_compute_velocity_divergence(b, &div));
        passes++;
        if (passes > 10) {
            b->st.n_pass[st] = passes;
            return pv_set_error(PV_ENOCONV, "Poisson solver did not converge (div %g after %d passes)", div,
                                passes);
        }
    } while (div > 1e-6);
    b->st.n_pass[st] = passes;
    b->st.div_max[st] = div;
    return PV_OK;
}

/* Temporal_int.c:563-1112 (no IBM / CONC / LES / RANS) */
int pv_Temporal_int_all_the_equations(pv_bag *b)
{
    int st;
    RC(begin(b));
    st = b->stage;
    RC(ch_step(b));
    RC(mom_terms(b, 'u', 2, 1));
    RC(mom_terms(b, 'v', 2, 1));
    pvk_psi_fsigma(G(b), &b->ph, FD(b, PV_F_C_L), FD(b, PV_F_C_G), NULL, FD(b, PV_F_PSI_LG), FD(b, PV_F_FSX_NEW),
                   FD(b, PV_F_FSY_NEW), S(b));
    LAUNCHCHK();
    if (b->m_warm && b->have_mprev[st]) {
        pvk_mom_warm(G(b), 0, b->dt / b->mprev_dt[st], FD(b, PV_F_U), FD(b, PV_F_V), FD(b, PV_FI_MUS),
                     FD(b, PV_FI_MVS), FD(b, PV_FI_MU0 + st), FD(b, PV_FI_MV0 + st), S(b));
        /* the CG's first residual reads ghosts (periodic image, wall mirrors)
         * of the start: make them those of the shifted start */
        pvk_vel_bc(G(b), FD(b, PV_F_U), 'u', PV_VEL_TYPE_NORMAL, S(b));
        pvk_vel_bc(G(b), FD(b, PV_F_V), 'v', PV_VEL_TYPE_NORMAL, S(b));
    } else if (b->m_warm)
        pvk_mom_warm(G(b), 0, 0.0, FD(b, PV_F_U), FD(b, PV_F_V), FD(b, PV_FI_MUS), FD(b, PV_FI_MVS),
                     FD(b, PV_FI_MU0 + st), FD(b, PV_FI_MV0 + st), S(b));
    RC(mom_solve(b, 3));
    if (b->m_warm) {
        pvk_mom_warm(G(b), 1, 0.0, FD(b, PV_F_U), FD(b, PV_F_V), FD(b, PV_FI_MUS), FD(b, PV_FI_MVS),
                     FD(b, PV_FI_MU0 + st), FD(b, PV_FI_MV0 + st), S(b));
        b->have_mprev[st] = 1;
        b->mprev_dt[st] = b->dt;
    }
    pvk_vel_bc(G(b), FD(b, PV_F_U), 'u', PV_VEL_TYPE_NORMAL, S(b));
    pvk_vel_bc(G(b), FD(b, PV_F_V), 'v', PV_VEL_TYPE_NORMAL, S(b));
    LAUNCHCHK();
    RC(pressure_passes(b));
    b->st.clamp_delta[st] = b->h_sc->clamp_delta;
    b->st.clamp_delta_interior[st] = b->h_sc->clamp_delta_int;
    RC(pv_VOF_swap_f_sigma(b));
    return PV_OK;
}

/* Dtime.c:24-249 */
int pv_Dtime_cfl(pv_bag *b, double *dt_out)
{
    RC(begin(b));
    pvk_cfl(G(b), FD(b, PV_F_U), FD(b, PV_F_V), &b->rd, S(b));
    pvk_cfl_finalize2(G(b), &b->ph, b->d_sc, b->dx_min, &b->rd, S(b));
    LAUNCHCHK();
    RC(pv_read_scalars(b));
    if (dt_out) *dt_out = b->h_sc->dt_new;
    return PV_OK;
}

/* Temporal_int.c:399-448 */
int pv_Temporal_int_rk3(pv_bag *b, int nsteps)
{
    int n, rk;
    RC(begin(b));
    for (n = 0; n < nsteps; n++) {
        for (rk = 0; rk < 3; rk++) {
            if (b->time_nonzero && rk == 0) {
                pvk_cfl(G(b), FD(b, PV_F_U), FD(b, PV_F_V), &b->rd, S(b));
                pvk_cfl_finalize2(G(b), &b->ph, b->d_sc, b->dx_min, &b->rd, S(b));
                pvk_take_dt(b->d_sc, S(b));
                LAUNCHCHK();
            }
            b->stage = rk;
            RC(pv_Temporal_int_all_the_equations(b));
            pvk_advance_time(b->d_sc, rk, S(b));
            LAUNCHCHK();
            b->time_nonzero = 1;
        }
        b->ntime++;
    }
    RC(pv_read_scalars(b));
    return PV_OK;
}

/* Cart3d.c:1094-1130, VOF_DIFFUSE.c:524-568, Temporal_int.c:163-199 */
int pv_init_from_F(pv_bag *b, const double *F_host)
{
    RC(begin(b));
    if (!F_host) return pv_set_error(PV_EINVAL, "null F");
    pvk_fill(b->slabmem, b->alloc_elems * PV_FI_END, 0.0, S(b));
    LAUNCHCHK();
    RC(pv_set_field(b, PV_F_F, F_host));
    pvk_bc_ch(G(b), FD(b, PV_F_F), S(b));
    RC(pv_VOF_DIFFUSE_init(b));
    pvk_psi_fsigma(G(b), &b->ph, FD(b, PV_F_C_L), FD(b, PV_F_C_G), NULL, FD(b, PV_F_PSI_LG), FD(b, PV_F_FSX_NEW),
                   FD(b, PV_F_FSY_NEW), S(b));
    pvk_copy(G(b), FD(b, PV_F_FSX_NEW), FD(b, PV_F_FSX_OLD), S(b));
    pvk_copy(G(b), FD(b, PV_F_FSY_NEW), FD(b, PV_F_FSY_OLD), S(b));
    LAUNCHCHK();
    memset(&b->st, 0, sizeof b->st);
    b->time = 0.0; b->dt = b->dt_old = b->prm.default_dt; b->stage = 0; b->ntime = 1; b->time_nonzero = 0;
    b->have_mprev[0] = b->have_mprev[1] = b->have_mprev[2] = 0;
    b->mg_bottom_age[0] = b->mg_bottom_age[1] = -1;                /* rebuilt at the next solve */
    memset(b->h_sc, 0, sizeof(pvk_scal));
    RC(pv_sync_scalars_to_device(b));
    RC(pv_Velocity_cell_center(b));
    HIPCHK(hipStreamSynchronize(S(b)));
    return PV_OK;
}

/* Replay the pressure FFT-PCG from hipGraphs of `chunk` iterations (on != 0)
 * or launch it eagerly.  Graphs are dropped when the setting changes. */
int pv_set_graphs(pv_bag *b, int on, int chunk)
{
    RC(begin(b));
    if (on && (chunk < 1 || chunk > 256)) return pv_set_error(PV_EINVAL, "graph chunk must be in [1, 256]");
    HIPCHK(hipStreamSynchronize(S(b)));
    if (!on || chunk != b->graph_chunk) pv_graphs_free(b);
    b->use_graphs = on ? 1 : 0;
    if (on) b->graph_chunk = chunk;
    return PV_OK;
}

/* Time one candidate chunk (0 = eager launches): a warm-up solve that
 * captures the graphs, then two timed solves. */
static int time_chunk(pv_bag *b, int chunk, float *ms)
{
    hipEvent_t e0 = NULL, e1 = NULL;
    int rep, its, rc;
    if ((rc = pv_set_graphs(b, chunk > 0, chunk > 0 ? chunk : 16))) return rc;
    if ((rc = pv_Pressure_solve_cg(b, &its))) return rc;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = pv_set_error(PV_EHIP, "autotune: hipEventCreate failed");
    } else if (hipEventRecord(e0, S(b)) != hipSuccess) {
        rc = pv_set_error(PV_EHIP, "autotune: hipEventRecord failed");
    } else {
        for (rep = 0; rep < 2 && rc == PV_OK; rep++) rc = pv_Pressure_solve_cg(b, &its);
        if (rc == PV_OK && (hipEventRecord(e1, S(b)) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                            hipEventElapsedTime(ms, e0, e1) != hipSuccess))
            rc = pv_set_error(PV_EHIP, "autotune: event timing failed");
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    return rc;
}

/* Pick the fastest of eager launches and graph chunks 8, 16, 32 for the
 * pressure solve of the current state (the solve restarts from phi = 0 and
 * its right-hand side is left as it is, so the trajectory is unchanged).  On
 * a failure the handle keeps the graph setting it had before the call. */
int pv_autotune(pv_bag *b, int *choice)
{
    static const int chunks[4] = {0, 8, 16, 32};
    double best = 1e300;
    int k, pick = 0, rc = PV_OK;
    const int use0 = b ? b->use_graphs : 0, chunk0 = b ? b->graph_chunk : 0;
    RC(begin(b));
    if (b->prm.pressure_solver != PV_PSOLVE_FFT_PCG) return pv_set_error(PV_EINVAL, "autotune is for FFT-PCG");
    for (k = 0; k < 4 && rc == PV_OK; k++) {
        float ms = 0.0f;
        if ((rc = time_chunk(b, chunks[k], &ms)) == PV_OK && ms < best) { best = ms; pick = k; }
    }
    if (rc != PV_OK) {
        char msg[512];
        snprintf(msg, sizeof msg, "%s", pv_last_error());
        pv_set_graphs(b, use0, use0 ? chunk0 : 16);
        return pv_set_error(rc, "%s", msg);
    }
    RC(pv_set_graphs(b, chunks[pick] > 0, chunks[pick] > 0 ? chunks[pick] : 16));
    if (choice) *choice = chunks[pick];
    return PV_OK;
}

/* Pick the fastest multigrid cycle (W on 0 .. 3 of the in-workgroup levels)
 * and sweep count (1, 2) for the MG-PCG pressure solve of the right-hand side
 * in place (call it after a step).  Each candidate solves from phi = 0 with
 * no temporal or network start; PHI and the iteration-chunk guess are put
 * back afterwards and the start bookkeeping is not touched, so the
 * trajectory is the one the chosen settings give without tuning. */
/* event-timed ms of `iters` momentum CG iterations (u and v) from the
 * current fields, stop test off (tolerance 0, so the solve runs exactly
 * iters iterations) */
static int mom_time(pv_bag *b, int iters, float *ms)
{
    hipEvent_t e0 = NULL, e1 = NULL;
    pvk_mcg w = b->mcg;
    int n, rc = PV_OK;
    w.exact = b->m_exact;
    w.tol = 0.0;
    w.maxit = iters;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = pv_set_error(PV_EHIP, "tile autotune: hipEventCreate failed");
    } else {
        hipEventRecord(e0, S(b));
        pvk_mcg_init(G(b), &b->ph, b->stage, b->d_sc, &w, S(b));
        for (n = 0; n <= iters; n++) {
            pvk_mcg_iter(G(b), &b->ph, b->stage, b->d_sc, &w, 0, n, S(b));
            pvk_mcg_iter(G(b), &b->ph, b->stage, b->d_sc, &w, 1, n, S(b));
        }
        hipEventRecord(e1, S(b));
        if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(ms, e0, e1) != hipSuccess)
            rc = pv_set_error(PV_EHIP, "tile autotune: momentum timing failed");
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    return rc;
}

/* Tile shapes: rows per block of the momentum CG (1, 2, 4), timed on 16
 * iterations of the current momentum system (u and v saved and restored),
 * and the level-0 tile shape of the multigrid (64 x 16, 32 x 16, 32 x 8),
 * timed on cold pressure solves in place as pv_autotune_mg does; the
 * fastest of each is kept.  A failure leaves both as they were. */
int pv_autotune_tiles(pv_bag *b, int *mom_rows, int *mg_tile0)
{
    const int rows_c[3] = {1, 2, 4};
    double *save = NULL;
    size_t bytes;
    int k, rep, rc = PV_OK, rows0 = b ? b->mcg.band : 2, sh00 = b ? b->mg.sh0 : -1, pick_r = rows0, pick_s = sh00;
    int p_warm, chunk;
    double start_res, best = 1e300;
    RC(begin(b));
    bytes = sizeof(double) * pv_field_elems(b);
    if (hipMalloc((void **)&save, 2 * bytes) != hipSuccess)
        return pv_set_error(PV_ENOMEM, "tile autotune: hipMalloc of %zu bytes failed", 2 * bytes);
    /* momentum rows */
    const int saved_uv =
        hipMemcpyAsync(save, FD(b, PV_F_U) - pv_field_front(b), bytes, hipMemcpyDeviceToDevice, S(b)) == hipSuccess &&
        hipMemcpyAsync(save + bytes / sizeof(double), FD(b, PV_F_V) - pv_field_front(b), bytes,
                       hipMemcpyDeviceToDevice, S(b)) == hipSuccess;
    if (!saved_uv) rc = pv_set_error(PV_EHIP, "tile autotune: saving u, v failed");
    for (k = 0; k < 3 && rc == PV_OK; k++) {
        float ms = 0.0f, ms2 = 0.0f;
        if ((rc = pv_set_momentum_rows(b, rows_c[k]))) break;
        if ((rc = mom_time(b, 16, &ms)) || (rc = mom_time(b, 16, &ms2))) break;   /* the first one warms up */
        if (ms2 < best) { best = ms2; pick_r = rows_c[k]; }
    }
    if (saved_uv && (hipMemcpyAsync(FD(b, PV_F_U) - pv_field_front(b), save, bytes, hipMemcpyDeviceToDevice,
                                    S(b)) != hipSuccess ||
                     hipMemcpyAsync(FD(b, PV_F_V) - pv_field_front(b), save + bytes / sizeof(double), bytes,
                                    hipMemcpyDeviceToDevice, S(b)) != hipSuccess ||
                     hipStreamSynchronize(S(b)) != hipSuccess))
        if (rc == PV_OK) rc = pv_set_error(PV_EHIP, "tile autotune: restoring u, v failed");
    if (rc == PV_OK) rc = pv_set_momentum_rows(b, pick_r);
    /* multigrid level-0 tiles */
    if (rc == PV_OK && b->mg_ok && b->mg.c0 > 0 &&
        (b->prm.pressure_solver == PV_PSOLVE_MG_PCG || b->prm.pressure_solver == PV_PSOLVE_LEARNED_MG)) {
        const int saved = (hipMemcpyAsync(save, FD(b, PV_F_PHI) - pv_field_front(b), bytes, hipMemcpyDeviceToDevice,
                                          S(b)) == hipSuccess);
        if (!saved) rc = pv_set_error(PV_EHIP, "tile autotune: saving PHI failed");
        p_warm = b->p_warm;
        chunk = b->mg_chunk;
        start_res = b->learn_start_res;
        b->p_warm = 0;
        best = 1e300;
        for (k = 0; k < 3 && rc == PV_OK; k++) {
            hipEvent_t e0 = NULL, e1 = NULL;
            float ms = 0.0f;
            if ((rc = pv_set_mg_tile0(b, k))) break;
            if ((rc = mg_solve(b, 0))) break;                   /* warm-up */
            if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
                rc = pv_set_error(PV_EHIP, "tile autotune: hipEventCreate failed");
            } else {
                hipEventRecord(e0, S(b));
                for (rep = 0; rep < 2 && rc == PV_OK; rep++) rc = mg_solve(b, 0);
                hipEventRecord(e1, S(b));
                if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                    rc = pv_set_error(PV_EHIP, "tile autotune: MG timing failed");
            }
            if (e0) hipEventDestroy(e0);
            if (e1) hipEventDestroy(e1);
            if (rc == PV_OK && ms < best) { best = ms; pick_s = k; }
        }
        b->p_warm = p_warm;
        b->mg_chunk = chunk;
        b->learn_start_res = start_res;
        if (saved && (hipMemcpyAsync(FD(b, PV_F_PHI) - pv_field_front(b), save, bytes, hipMemcpyDeviceToDevice,
                                     S(b)) != hipSuccess || hipStreamSynchronize(S(b)) != hipSuccess))
            if (rc == PV_OK) rc = pv_set_error(PV_EHIP, "tile autotune: restoring PHI failed");
        if (rc == PV_OK) rc = pv_set_mg_tile0(b, pick_s);
    }
    hipFree(save);
    if (rc) {                                  /* a failed tuning leaves the shapes as they were */
        pv_set_momentum_rows(b, rows0);
        if (b->mg_ok) pv_set_mg_tile0(b, sh00);
        return rc;
    }
    if (mom_rows) *mom_rows = b->mcg.band;
    if (mg_tile0) *mg_tile0 = b->mg_ok ? b->mg.sh0 : -1;
    return PV_OK;
}

int pv_autotune_mg(pv_bag *b, int *wlevels, int *sweeps)
{
    double best = 1e300, *phi_save = NULL;
    int w, nu, rep, pick_w = 0, pick_nu = 2, p_warm, chunk, rc = PV_OK, mask0, nu0;
    double start_res;
    size_t bytes;
    RC(begin(b));
    if (b->prm.pressure_solver != PV_PSOLVE_MG_PCG && b->prm.pressure_solver != PV_PSOLVE_LEARNED_MG)
        return pv_set_error(PV_EINVAL, "MG autotune is for the MG-PCG solvers");
    if (!b->mg_ok) return pv_set_error(PV_EINVAL, "MG-PCG is not available for this grid size");
    bytes = sizeof(double) * pv_field_elems(b);
    if (hipMalloc((void **)&phi_save, bytes) != hipSuccess)
        return pv_set_error(PV_ENOMEM, "MG autotune: hipMalloc of %zu bytes failed", bytes);
    if (hipMemcpyAsync(phi_save, FD(b, PV_F_PHI) - pv_field_front(b), bytes, hipMemcpyDeviceToDevice, S(b)) !=
        hipSuccess) {
        hipFree(phi_save);
        return pv_set_error(PV_EHIP, "MG autotune: saving PHI failed");
    }
    p_warm = b->p_warm;
    chunk = b->mg_chunk;
    start_res = b->learn_start_res;
    mask0 = b->mg.wmask;
    nu0 = b->mg.nu;
    b->p_warm = 0;
    for (w = 0; w <= 3 && rc == PV_OK; w++)
        for (nu = 1; nu <= 2 && rc == PV_OK; nu++) {
            hipEvent_t e0 = NULL, e1 = NULL;
            float ms = 0.0f;
            if ((rc = pv_set_mg_cycle(b, w)) || (rc = pv_set_mg_sweeps(b, nu))) break;
            if ((rc = mg_solve(b, 0))) break;                   /* warm-up; captures graphs if on */
            if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
                rc = pv_set_error(PV_EHIP, "MG autotune: hipEventCreate failed");
            } else {
                hipEventRecord(e0, S(b));
                for (rep = 0; rep < 2 && rc == PV_OK; rep++) rc = mg_solve(b, 0);
                hipEventRecord(e1, S(b));
                if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                    rc = pv_set_error(PV_EHIP, "MG autotune: event timing failed");
            }
            if (e0) hipEventDestroy(e0);
            if (e1) hipEventDestroy(e1);
            if (rc == PV_OK && ms < best) { best = ms; pick_w = w; pick_nu = nu; }
        }
    b->p_warm = p_warm;
    b->mg_chunk = chunk;
    b->learn_start_res = start_res;
    if (hipMemcpyAsync(FD(b, PV_F_PHI) - pv_field_front(b), phi_save, bytes, hipMemcpyDeviceToDevice, S(b)) !=
            hipSuccess || hipStreamSynchronize(S(b)) != hipSuccess)
        if (rc == PV_OK) rc = pv_set_error(PV_EHIP, "MG autotune: restoring PHI failed");
    hipFree(phi_save);
    if (rc) {                          /* a failed tuning leaves the cycle as it was */
        b->mg.wmask = b->mgf.wmask = mask0;
        b->mg.nu = b->mgf.nu = nu0;
        return rc;
    }
    RC(pv_set_mg_cycle(b, pick_w));
    RC(pv_set_mg_sweeps(b, pick_nu));
    if (wlevels) *wlevels = pick_w;
    if (sweeps) *sweeps = pick_nu;
    return PV_OK;
}
