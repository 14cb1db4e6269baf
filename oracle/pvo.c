/*
 * pvo.c -- CPU oracle (serial C restatement) of PARTIES' 2-D two-phase RK3
 * substep.  TEST INFRASTRUCTURE ONLY: tests/, smoke() and bench.py's
 * cpu_baseline leg use it as the checker; the product never links it.
 *
 * Build: gcc -O2 -ffp-contract=off -fPIC -shared (oracle/Makefile).  No FMA
 * contraction, so every expression rounds exactly as written; operand order
 * follows the reference expression by expression.
 *
 * Configuration restated: TWOD_CARTESIAN, PERIODIC_NOSLIP_BOX (x periodic,
 * no-slip y walls, z periodic/collapsed), GRID_UNIFORM, VOF + VOF_DIFFUSE +
 * SURFACE_TENSION + VOF_GRAVITY, CG_SOLVE (=> FULLY_IMPLICIT), USE_HYPRE off,
 * no VOF_IBM / LAG_PARTICLE_RESOLVED (C_S == 0 everywhere)
 * (Include/Boundary.h:35,88-92,230-258,366,410-420).
 *
 * Storage: one k-slab of the reference's [k][j][i] block (Memory.c:36-96),
 * canonical layout a[NYM+7][NXM+7], (i,j) -> a[(j+3)*(NXM+7) + i+3].  The
 * reference's second z slab (NZ = 2) is a halo copy of the first in every CH
 * array, so sums over the G range count it by repeating the k = 0 sequence.
 *
 * Citations are relative to PARTIES/src/.
 *
 * Pinning.  The reference holds no golden vectors for this path and cannot be
 * built here, so every exported routine is checked against what the
 * mathematics fixes (tests/test_oracle.py, tests/test_oracle_pins.py:
 * discrete eigenpairs, exactness classes, spectral solves, operator
 * identities, force balances, the published RK3 and capillary time-step
 * formulas), and tests/test_oracle_mutants.py plants 41 plausible mistakes
 * in this file, each of which one of those checks rejects.  Parity unpinned:
 * the reference's own iteration counts and dt sequence, and the effect of
 * the G-range sums of VOF_DIFFUSE.c:128-166 over a whole trajectory (the
 * code follows the reference line by line there, but only a run of the
 * reference could confirm the numbers).
 */
#include "pvo.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

/* Built with -fopenmp (libpvo_omp.so, the timed CPU baseline) the sweeps run
 * over rows on the host's cores; without it (libpvo.so, the checker) these
 * macros are empty and every sum runs in the serial order. */
#ifdef _OPENMP
#include <omp.h>
#define PVO_PRAGMA(x) _Pragma(#x)
#define PVO_FOR PVO_PRAGMA(omp parallel for schedule(static) private(i))
#define PVO_FOR_SUM(...) PVO_PRAGMA(omp parallel for schedule(static) private(i) reduction(+ : __VA_ARGS__))
#define PVO_FOR_COLS PVO_PRAGMA(omp parallel for schedule(static) private(j))
#define PVO_TID omp_get_thread_num()
#define PVO_NTHREADS omp_get_max_threads()
#else
#define PVO_FOR
#define PVO_FOR_SUM(...)
#define PVO_FOR_COLS
#define PVO_TID 0
#define PVO_NTHREADS 1
#endif

#define PI_REF 3.14159265358979323846          /* Include/definitions.h:7 */
/* Include/definitions.h:198-204 */
static const double BET[3]  = { 4.0 / 15.0, 1.0 / 15.0, 1.0 / 6.0 };
static const double BET2[3] = { 8.0 / 15.0, 2.0 / 15.0, 1.0 / 3.0 };
static const double GAMB[3] = { 2.0, 25.0 / 4.0, 9.0 / 2.0 };
static const double ZETB[3] = { 0.0, -17.0 / 4.0, -2.5 };

#define MAXD(a, b) ((a) < (b) ? (b) : (a))     /* definitions.h:44 */

/* RK3 tables as used here: out[0..2] BETA, [3..5] BETA2, [6..8] GAMBETA, [9..11] ZETBETA */
void pvo_rk3_tables(double *out)
{
    int k;
    for (k = 0; k < 3; k++) {
        out[k] = BET[k]; out[3 + k] = BET2[k]; out[6 + k] = GAMB[k]; out[9 + k] = ZETB[k];
    }
}
#define MIND(a, b) ((a) > (b) ? (b) : (a))     /* definitions.h:45 */

struct pvo_state {
    pv_params prm;
    int NXM, NYM, NX, NY, P, NYT;
    size_t nel;
    double *f[PV_NFIELDS];
    double *m_xu, *m_xc, *m_yv, *m_yc;
    double *m_dx_u, *m_dx_c, *m_idx_u, *m_idx_c;
    double *m_dy_v, *m_dy_c, *m_idy_v, *m_idy_c;
    double slab;
    double time, dt, dt_old;
    int stage, ntime;
    pv_stats st;
    double *zbuf;              /* psolve_cg.c:164 per-call z */
    double *twc, *tws;         /* cos/sin(2 pi m / NXM), m < NXM */
    struct pvo_ws *ws;         /* one FFT / Thomas workspace per host thread */
    int nws;
    double clamp_int;          /* last clamp correction over interior cells only */
    int mom_exact;             /* 1: momentum CG steps without the EPS denominators */
};

struct pvo_ws {
    double *wre, *wim, *wrow;  /* FFT workspace (length NXM) */
    double *fre, *fim;         /* radix-2 FFT work (length NXM) */
    double *tsa, *tsb, *tsc, *tsr, *tsu, *tsg; /* Thomas workspace (NYM) */
};

#define AT(a, i, j) (a)[((j) + 3) * P + (i) + 3]
#define MT(a, i)    (a)[(i) + 3]

/* ------------------------------------------------------------------ */
/* parameters and grid                                                 */
/* ------------------------------------------------------------------ */

/* IO/Input.c:20-127 derived values (2-D slab, Cn, Pe_CH) */
static void derive_params(pv_params *p, double *slab)
{
    double Lx = p->xmax - p->xmin, Ly = p->ymax - p->ymin;
    double hx = Lx / (double)p->NXM, hy = Ly / (double)p->NYM;
    double s = p->twod_slab_thickness, href = hx, hz, h_ref;
    if (hy < href) href = hy;
    if (s <= 0.0) s = href;
    if (s <= 0.0) s = 1.0;
    p->twod_slab_thickness = s;
    *slab = s;
    hz = s / 1.0;
    h_ref = hx;
    if (hy < h_ref) h_ref = hy;
    if (hz < h_ref) h_ref = hz;
    if (p->Cn <= 0.0) p->Cn = 0.75 * h_ref;
    if (p->Pe_CH <= 0.0 && p->Cn > 0.0) p->Pe_CH = 0.9 / p->Cn;
}

/* Grid.c:797-861 (uniform coordinates) and Grid.c:385-420 (metrics) */
static void build_metrics(pvo_state *s)
{
    const pv_params *p = &s->prm;
    int NX = s->NX, NY = s->NY, i, j;
    double dx = (p->xmax - p->xmin) / (double)(NX - 1);
    double dy = (p->ymax - p->ymin) / (double)(NY - 1);
    for (i = -3; i < NX + 3; i++) {
        MT(s->m_xu, i) = p->xmin + i * dx;
        MT(s->m_xc, i) = p->xmin + (0.5 + i) * dx;
    }
    for (j = -3; j < NY + 3; j++) {
        MT(s->m_yv, j) = p->ymin + j * dy;
        MT(s->m_yc, j) = p->ymin + (0.5 + j) * dy;
    }
    for (i = -3; i < NX + 2; i++) {
        MT(s->m_dx_u, i) = MT(s->m_xu, i + 1) - MT(s->m_xu, i);
        MT(s->m_dx_c, i) = MT(s->m_xc, i + 1) - MT(s->m_xc, i);
        MT(s->m_idx_u, i) = 1.0 / MT(s->m_dx_u, i);
        MT(s->m_idx_c, i) = 1.0 / MT(s->m_dx_c, i);
    }
    for (j = -3; j < NY + 2; j++) {
        MT(s->m_dy_v, j) = MT(s->m_yv, j + 1) - MT(s->m_yv, j);
        MT(s->m_dy_c, j) = MT(s->m_yc, j + 1) - MT(s->m_yc, j);
        MT(s->m_idy_v, j) = 1.0 / MT(s->m_dy_v, j);
        MT(s->m_idy_c, j) = 1.0 / MT(s->m_dy_c, j);
    }
}

pvo_state *pvo_create(const pv_params *prm)
{
    pvo_state *s = (pvo_state *)calloc(1, sizeof(pvo_state));
    int k;
    if (!s) return NULL;
    s->prm = *prm;
    derive_params(&s->prm, &s->slab);
    s->NXM = prm->NXM; s->NYM = prm->NYM;
    s->NX = s->NXM + 1; s->NY = s->NYM + 1;
    s->P = s->NXM + 7; s->NYT = s->NYM + 7;
    s->nel = (size_t)s->P * (size_t)s->NYT;
    for (k = 0; k < PV_NFIELDS; k++) s->f[k] = (double *)calloc(s->nel, sizeof(double));
#define MA(x, n) s->x = (double *)calloc((size_t)(n), sizeof(double))
    MA(m_xu, s->P); MA(m_xc, s->P); MA(m_dx_u, s->P); MA(m_dx_c, s->P);
    MA(m_idx_u, s->P); MA(m_idx_c, s->P);
    MA(m_yv, s->NYT); MA(m_yc, s->NYT); MA(m_dy_v, s->NYT); MA(m_dy_c, s->NYT);
    MA(m_idy_v, s->NYT); MA(m_idy_c, s->NYT);
    MA(zbuf, s->nel); MA(twc, s->NXM); MA(tws, s->NXM);
#undef MA
    s->nws = PVO_NTHREADS;
    s->ws = (struct pvo_ws *)calloc((size_t)s->nws, sizeof(struct pvo_ws));
    for (k = 0; k < s->nws; k++) {
        struct pvo_ws *w = &s->ws[k];
#define MW(x, n) w->x = (double *)calloc((size_t)(n), sizeof(double))
        MW(wre, s->NXM); MW(wim, s->NXM); MW(wrow, s->NXM); MW(fre, s->NXM); MW(fim, s->NXM);
        MW(tsa, s->NYM); MW(tsb, s->NYM); MW(tsc, s->NYM); MW(tsr, s->NYM); MW(tsu, s->NYM); MW(tsg, s->NYM);
#undef MW
    }
    build_metrics(s);
    for (k = 0; k < s->NXM; k++) {
        double ang = 2.0 * PI_REF * (double)k / (double)s->NXM;
        s->twc[k] = cos(ang); s->tws[k] = sin(ang);
    }
    s->dt = s->dt_old = s->prm.default_dt;
    return s;
}

void pvo_destroy(pvo_state *s)
{
    int k;
    if (!s) return;
    for (k = 0; k < PV_NFIELDS; k++) free(s->f[k]);
    free(s->m_xu); free(s->m_xc); free(s->m_dx_u); free(s->m_dx_c);
    free(s->m_idx_u); free(s->m_idx_c); free(s->m_yv); free(s->m_yc);
    free(s->m_dy_v); free(s->m_dy_c); free(s->m_idy_v); free(s->m_idy_c);
    free(s->zbuf); free(s->twc); free(s->tws);
    for (k = 0; k < s->nws; k++) {
        struct pvo_ws *w = &s->ws[k];
        free(w->wre); free(w->wim); free(w->wrow); free(w->fre); free(w->fim);
        free(w->tsa); free(w->tsb); free(w->tsc); free(w->tsr); free(w->tsu); free(w->tsg);
    }
    free(s->ws);
    free(s);
}

void pvo_get_params(const pvo_state *s, pv_params *out) { *out = s->prm; }

/* Host threads of the OpenMP build (libpvo_omp.so); the serial build has one. */
int pvo_threads(void) { return PVO_NTHREADS; }
double *pvo_field(pvo_state *s, int field) { return (field >= 0 && field < PV_NFIELDS) ? s->f[field] : NULL; }
int pvo_nxt(const pvo_state *s) { return s->P; }
int pvo_nyt(const pvo_state *s) { return s->NYT; }

/* The momentum CG's steps without EPS (alpha = rz / dAd, beta = rz_new / rz,
 * zero when the denominator is zero): the library's pv_set_momentum_cg_steps
 * variant, not a reference routine. */
void pvo_set_momentum_cg_steps(pvo_state *s, int exact)
{
    s->mom_exact = exact;
}

void pvo_set_time_state(pvo_state *s, double time, double dt, double dt_old, int stage)
{
    s->time = time; s->dt = dt; s->dt_old = dt_old; s->stage = stage;
}

void pvo_get_stats(const pvo_state *s, pv_stats *out)
{
    *out = s->st;
    out->time = s->time; out->dt = s->dt; out->dt_old = s->dt_old;
    out->ntime = s->ntime; out->which_stage = s->stage;
}

/* ------------------------------------------------------------------ */
/* ghost fills                                                         */
/* ------------------------------------------------------------------ */

/* Communication.c:146-295, cell-type component with XPERIODIC: one rank is
 * its own x neighbour; period NXM over cells 0..NXM-1, 3 layers. */
static void halo_cell(pvo_state *s, double *f)
{
    int P = s->P, NXM = s->NXM, j, g;
    for (j = -3; j < s->NY + 3; j++) {
        for (g = 0; g < 3; g++) AT(f, -1 - g, j) = AT(f, NXM - 1 - g, j);
        for (g = 0; g < 3; g++) AT(f, NXM + g, j) = AT(f, g, j);
    }
}

/* Communication.c:162-169: 'u', U_VELOCITY and FLUX_X receive one extra
 * layer; unknowns 1..NXM, f[0] is the image of f[NXM]. */
static void halo_u(pvo_state *s, double *f)
{
    int P = s->P, NXM = s->NXM, j, g;
    for (j = -3; j < s->NY + 3; j++) {
        for (g = 0; g < 4; g++) AT(f, -g, j) = AT(f, NXM - g, j);
        for (g = 0; g < 3; g++) AT(f, NXM + 1 + g, j) = AT(f, 1 + g, j);
    }
}

/* Communication.c:128-295 on one rank that is its own x neighbour, any
 * pnodes: rows -pnodes .. NY-1+pnodes; 'u'-type components ('u', 'd', 'e',
 * 'p') start at Is = 1 and receive pnodes + 1 layers (:162-169); the other
 * components end at Ie = NX - 1.  No y neighbour; z is the k = 0 slab. */
void pvo_Communication_update_ghost_nodes_flow_variable(pvo_state *s, int field, int component, int pnodes)
{
    double *f = s->f[field];
    const int P = s->P, NX = s->NX;
    const int ulike = (component == 'u' || component == 'd' || component == 'e' || component == 'p');
    const int Is = ulike ? 1 : 0, Ie = ulike ? NX : NX - 1;
    const int precv = ulike ? pnodes + 1 : pnodes, psend = precv;
    int i, j;
    (void)psend;
    for (j = -pnodes; j < s->NY + pnodes; j++) {
        for (i = 0; i < precv; i++) AT(f, Is - i - 1, j) = AT(f, Ie - i - 1, j);   /* from the upper side */
        for (i = 0; i < pnodes; i++) AT(f, Ie + i, j) = AT(f, Is + i, j);           /* from the lower side */
    }
}

/* VOF_DIFFUSE.c:571-683: zero-gradient y walls over the L range in i, all
 * three ghost layers, then the cell-type exchange. */
static void bc_ch(pvo_state *s, double *f)
{
    int P = s->P, NY = s->NY, i, g;
    for (i = -3; i < s->NX + 3; i++) {
        for (g = 1; g <= 3; g++) AT(f, i, -g) = AT(f, i, 0);
        for (g = 0; g < 3; g++) AT(f, i, NY - 1 + g) = AT(f, i, NY - 2);
    }
    halo_cell(s, f);
}

void pvo_VOF_DIFFUSE_set_boundary_values(pvo_state *s, int field) { bc_ch(s, s->f[field]); }

/* lsolver/psolve_cg.c:358-529, y part (Neumann copy of one layer). */
static void p_bc(pvo_state *s, double *f)
{
    int P = s->P, NY = s->NY, i;
    for (i = -3; i < s->NX + 3; i++) {
        AT(f, i, -1) = AT(f, i, 0);
        AT(f, i, NY - 1) = AT(f, i, NY - 2);
    }
}

void pvo_Pressure_apply_BCs(pvo_state *s, int field)
{
    p_bc(s, s->f[field]);
    halo_cell(s, s->f[field]);
}

/* Velocity.c:2427-2805 for PERIODIC_NOSLIP_BOX. */
static void vel_bc(pvo_state *s, double *f, int comp, int type)
{
    int P = s->P, NY = s->NY, i;
    if (comp == 'u') {
        for (i = 0; i < s->NX; i++) {           /* G range, :2600-2606, :2676-2682 */
            AT(f, i, -1) = -AT(f, i, 0);
            AT(f, i, NY - 1) = -AT(f, i, NY - 2);
        }
        halo_u(s, f);
    } else {
        if (type == PV_VEL_TYPE_NORMAL) {       /* :2566-2574, :2636-2644 */
            for (i = -3; i < s->NX + 3; i++) {
                AT(f, i, 0) = 0.0;
                AT(f, i, NY - 1) = 0.0;
            }
        }
        halo_cell(s, f);
    }
}

static int comp_of_field(int field)
{
    switch (field) {
    case PV_F_U: case PV_F_U_D: return 'u';
    default: return 'v';
    }
}

void pvo_Velocity_update_boundaries(pvo_state *s, int field, int type)
{
    vel_bc(s, s->f[field], comp_of_field(field), type);
}

/* ------------------------------------------------------------------ */
/* Cahn-Hilliard                                                       */
/* ------------------------------------------------------------------ */

/* TwodOps.c:159-199, planar branch */
static inline double lap_at(const pvo_state *s, const double *f, int i, int j)
{
    int P = s->P;
    double dfdxE = (AT(f, i + 1, j) - AT(f, i, j)) * MT(s->m_idx_c, i);
    double dfdxW = (AT(f, i, j) - AT(f, i - 1, j)) *
                   ((i != 0) ? MT(s->m_idx_c, i - 1) : MT(s->m_idx_c, i));
    double dfdyN = (AT(f, i, j + 1) - AT(f, i, j)) * MT(s->m_idy_c, j);
    double dfdyS = (AT(f, i, j) - AT(f, i, j - 1)) *
                   ((j != 0) ? MT(s->m_idy_c, j - 1) : MT(s->m_idy_c, j));
    return (dfdxE - dfdxW) * MT(s->m_idx_u, i) + (dfdyN - dfdyS) * MT(s->m_idy_v, j);
}

/* VOF_DIFFUSE.c:838-863 */
static void compute_laplacian(pvo_state *s, double *in, double *lap)
{
    int P = s->P, i, j;
    bc_ch(s, in);
    PVO_FOR
    for (j = 0; j < s->NY; j++)
        for (i = 0; i < s->NX; i++)
            AT(lap, i, j) = lap_at(s, in, i, j);
    bc_ch(s, lap);
}

void pvo_VOF_DIFFUSE_compute_laplacian(pvo_state *s, int in, int out)
{
    compute_laplacian(s, s->f[in], s->f[out]);
}

/* VOF_DIFFUSE.c:865-885 (C_S = 0) */
void pvo_VOF_DIFFUSE_compute_bulk_S(pvo_state *s)
{
    int P = s->P, i, j;
    double *C = s->f[PV_F_C_L], *S = s->f[PV_F_BULK_S];
    const double cs = 0.0;
    PVO_FOR
    for (j = 0; j < s->NY; j++)
        for (i = 0; i < s->NX; i++) {
            double c = AT(C, i, j);
            AT(S, i, j) = 0.5 * c - 1.5 * c * c + c * c * c - c * cs * (1.0 - c - cs);
        }
    bc_ch(s, S);
}

/* VOF_DIFFUSE.c:907-929 */
void pvo_VOF_DIFFUSE_compute_psi_LG(pvo_state *s)
{
    int P = s->P, i, j;
    double *C = s->f[PV_F_C_L], *L = s->f[PV_F_LAP_C], *psi = s->f[PV_F_PSI_LG];
    double Cn = s->prm.Cn;
    compute_laplacian(s, C, L);
    PVO_FOR
    for (j = 0; j < s->NY; j++)
        for (i = 0; i < s->NX; i++) {
            double c = AT(C, i, j);
            AT(psi, i, j) = c * c * c - 1.5 * c * c + 0.5 * c - Cn * Cn * AT(L, i, j);
        }
    bc_ch(s, psi);
}

/* VOF_DIFFUSE.c:431-453 (pow kept as in the reference) */
static inline double weno5_left(double v0, double v1, double v2, double v3, double v4)
{
    const double eps = 1e-6;
    double q0 = (2.0 * v0 - 7.0 * v1 + 11.0 * v2) / 6.0;
    double q1 = (-v1 + 5.0 * v2 + 2.0 * v3) / 6.0;
    double q2 = (2.0 * v2 + 5.0 * v3 - v4) / 6.0;
    double b0 = 13.0 / 12.0 * pow(v0 - 2.0 * v1 + v2, 2.0) + 0.25 * pow(v0 - 4.0 * v1 + 3.0 * v2, 2.0);
    double b1 = 13.0 / 12.0 * pow(v1 - 2.0 * v2 + v3, 2.0) + 0.25 * pow(v1 - v3, 2.0);
    double b2 = 13.0 / 12.0 * pow(v2 - 2.0 * v3 + v4, 2.0) + 0.25 * pow(3.0 * v2 - 4.0 * v3 + v4, 2.0);
    double a0 = 0.1 / ((eps + b0) * (eps + b0));
    double a1 = 0.6 / ((eps + b1) * (eps + b1));
    double a2 = 0.3 / ((eps + b2) * (eps + b2));
    double asum = a0 + a1 + a2;
    return (a0 * q0 + a1 * q1 + a2 * q2) / asum;
}

/* VOF_DIFFUSE.c:931-996 with TwodOps.c:201-229 (planar) */
void pvo_VOF_DIFFUSE_advect_WENO5(pvo_state *s, int rhs_out)
{
    int P = s->P, i, j, wo = s->prm.weno_order;
    double *C = s->f[PV_F_C_L], *u = s->f[PV_F_U], *v = s->f[PV_F_V];
    double *fx = s->f[PV_F_FLUX_X], *fy = s->f[PV_F_FLUX_Y], *rhs = s->f[rhs_out];
    bc_ch(s, C);
    memset(fx, 0, s->nel * sizeof(double));
    memset(fy, 0, s->nel * sizeof(double));
    PVO_FOR
    for (j = 0; j < s->NY; j++)
        for (i = 0; i < s->NX; i++) {
            double vel = AT(u, i, j), st;
            if (wo == 5)
                st = (vel >= 0.0)
                    ? weno5_left(AT(C, i - 3, j), AT(C, i - 2, j), AT(C, i - 1, j), AT(C, i, j), AT(C, i + 1, j))
                    : weno5_left(AT(C, i + 2, j), AT(C, i + 1, j), AT(C, i, j), AT(C, i - 1, j), AT(C, i - 2, j));
            else
                st = (vel >= 0.0) ? AT(C, i - 1, j) : AT(C, i, j);
            AT(fx, i, j) = vel * st;
            vel = AT(v, i, j);
            if (wo == 5)
                st = (vel >= 0.0)
                    ? weno5_left(AT(C, i, j - 3), AT(C, i, j - 2), AT(C, i, j - 1), AT(C, i, j), AT(C, i, j + 1))
                    : weno5_left(AT(C, i, j + 2), AT(C, i, j + 1), AT(C, i, j), AT(C, i, j - 1), AT(C, i, j - 2));
            else
                st = (vel >= 0.0) ? AT(C, i, j - 1) : AT(C, i, j);
            AT(fy, i, j) = vel * st;
        }
    halo_u(s, fx);        /* FLUX_X: u-type exchange */
    halo_cell(s, fy);     /* FLUX_Y: cell-type in x */
    PVO_FOR
    for (j = 0; j < s->NY; j++)
        for (i = 0; i < s->NX; i++) {
            double div_x = (AT(fx, i + 1, j) - AT(fx, i, j)) / MT(s->m_dx_c, i);
            double div_y = (AT(fy, i, j + 1) - AT(fy, i, j)) / MT(s->m_dy_c, j);
            double div = div_x + div_y;
            AT(rhs, i, j) = -div;
        }
    bc_ch(s, rhs);
}

/* VOF_DIFFUSE.c:998-1016 */
void pvo_VOF_DIFFUSE_explicit_diffusion(pvo_state *s, int rhs_out)
{
    int P = s->P, i, j;
    double *aux = s->f[PV_F_CH_AUX1], *rhs = s->f[rhs_out];
    pvo_VOF_DIFFUSE_compute_bulk_S(s);
    compute_laplacian(s, s->f[PV_F_BULK_S], aux);
    PVO_FOR
    for (j = 0; j < s->NY; j++)
        for (i = 0; i < s->NX; i++)
            AT(rhs, i, j) = AT(aux, i, j) / s->prm.Pe_CH;
    bc_ch(s, rhs);
}

/* VOF_DIFFUSE.c:128-146 with TwodOps.c:323-345: measure-weighted sum over
 * the G range (k = 0, 1; i = 0..NX-1; j = 0..NY-1). */
static double ch_inner(const pvo_state *s, const double *a, const double *b)
{
    int P = s->P, i, j, k;
    double sum = 0.0;
    for (k = 0; k < 2; k++) {
        PVO_FOR_SUM(sum)
        for (j = 0; j < s->NY; j++)
            for (i = 0; i < s->NX; i++) {
                double measure = MT(s->m_dx_u, i) * MT(s->m_dy_v, j) * s->slab;
                sum += measure * AT(a, i, j) * AT(b, i, j);
            }
    }
    return sum;
}

/* VOF_DIFFUSE.c:318-337 */
static void apply_biharmonic(pvo_state *s, double *x, double *Ax, double a_dt, double bih)
{
    int P = s->P, i, j;
    double *lap_x = s->f[PV_F_CH_AUX1], *bih_x = s->f[PV_F_PSI_LG];
    compute_laplacian(s, x, lap_x);
    compute_laplacian(s, lap_x, bih_x);
    PVO_FOR
    for (j = 0; j < s->NY; j++)
        for (i = 0; i < s->NX; i++)
            AT(Ax, i, j) = a_dt * AT(x, i, j) + bih * AT(bih_x, i, j);
}

/* VOF_DIFFUSE.c:339-429: unpreconditioned CG (z = r) on [a_dt I + c L^2]. */
static int biharmonic_cg(pvo_state *s, double *x, double *b, double a_dt, double bih, double *final_rel)
{
    int P = s->P, i, j, iter = 0;
    double *r = s->f[PV_F_CH_RES], *p = s->f[PV_F_CH_DIR];
    double *z = s->f[PV_F_BULK_S], *Ap = s->f[PV_F_LAP_C];
    int maxit = (s->prm.ch_iter_max > 0) ? s->prm.ch_iter_max : s->prm.CG_MAXIT;
    double tol = (s->prm.ch_tol > 0.0) ? s->prm.ch_tol : s->prm.CG_ETOL;
    double rhs_norm, rz, rel;

    apply_biharmonic(s, x, Ap, a_dt, bih);
    PVO_FOR
    for (j = 0; j < s->NY; j++)
        for (i = 0; i < s->NX; i++) {
            AT(r, i, j) = AT(b, i, j) - AT(Ap, i, j);
            AT(z, i, j) = AT(r, i, j);
            AT(p, i, j) = AT(z, i, j);
        }
    rhs_norm = sqrt(ch_inner(s, b, b));
    if (rhs_norm < 1e-30) rhs_norm = 1.0;
    rz = ch_inner(s, r, z);
    rel = sqrt(ch_inner(s, r, r)) / rhs_norm;
    if (rel < tol) {
        *final_rel = rel;
        bc_ch(s, x);
        return 0;
    }
    while (iter < maxit) {
        double pAp, alpha, rz_new, beta;
        apply_biharmonic(s, p, Ap, a_dt, bih);
        pAp = ch_inner(s, p, Ap);
        if (fabs(pAp) < 1e-30) break;
        alpha = rz / pAp;
        PVO_FOR
        for (j = 0; j < s->NY; j++)
            for (i = 0; i < s->NX; i++) {
                AT(x, i, j) += alpha * AT(p, i, j);
                AT(r, i, j) -= alpha * AT(Ap, i, j);
                AT(z, i, j) = AT(r, i, j);
            }
        ++iter;
        rel = sqrt(ch_inner(s, r, r)) / rhs_norm;
        if (rel < tol) break;
        rz_new = ch_inner(s, r, z);
        beta = rz_new / (rz + 1e-30);
        PVO_FOR
        for (j = 0; j < s->NY; j++)
            for (i = 0; i < s->NX; i++)
                AT(p, i, j) = AT(z, i, j) + beta * AT(p, i, j);
        rz = rz_new;
    }
    *final_rel = rel;
    bc_ch(s, x);
    return iter;
}

/* VOF_DIFFUSE.c:1018-1069 */
int pvo_VOF_DIFFUSE_solve_implicit_biharmonic(pvo_state *s, int rhs, double *rel_res)
{
    int P = s->P, i, j, it;
    double a_dt = 1.0 / (s->dt * BET[s->stage]);
    double bih = s->prm.Cn * s->prm.Cn / s->prm.Pe_CH;
    double *b = s->f[rhs], *psi = s->f[PV_F_PSI], *psiLG = s->f[PV_F_PSI_LG];
    double rr = 0.0;
    memcpy(psi, s->f[PV_F_C_L], s->nel * sizeof(double));
    compute_laplacian(s, psi, s->f[PV_F_CH_AUX1]);
    compute_laplacian(s, s->f[PV_F_CH_AUX1], psiLG);
    PVO_FOR
    for (j = 0; j < s->NY; j++)
        for (i = 0; i < s->NX; i++)
            AT(b, i, j) -= bih * AT(psiLG, i, j);
    it = biharmonic_cg(s, s->f[PV_F_C_L], b, a_dt, bih, &rr);
    if (rel_res) *rel_res = rr;
    s->st.n_ch[s->stage] = it;
    s->st.ch_res[s->stage] = rr;
    return it;
}

/* VOF_DIFFUSE.c:1093-1127 with diffuse_physical_cell_measure (:148-160).
 * The k = 1 slab holds the same pre-clamp values as k = 0, so its
 * contributions repeat the k = 0 sequence. */
double pvo_VOF_DIFFUSE_bound_liquid_fraction(pvo_state *s)
{
    int P = s->P, i, j, k;
    double *C = s->f[PV_F_C_L], *G = s->f[PV_F_C_G];
    double delta = 0.0, delta_int = 0.0;
    const double cs = 0.0;
    for (k = 0; k < 2; k++) {
        PVO_FOR_SUM(delta, delta_int)
        for (j = 0; j < s->NY; j++)
            for (i = 0; i < s->NX; i++) {
                double old_cl = AT(C, i, j);
                double max_cl = 1.0 - cs, new_cl;
                if (max_cl < 0.0) max_cl = 0.0;
                if (max_cl > 1.0) max_cl = 1.0;
                new_cl = old_cl;
                if (new_cl < 0.0) new_cl = 0.0;
                if (new_cl > max_cl) new_cl = max_cl;
                if (cs < 0.05) {
                    double measure = MT(s->m_dx_u, i) * MT(s->m_dy_v, j) * s->slab;
                    measure /= 2.0;
                    delta += (new_cl - old_cl) * measure;
                    if (i < s->NXM && j < s->NYM) delta_int += (new_cl - old_cl) * measure;
                }
                if (k == 1) {
                    double g = 1.0 - new_cl - cs;
                    AT(C, i, j) = new_cl;
                    AT(G, i, j) = (g < 0.0) ? 0.0 : ((g > 1.0) ? 1.0 : g);
                }
            }
    }
    s->clamp_int = delta_int;
    return delta;
}

/* The last clamp's correction over interior cells only (i < NXM, j < NYM):
 * the reference's sum above also counts the periodic image column i = NXM
 * and the ghost row j = NYM, which hold copies of column 0 and row NYM-1. */
double pvo_clamp_interior(const pvo_state *s) { return s->clamp_int; }

/* VOF_DIFFUSE.c:108-126 (the F mirror is not kept) */
void pvo_VOF_DIFFUSE_update_phase_cache(pvo_state *s)
{
    double *C = s->f[PV_F_C_L], *G = s->f[PV_F_C_G];
    size_t n;
    for (n = 0; n < s->nel; n++) {
        double g = 1.0 - C[n] - 0.0;
        G[n] = (g < 0.0) ? 0.0 : ((g > 1.0) ? 1.0 : g);
    }
}

/* VOF_DIFFUSE.c:1262-1359 (VOF_DIFFUSE_RESTORE_MASS_EVERY_STAGE = 0, no IBM) */
void pvo_VOF_DIFFUSE_step(pvo_state *s)
{
    int P = s->P, i, j, st = s->stage;
    double a_dt = 1.0 / (s->dt * BET[st]);
    double *rc = s->f[PV_F_CH_AUX2], *rt = s->f[PV_F_CH_AUX1], *rv = s->f[PV_F_CH_RHS_NM1];
    double *C = s->f[PV_F_C_L], *rn = s->f[PV_F_CH_RHS_N];

    pvo_VOF_DIFFUSE_advect_WENO5(s, PV_F_CH_AUX2);
    pvo_VOF_DIFFUSE_explicit_diffusion(s, PV_F_CH_AUX1);
    PVO_FOR
    for (j = 0; j < s->NY; j++)
        for (i = 0; i < s->NX; i++)
            AT(rc, i, j) += AT(rt, i, j);
    bc_ch(s, rc);
    if (st == 0) {
        PVO_FOR
        for (j = 0; j < s->NY; j++)
            for (i = 0; i < s->NX; i++)
                AT(rv, i, j) = a_dt * AT(C, i, j) + GAMB[0] * AT(rc, i, j);
    } else {
        PVO_FOR
        for (j = 0; j < s->NY; j++)
            for (i = 0; i < s->NX; i++)
                AT(rv, i, j) = a_dt * AT(C, i, j) + GAMB[st] * AT(rc, i, j) + ZETB[st] * AT(rn, i, j);
    }
    bc_ch(s, rv);
    pvo_VOF_DIFFUSE_solve_implicit_biharmonic(s, PV_F_CH_RHS_NM1, NULL);
    s->st.clamp_delta[st] = pvo_VOF_DIFFUSE_bound_liquid_fraction(s);
    s->st.clamp_delta_interior[st] = s->clamp_int;
    bc_ch(s, C);
    pvo_VOF_DIFFUSE_update_phase_cache(s);
    memcpy(s->f[PV_F_CH_RHS_NM1], rn, s->nel * sizeof(double));
    memcpy(rn, rc, s->nel * sizeof(double));
}

/* VOF_DIFFUSE.c:1453-1488 (C_S = 0, lambda_S = 1) */
void pvo_VOF_update_density_viscosity(pvo_state *s)
{
    int P = s->P, i, j;
    double *C = s->f[PV_F_C_L], *G = s->f[PV_F_C_G], *rho = s->f[PV_F_RHO], *mu = s->f[PV_F_MU];
    const double rho_g = s->prm.rho2 / s->prm.rho1, mu_g = s->prm.mu2 / s->prm.mu1;
    const double rho_s = 1.0, mu_s = 1.0, cs = 0.0;
    PVO_FOR
    for (j = 0; j < s->NY; j++)
        for (i = 0; i < s->NX; i++) {
            AT(rho, i, j) = AT(C, i, j) + rho_s * cs + rho_g * AT(G, i, j);
            AT(mu, i, j) = AT(C, i, j) + mu_s * cs + mu_g * AT(G, i, j);
        }
    bc_ch(s, rho);
    bc_ch(s, mu);
}

/* VOF_DIFFUSE.c:1363-1445 */
void pvo_VOF_DIFFUSE_compute_f_sigma(pvo_state *s)
{
    int P = s->P, i, j;
    double *C = s->f[PV_F_C_L], *G = s->f[PV_F_C_G], *psi = s->f[PV_F_PSI_LG];
    double *fx = s->f[PV_F_FSX_NEW], *fy = s->f[PV_F_FSY_NEW];
    const double scale = 2.0 * 6.0 * sqrt(2.0) / (s->prm.Cn * s->prm.We + 1e-30);
    const double cs = 0.0;
    memset(fx, 0, s->nel * sizeof(double));
    memset(fy, 0, s->nel * sizeof(double));
    PVO_FOR
    for (j = 0; j < s->NY - 1; j++)
        for (i = 1; i < s->NX; i++) {
            double cl = 0.5 * (AT(C, i, j) + AT(C, i - 1, j));
            double cg = 0.5 * (AT(G, i, j) + AT(G, i - 1, j));
            double psi_f, gradC;
            if (cl < 0.005 || cl > 0.995 || cg < 0.005 || cg > 0.995 || cs > 0.05) continue;
            psi_f = 0.5 * (AT(psi, i, j) + AT(psi, i - 1, j));
            gradC = (AT(C, i, j) - AT(C, i - 1, j)) * MT(s->m_idx_c, i - 1);
            AT(fx, i, j) = scale * psi_f * gradC;
        }
    PVO_FOR
    for (j = 1; j < s->NY - 1; j++)
        for (i = 0; i < s->NX - 1; i++) {
            double cl = 0.5 * (AT(C, i, j) + AT(C, i, j - 1));
            double cg = 0.5 * (AT(G, i, j) + AT(G, i, j - 1));
            double psi_f, gradC;
            if (cl < 0.005 || cl > 0.995 || cg < 0.005 || cg > 0.995 || cs > 0.05) continue;
            psi_f = 0.5 * (AT(psi, i, j) + AT(psi, i, j - 1));
            gradC = (AT(C, i, j) - AT(C, i, j - 1)) * MT(s->m_idy_c, j - 1);
            AT(fy, i, j) = scale * psi_f * gradC;
        }
}

/* VOF_SurfaceTension.c:1983-2041 */
void pvo_VOF_apply_f_sigma_old(pvo_state *s)
{
    int P = s->P, i, j;
    double *ru = s->f[PV_F_U_RHS], *rv = s->f[PV_F_V_RHS];
    double *fx = s->f[PV_F_FSX_OLD], *fy = s->f[PV_F_FSY_OLD];
    PVO_FOR
    for (j = 0; j < s->NY - 1; j++)
        for (i = 1; i < s->NX; i++) AT(ru, i, j) += AT(fx, i, j);
    PVO_FOR
    for (j = 1; j < s->NY - 1; j++)
        for (i = 0; i < s->NX - 1; i++) AT(rv, i, j) += AT(fy, i, j);
}

/* VOF_SurfaceTension.c:2055-2071 */
void pvo_VOF_swap_f_sigma(pvo_state *s)
{
    double *t = s->f[PV_F_FSX_OLD];
    s->f[PV_F_FSX_OLD] = s->f[PV_F_FSX_NEW]; s->f[PV_F_FSX_NEW] = t;
    t = s->f[PV_F_FSY_OLD];
    s->f[PV_F_FSY_OLD] = s->f[PV_F_FSY_NEW]; s->f[PV_F_FSY_NEW] = t;
}

/* VOF_SurfaceTension.c:2075-2154 (twod_mode_enabled) */
void pvo_Velocity_add_gravity_2_RHS(pvo_state *s)
{
    int P = s->P, i, j;
    double *rho = s->f[PV_F_RHO], *ru = s->f[PV_F_U_RHS], *rv = s->f[PV_F_V_RHS];
    const double *grav = s->prm.grav;
    double Ri = s->prm.richardson, gm, gx = 0, gy = 0;
    gm = sqrt(grav[0] * grav[0] + grav[1] * grav[1]);
    if (gm > 0.0) { gx = grav[0] / gm; gy = grav[1] / gm; }
    PVO_FOR
    for (j = 0; j < s->NY - 1; j++)
        for (i = 1; i < s->NX; i++) {
            double rho_f = 0.5 * (AT(rho, i, j) + AT(rho, i - 1, j));
            AT(ru, i, j) += 2.0 * rho_f * Ri * gx;
        }
    PVO_FOR
    for (j = 1; j < s->NY - 1; j++)
        for (i = 0; i < s->NX - 1; i++) {
            double rho_f = 0.5 * (AT(rho, i, j) + AT(rho, i, j - 1));
            AT(rv, i, j) += 2.0 * rho_f * Ri * gy;
        }
}

/* ------------------------------------------------------------------ */
/* momentum                                                            */
/* ------------------------------------------------------------------ */

/* Velocity.c:278-357 (over the L range minus one) */
void pvo_Velocity_cell_center(pvo_state *s)
{
    int P = s->P, i, j;
    double *u = s->f[PV_F_U], *v = s->f[PV_F_V], *ub = s->f[PV_F_U_BC], *vb = s->f[PV_F_V_BC];
    halo_u(s, u);
    halo_cell(s, v);
    PVO_FOR
    for (j = -3; j < s->NY + 2; j++)
        for (i = -3; i < s->NX + 2; i++) {
            AT(ub, i, j) = 0.5 * (AT(u, i, j) + AT(u, i + 1, j));
            AT(vb, i, j) = 0.5 * (AT(v, i, j) + AT(v, i, j + 1));
        }
}

#define HARM(a, b) (2.0 * (a) * (b) / ((a) + (b) + 1e-12))   /* msolve_cg.c:73 */
#define AVG2(a, b) (0.5 * ((a) + (b)))                        /* msolve_cg.c:74 */

/* Velocity.c:366-762 (VOF, VOF_DIFFUSE velocity-form convection, planar) */
void pvo_Velocity_u_set_implicit_explicit(pvo_state *s)
{
    int P = s->P, i, j;
    double *u = s->f[PV_F_U], *v = s->f[PV_F_V], *ub = s->f[PV_F_U_BC], *mu = s->f[PV_F_MU];
    double *ex = s->f[PV_F_U_EXPL], *im = s->f[PV_F_U_IMPL], *ve = s->f[PV_F_U_VISC];
    double iRe = 1.0 / s->prm.Re;
    PVO_FOR
    for (j = 0; j < s->NY - 1; j++)
        for (i = 1; i < s->NX; i++) {
            double muE = AT(mu, i, j), muW = AT(mu, i - 1, j);
            double mu_yp = 2.0 * (0.5 * (AT(mu, i - 1, j + 1) + AT(mu, i, j + 1))) *
                           (0.5 * (AT(mu, i - 1, j) + AT(mu, i, j))) /
                           ((0.5 * (AT(mu, i - 1, j + 1) + AT(mu, i, j + 1))) +
                            (0.5 * (AT(mu, i - 1, j) + AT(mu, i, j))) + 1e-12);
            double mu_ym = 2.0 * (0.5 * (AT(mu, i - 1, j) + AT(mu, i, j))) *
                           (0.5 * (AT(mu, i - 1, j - 1) + AT(mu, i, j - 1))) /
                           ((0.5 * (AT(mu, i - 1, j) + AT(mu, i, j))) +
                            (0.5 * (AT(mu, i - 1, j - 1) + AT(mu, i, j - 1))) + 1e-12);
            double dudxE = (AT(u, i + 1, j) - AT(u, i, j)) * MT(s->m_idx_u, i);
            double dudxW = (AT(u, i, j) - AT(u, i - 1, j)) * MT(s->m_idx_u, i - 1);
            double dudyN = (AT(u, i, j + 1) - AT(u, i, j)) * MT(s->m_idy_c, j);
            double dudyS = (AT(u, i, j) - AT(u, i, j - 1)) * MT(s->m_idy_c, j - 1);
            double d2udy2 = iRe * (mu_yp * dudyN - mu_ym * dudyS) * MT(s->m_idy_v, j);
            double d2udz2 = 0.0;
            double dvdxN = (AT(v, i, j + 1) - AT(v, i - 1, j + 1)) * MT(s->m_idx_c, i - 1);
            double dvdxS = (AT(v, i, j) - AT(v, i - 1, j)) * MT(s->m_idx_c, i - 1);
            double d2udx2 = iRe * ((muE * dudxE - muW * dudxW) * MT(s->m_idx_c, i - 1));
            double ddxdudx = d2udx2;
            double ddydvdx = iRe * (mu_yp * dvdxN - mu_ym * dvdxS) * MT(s->m_idy_v, j);
            double ddzdwdx = 0.0, axisym = 0.0, duwdz = 0.0;
            double vCN = 0.5 * (AT(v, i, j + 1) + AT(v, i - 1, j + 1));
            double vCS = 0.5 * (AT(v, i, j) + AT(v, i - 1, j));
            double uuE = AT(ub, i, j) * AT(ub, i, j);
            double uuW = AT(ub, i - 1, j) * AT(ub, i - 1, j);
            double uvN = 0.5 * (AT(u, i, j) + AT(u, i, j + 1)) * vCN;
            double uvS = 0.5 * (AT(u, i, j) + AT(u, i, j - 1)) * vCS;
            double duudx = (uuE - uuW) * MT(s->m_idx_c, i - 1);
            double duvdy = (uvN - uvS) * MT(s->m_idy_v, j);
            AT(ex, i, j) = -(duudx + duvdy + duwdz);
            AT(ve, i, j) = ddxdudx + ddydvdx + ddzdwdx;
            AT(im, i, j) = d2udx2 + d2udy2 + d2udz2 + axisym;
        }
}

/* Velocity.c:771-1144 */
void pvo_Velocity_v_set_implicit_explicit(pvo_state *s)
{
    int P = s->P, i, j;
    double *u = s->f[PV_F_U], *v = s->f[PV_F_V], *vb = s->f[PV_F_V_BC], *mu = s->f[PV_F_MU];
    double *ex = s->f[PV_F_V_EXPL], *im = s->f[PV_F_V_IMPL], *ve = s->f[PV_F_V_VISC];
    double iRe = 1.0 / s->prm.Re;
    PVO_FOR
    for (j = 1; j < s->NY - 1; j++)
        for (i = 0; i < s->NX - 1; i++) {
            double mu_xp = 2.0 * (0.5 * (AT(mu, i + 1, j - 1) + AT(mu, i + 1, j))) *
                           (0.5 * (AT(mu, i, j - 1) + AT(mu, i, j))) /
                           ((0.5 * (AT(mu, i + 1, j - 1) + AT(mu, i + 1, j))) +
                            (0.5 * (AT(mu, i, j - 1) + AT(mu, i, j))) + 1e-12);
            double mu_xm = 2.0 * (0.5 * (AT(mu, i, j - 1) + AT(mu, i, j))) *
                           (0.5 * (AT(mu, i - 1, j - 1) + AT(mu, i - 1, j))) /
                           ((0.5 * (AT(mu, i, j - 1) + AT(mu, i, j))) +
                            (0.5 * (AT(mu, i - 1, j - 1) + AT(mu, i - 1, j))) + 1e-12);
            double mu_yp = AT(mu, i, j), mu_ym = AT(mu, i, j - 1);
            double dvdxE = (AT(v, i + 1, j) - AT(v, i, j)) * MT(s->m_idx_c, i);
            double dvdxW = (AT(v, i, j) - AT(v, i - 1, j)) * MT(s->m_idx_c, i - 1);
            double dvdyN = (AT(v, i, j + 1) - AT(v, i, j)) * MT(s->m_idy_v, j);
            double dvdyS = (AT(v, i, j) - AT(v, i, j - 1)) * MT(s->m_idy_v, j - 1);
            double d2vdx2 = iRe * ((mu_xp * dvdxE - mu_xm * dvdxW) * MT(s->m_idx_u, i));
            double d2vdy2 = iRe * (mu_yp * dvdyN - mu_ym * dvdyS) * MT(s->m_idy_c, j - 1);
            double d2vdz2 = 0.0;
            double dudyE = (AT(u, i + 1, j) - AT(u, i + 1, j - 1)) * MT(s->m_idy_c, j - 1);
            double dudyW = (AT(u, i, j) - AT(u, i, j - 1)) * MT(s->m_idy_c, j - 1);
            double ddxdudy = iRe * ((mu_xp * dudyE - mu_xm * dudyW) * MT(s->m_idx_u, i));
            double ddydvdy = d2vdy2, ddzdwdy = 0.0, dvwdz = 0.0;
            double uCE = 0.5 * (AT(u, i + 1, j) + AT(u, i + 1, j - 1));
            double uCW = 0.5 * (AT(u, i, j) + AT(u, i, j - 1));
            double vuE = 0.5 * (AT(v, i, j) + AT(v, i + 1, j)) * uCE;
            double vuW = 0.5 * (AT(v, i, j) + AT(v, i - 1, j)) * uCW;
            double vvN = AT(vb, i, j) * AT(vb, i, j);
            double vvS = AT(vb, i, j - 1) * AT(vb, i, j - 1);
            double dvudx = (vuE - vuW) * MT(s->m_idx_u, i);
            double dvvdy = (vvN - vvS) * MT(s->m_idy_c, j - 1);
            AT(ex, i, j) = -(dvudx + dvvdy + dvwdz);
            AT(ve, i, j) = ddxdudy + ddydvdy + ddzdwdy;
            AT(im, i, j) = d2vdx2 + d2vdy2 + d2vdz2;
        }
}

/* Velocity.c:1494-1728 (VOF, VOF_DIFFUSE, XPERIODIC && !BOUSSINESQ) */
void pvo_Velocity_u_set_RHS(pvo_state *s)
{
    int P = s->P, i, j, rk = s->stage;
    double *u = s->f[PV_F_U], *p = s->f[PV_F_P], *rho = s->f[PV_F_RHO];
    double *rhs = s->f[PV_F_U_RHS], *ex = s->f[PV_F_U_EXPL], *exo = s->f[PV_F_U_EXPL_OLD];
    double *ve = s->f[PV_F_U_VISC], *veo = s->f[PV_F_U_VISC_OLD], *im = s->f[PV_F_U_IMPL];
    double a_dt = 1.0 / (s->dt * BET[rk]);
    double dp_dx_source = -s->prm.dp_dx;
    PVO_FOR
    for (j = 0; j < s->NY - 1; j++)
        for (i = 1; i < s->NX; i++) {
            double dpdx = (AT(p, i, j) - AT(p, i - 1, j)) * MT(s->m_idx_c, i - 1);
            double rho_face = 0.5 * (AT(rho, i, j) + AT(rho, i - 1, j));
            double conv = GAMB[rk] * AT(ex, i, j) + ZETB[rk] * AT(exo, i, j);
            conv *= rho_face;
            AT(rhs, i, j) = rho_face * a_dt * AT(u, i, j) + conv + AT(im, i, j) +
                            (GAMB[rk] * AT(ve, i, j) + ZETB[rk] * AT(veo, i, j)) - 2.0 * dpdx;
            AT(rhs, i, j) += 2.0 * dp_dx_source;
            AT(exo, i, j) = AT(ex, i, j);
            AT(veo, i, j) = AT(ve, i, j);
        }
}

/* Velocity.c:1739-1947 */
void pvo_Velocity_v_set_RHS(pvo_state *s)
{
    int P = s->P, i, j, rk = s->stage;
    double *v = s->f[PV_F_V], *p = s->f[PV_F_P], *rho = s->f[PV_F_RHO];
    double *rhs = s->f[PV_F_V_RHS], *ex = s->f[PV_F_V_EXPL], *exo = s->f[PV_F_V_EXPL_OLD];
    double *ve = s->f[PV_F_V_VISC], *veo = s->f[PV_F_V_VISC_OLD], *im = s->f[PV_F_V_IMPL];
    double a_dt = 1.0 / (s->dt * BET[rk]);
    PVO_FOR
    for (j = 1; j < s->NY - 1; j++)
        for (i = 0; i < s->NX - 1; i++) {
            double dpdy = (AT(p, i, j) - AT(p, i, j - 1)) * MT(s->m_idy_c, j - 1);
            double rho_face = 0.5 * (AT(rho, i, j) + AT(rho, i, j - 1));
            double conv = GAMB[rk] * AT(ex, i, j) + ZETB[rk] * AT(exo, i, j);
            conv *= rho_face;
            AT(rhs, i, j) = rho_face * a_dt * AT(v, i, j) + conv + AT(im, i, j) +
                            (GAMB[rk] * AT(ve, i, j) + ZETB[rk] * AT(veo, i, j)) - 2.0 * dpdy;
            AT(exo, i, j) = AT(ex, i, j);
            AT(veo, i, j) = AT(ve, i, j);
        }
}

/* lsolver/msolve_cg.c:103-249 (VOF branch, planar) */
static inline double vel_diag(const pvo_state *s, int i, int j, int comp, double alpha_k)
{
    int P = s->P;
    const double *rho = s->f[PV_F_RHO], *mu = s->f[PV_F_MU];
    const double iRe = 1.0 / s->prm.Re;
    double rho_face, d_time, diag_x, diag_y, diag_z = 0.0, diag_axisym = 0.0;
    double mu_xp, mu_xm, mu_yp, mu_ym;
    if (comp == 'u') {
        rho_face = 0.5 * (AT(rho, i, j) + AT(rho, i - 1, j));
        d_time = rho_face / (alpha_k * s->dt);
        mu_xp = AT(mu, i, j);
        mu_xm = AT(mu, i - 1, j);
        mu_yp = HARM(AVG2(AT(mu, i - 1, j + 1), AT(mu, i, j + 1)), AVG2(AT(mu, i - 1, j), AT(mu, i, j)));
        mu_ym = HARM(AVG2(AT(mu, i - 1, j), AT(mu, i, j)), AVG2(AT(mu, i - 1, j - 1), AT(mu, i, j - 1)));
        diag_x = iRe * ((mu_xp * MT(s->m_idx_u, i) + mu_xm * MT(s->m_idx_u, i - 1)) * MT(s->m_idx_c, i - 1));
        diag_y = iRe * (mu_yp * MT(s->m_idy_c, j) + mu_ym * MT(s->m_idy_c, j - 1)) * MT(s->m_idy_v, j);
    } else {
        rho_face = 0.5 * (AT(rho, i, j) + AT(rho, i, j - 1));
        d_time = rho_face / (alpha_k * s->dt);
        mu_yp = AT(mu, i, j);
        mu_ym = AT(mu, i, j - 1);
        mu_xp = HARM(AVG2(AT(mu, i + 1, j - 1), AT(mu, i + 1, j)), AVG2(AT(mu, i, j - 1), AT(mu, i, j)));
        mu_xm = HARM(AVG2(AT(mu, i, j - 1), AT(mu, i, j)), AVG2(AT(mu, i - 1, j - 1), AT(mu, i - 1, j)));
        diag_x = iRe * ((mu_xp * MT(s->m_idx_c, i) + mu_xm * MT(s->m_idx_c, i - 1)) * MT(s->m_idx_u, i));
        diag_y = iRe * (mu_yp * MT(s->m_idy_v, j) + mu_ym * MT(s->m_idy_v, j - 1)) * MT(s->m_idy_c, j - 1);
    }
    return d_time + diag_x + diag_y + diag_z + diag_axisym;
}

/* lsolver/msolve_cg.c:255-493 (VOF branch, planar) */
static void matvec(pvo_state *s, double *Ax, const double *x, int comp)
{
    int P = s->P, i, j;
    const double *rho = s->f[PV_F_RHO], *mu = s->f[PV_F_MU];
    const double iRe = 1.0 / s->prm.Re, alpha_k = BET[s->stage], dt = s->dt;
    if (comp == 'u') {
        PVO_FOR
        for (j = 0; j < s->NY - 1; j++)
            for (i = 1; i < s->NX; i++) {
                double rho_face = 0.5 * (AT(rho, i, j) + AT(rho, i - 1, j));
                double factor_time = rho_face / (alpha_k * dt);
                double Ax_val = factor_time * AT(x, i, j), op = 0.0;
                double mu_xp = AT(mu, i, j), mu_xm = AT(mu, i - 1, j);
                double mu_yp = HARM(AVG2(AT(mu, i - 1, j + 1), AT(mu, i, j + 1)), AVG2(AT(mu, i - 1, j), AT(mu, i, j)));
                double mu_ym = HARM(AVG2(AT(mu, i - 1, j), AT(mu, i, j)), AVG2(AT(mu, i - 1, j - 1), AT(mu, i, j - 1)));
                double flux_xp = mu_xp * (AT(x, i + 1, j) - AT(x, i, j)) * MT(s->m_idx_u, i);
                double flux_xm = mu_xm * (AT(x, i, j) - AT(x, i - 1, j)) * MT(s->m_idx_u, i - 1);
                double flux_yp = mu_yp * (AT(x, i, j + 1) - AT(x, i, j)) * MT(s->m_idy_c, j);
                double flux_ym = mu_ym * (AT(x, i, j) - AT(x, i, j - 1)) * MT(s->m_idy_c, j - 1);
                op += iRe * ((flux_xp - flux_xm) * MT(s->m_idx_c, i - 1));
                op += iRe * (flux_yp - flux_ym) * MT(s->m_idy_v, j);
                op += 0.0;
                AT(Ax, i, j) = Ax_val - op;
            }
    } else {
        PVO_FOR
        for (j = 1; j < s->NY - 1; j++)
            for (i = 0; i < s->NX - 1; i++) {
                double rho_face = 0.5 * (AT(rho, i, j) + AT(rho, i, j - 1));
                double factor_time = rho_face / (alpha_k * dt);
                double Ax_val = factor_time * AT(x, i, j), op = 0.0;
                double mu_yp = AT(mu, i, j), mu_ym = AT(mu, i, j - 1);
                double mu_xp = HARM(AVG2(AT(mu, i + 1, j - 1), AT(mu, i + 1, j)), AVG2(AT(mu, i, j - 1), AT(mu, i, j)));
                double mu_xm = HARM(AVG2(AT(mu, i, j - 1), AT(mu, i, j)), AVG2(AT(mu, i - 1, j - 1), AT(mu, i - 1, j)));
                double flux_xp = mu_xp * (AT(x, i + 1, j) - AT(x, i, j)) * MT(s->m_idx_c, i);
                double flux_xm = mu_xm * (AT(x, i, j) - AT(x, i - 1, j)) * MT(s->m_idx_c, i - 1);
                double flux_yp = mu_yp * (AT(x, i, j + 1) - AT(x, i, j)) * MT(s->m_idy_v, j);
                double flux_ym = mu_ym * (AT(x, i, j) - AT(x, i, j - 1)) * MT(s->m_idy_v, j - 1);
                op += iRe * ((flux_xp - flux_xm) * MT(s->m_idx_u, i));
                op += iRe * (flux_yp - flux_ym) * MT(s->m_idy_c, j - 1);
                AT(Ax, i, j) = Ax_val - op;
            }
    }
}

void pvo_matVec(pvo_state *s, int out, int in, int component)
{
    matvec(s, s->f[out], s->f[in], component);
}

/* lsolver/msolve_cg.c:34-71, u/v index ranges */
static void vel_range(const pvo_state *s, int comp, int *is, int *ie, int *js, int *je)
{
    if (comp == 'u') { *is = 1; *ie = s->NX; *js = 0; *je = s->NY - 1; }
    else             { *is = 0; *ie = s->NX - 1; *js = 1; *je = s->NY - 1; }
}

/* lsolver/msolve_cg.c:525-722 */
int pvo_Velocity_solve_cg(pvo_state *s, int comp)
{
    int P = s->P, i, j, is, ie, js, je, iters = 0;
    const double EPS = 1e-14, alpha_k = BET[s->stage];
    double *data = s->f[comp == 'u' ? PV_F_U : PV_F_V];
    double *rhs = s->f[comp == 'u' ? PV_F_U_RHS : PV_F_V_RHS];
    double *r = s->f[comp == 'u' ? PV_F_U_R : PV_F_V_R];
    double *d = s->f[comp == 'u' ? PV_F_U_D : PV_F_V_D];
    double *Ad = s->f[comp == 'u' ? PV_F_U_AD : PV_F_V_AD];
    double in_tot, rz = 0.0, RMS = 0.0;
    vel_range(s, comp, &is, &ie, &js, &je);
    if (comp == 'u') in_tot = 1.0 / (double)((s->NX - 1) * (s->NY - 1) * 1);
    else             in_tot = 1.0 / (double)((s->NX - 1) * (s->NY - 2) * 1);

    matvec(s, Ad, data, comp);
    PVO_FOR
    for (j = js; j < je; j++)
        for (i = is; i < ie; i++) {
            double diag;
            AT(r, i, j) = AT(rhs, i, j) - AT(Ad, i, j);
            diag = vel_diag(s, i, j, comp, alpha_k);
            AT(d, i, j) = AT(r, i, j) / (diag + EPS);
        }
    PVO_FOR_SUM(rz)
    for (j = js; j < je; j++)
        for (i = is; i < ie; i++) rz += AT(r, i, j) * AT(d, i, j);
    while (iters < s->prm.CG_MAXIT) {
        double dAd = 0.0, alpha, lv0 = 0.0, lv1 = 0.0, beta;
        iters++;
        vel_bc(s, d, comp, PV_VEL_TYPE_CG);
        matvec(s, Ad, d, comp);
        PVO_FOR_SUM(dAd)
        for (j = js; j < je; j++)
            for (i = is; i < ie; i++) dAd += AT(d, i, j) * AT(Ad, i, j);
        if (s->mom_exact) alpha = (dAd != 0.0) ? rz / dAd : 0.0;
        else alpha = rz / (dAd + EPS);
        PVO_FOR_SUM(lv0, lv1)
        for (j = js; j < je; j++)
            for (i = is; i < ie; i++) {
                double ri, diag;
                AT(data, i, j) += alpha * AT(d, i, j);
                AT(r, i, j) -= alpha * AT(Ad, i, j);
                ri = AT(r, i, j);
                diag = vel_diag(s, i, j, comp, alpha_k);
                lv0 += ri * ri / (diag + EPS);
                lv1 += ri * ri;
            }
        RMS = sqrt(in_tot * lv1);
        if (RMS < s->prm.CG_ETOL) break;
        if (s->mom_exact) beta = (rz != 0.0) ? lv0 / rz : 0.0;
        else beta = lv0 / (rz + EPS);
        rz = lv0;
        PVO_FOR
        for (j = js; j < je; j++)
            for (i = is; i < ie; i++) {
                double diag = vel_diag(s, i, j, comp, alpha_k);
                AT(d, i, j) = AT(r, i, j) / (diag + EPS) + beta * AT(d, i, j);
            }
    }
    if (comp == 'u') s->st.n_u[s->stage] = iters; else s->st.n_v[s->stage] = iters;
    return iters;
}

/* ------------------------------------------------------------------ */
/* pressure                                                            */
/* ------------------------------------------------------------------ */

/* Pressure.c:100-196 (GRID_UNIFORM: idx_u[1] used for both directions) */
void pvo_Pressure_set_RHS(pvo_state *s)
{
    int P = s->P, i, j;
    double *u = s->f[PV_F_U], *v = s->f[PV_F_V], *rhs = s->f[PV_F_P_RHS];
    double a_dt = 1.0 / (s->dt * 2.0 * BET[s->stage]);
    double ihdt = MT(s->m_idx_u, 1) * a_dt;
    PVO_FOR
    for (j = 0; j < s->NY - 1; j++)
        for (i = 0; i < s->NX - 1; i++) {
            double dudx = (AT(u, i + 1, j) - AT(u, i, j)) * ihdt;
            double dvdy = (AT(v, i, j + 1) - AT(v, i, j)) * ihdt;
            AT(rhs, i, j) = dudx + dvdy;
        }
}

/* lsolver/psolve_cg.c:63-101 with TwodOps.c:45-114 */
void pvo_Pressure_compute_preconditioner(pvo_state *s)
{
    int P = s->P, i, j;
    double *rho = s->f[PV_F_RHO], *M = s->f[PV_F_M_INV];
    const double idx2 = MT(s->m_idx_c, 1) * MT(s->m_idx_c, 1);
    const double idy2 = MT(s->m_idy_c, 1) * MT(s->m_idy_c, 1);
    PVO_FOR
    for (j = 0; j < s->NY - 1; j++)
        for (i = 0; i < s->NX - 1; i++) {
            double r0 = AT(rho, i, j);
            double bxp = 2.0 / (r0 + AT(rho, i + 1, j)), bxm = 2.0 / (r0 + AT(rho, i - 1, j));
            double byp = 2.0 / (r0 + AT(rho, i, j + 1)), bym = 2.0 / (r0 + AT(rho, i, j - 1));
            double aii = bxm * idx2 + bxp * idx2 + bym * idy2 + byp * idy2 + 0.0 + 0.0;
            AT(M, i, j) = 1.0 / (1.0 * aii + 1e-12);
        }
}

/* lsolver/psolve_cg.c:302-344 with TwodOps.c:116-142 */
static void p_operator(pvo_state *s, double *Aphi, const double *phi)
{
    int P = s->P, i, j;
    const double *rho = s->f[PV_F_RHO];
    const double idx2 = MT(s->m_idx_c, 1) * MT(s->m_idx_c, 1);
    const double idy2 = MT(s->m_idy_c, 1) * MT(s->m_idy_c, 1);
    PVO_FOR
    for (j = 0; j < s->NY - 1; j++)
        for (i = 0; i < s->NX - 1; i++) {
            double r0 = AT(rho, i, j), c0 = AT(phi, i, j);
            double cxp = (2.0 / (r0 + AT(rho, i + 1, j))) * idx2;
            double cxm = (2.0 / (r0 + AT(rho, i - 1, j))) * idx2;
            double cyp = (2.0 / (r0 + AT(rho, i, j + 1))) * idy2;
            double cym = (2.0 / (r0 + AT(rho, i, j - 1))) * idy2;
            double op = cxp * (AT(phi, i + 1, j) - c0) - cxm * (c0 - AT(phi, i - 1, j)) +
                        cyp * (AT(phi, i, j + 1) - c0) - cym * (c0 - AT(phi, i, j - 1)) +
                        0.0 * 0.0 - 0.0 * 0.0;
            AT(Aphi, i, j) = 1.0 * op;
        }
}

void pvo_Pressure_operator_variableCoeff(pvo_state *s, int out, int in)
{
    p_operator(s, s->f[out], s->f[in]);
}

static double p_inner(const pvo_state *s, const double *a, const double *b)
{
    int P = s->P, i, j;
    double sum = 0.0;
    PVO_FOR_SUM(sum)
    for (j = 0; j < s->NY - 1; j++)
        for (i = 0; i < s->NX - 1; i++) sum += AT(a, i, j) * AT(b, i, j);
    return sum;
}

/* lsolver/psolve_cg.c:130-293 */
static int p_solve_jacobi(pvo_state *s, double *final_res)
{
    int P = s->P, i, j, iters;
    double *phi = s->f[PV_F_PHI], *rhs = s->f[PV_F_P_RHS], *res = s->f[PV_F_P_RES];
    double *d = s->f[PV_F_P_D], *Ad = s->f[PV_F_P_AD], *M = s->f[PV_F_M_INV], *z = s->zbuf;
    double tol = s->prm.P_CG_ETOL, gamma, gamma_new = 0.0;
    int maxIters = s->prm.P_CG_MAXIT;
    memset(z, 0, s->nel * sizeof(double));
    p_bc(s, phi);
    halo_cell(s, phi);
    p_operator(s, Ad, phi);
    PVO_FOR
    for (j = 0; j < s->NY - 1; j++)
        for (i = 0; i < s->NX - 1; i++) AT(res, i, j) = AT(rhs, i, j) - AT(Ad, i, j);
    PVO_FOR
    for (j = 0; j < s->NY - 1; j++)
        for (i = 0; i < s->NX - 1; i++) {
            AT(z, i, j) = AT(res, i, j) * AT(M, i, j);
            AT(d, i, j) = AT(z, i, j);
        }
    gamma = p_inner(s, res, z);
    if (sqrt(gamma) < tol) { *final_res = sqrt(gamma); return 0; }
    for (iters = 1; iters <= maxIters; iters++) {
        double dAd, alpha, beta;
        halo_cell(s, d);
        p_bc(s, d);
        p_operator(s, Ad, d);
        dAd = p_inner(s, d, Ad);
        alpha = gamma / (dAd + 1e-30);
        PVO_FOR
        for (j = 0; j < s->NY - 1; j++)
            for (i = 0; i < s->NX - 1; i++) {
                AT(phi, i, j) += alpha * AT(d, i, j);
                AT(res, i, j) -= alpha * AT(Ad, i, j);
            }
        PVO_FOR
        for (j = 0; j < s->NY - 1; j++)
            for (i = 0; i < s->NX - 1; i++) AT(z, i, j) = AT(res, i, j) * AT(M, i, j);
        gamma_new = p_inner(s, res, z);
        if (sqrt(gamma_new) < tol) break;
        beta = gamma_new / (gamma + 1e-30);
        PVO_FOR
        for (j = 0; j < s->NY - 1; j++)
            for (i = 0; i < s->NX - 1; i++) AT(d, i, j) = AT(z, i, j) + beta * AT(d, i, j);
        gamma = gamma_new;
    }
    {   /* NEED_REFERENCE_PRESSURE gauge, :255-279 */
        double sum = 0.0, avg;
        long cnt = 0;
        PVO_FOR_SUM(sum, cnt)
        for (j = 0; j < s->NY - 1; j++)
            for (i = 0; i < s->NX - 1; i++) { sum += AT(phi, i, j); ++cnt; }
        avg = sum / (double)cnt;
        PVO_FOR
        for (j = 0; j < s->NY - 1; j++)
            for (i = 0; i < s->NX - 1; i++) AT(phi, i, j) -= avg;
        halo_cell(s, phi);
    }
    if (iters > maxIters) iters = maxIters;
    *final_res = sqrt(gamma_new);
    return iters;
}

/* ---- FFT-x / tridiagonal-y constant-coefficient solve (A23) -------- */

/* In-place iterative radix-2 complex FFT of length n (power of two) on
 * w->fre / w->fim; sign = -1 forward, +1 inverse (unnormalised). */
static void fft_pow2(const pvo_state *s, struct pvo_ws *w, int n, int sign)
{
    double *re = w->fre, *im = w->fim;
    int i, j, k, len;
    for (i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { double t = re[i]; re[i] = re[j]; re[j] = t; t = im[i]; im[i] = im[j]; im[j] = t; }
    }
    for (len = 2; len <= n; len <<= 1) {
        int step = n / len;
        for (i = 0; i < n; i += len)
            for (k = 0; k < len / 2; k++) {
                double wr = s->twc[k * step], wi = sign * s->tws[k * step];
                double xr = re[i + k + len / 2], xi = im[i + k + len / 2];
                double tr = wr * xr - wi * xi, ti = wr * xi + wi * xr;
                re[i + k + len / 2] = re[i + k] - tr; im[i + k + len / 2] = im[i + k] - ti;
                re[i + k] += tr; im[i + k] += ti;
            }
    }
}

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* Real DFT of length n in FFTW R2HC layout: hc[k] = Re X_k (k <= n/2),
 * hc[n-k] = Im X_k (0 < k < n/2), X_k = sum_m x_m exp(-2 pi i k m / n).
 * Radix-2 FFT for powers of two, direct summation otherwise. */
static void r2hc(const pvo_state *s, struct pvo_ws *w, const double *x, double *hc, int n)
{
    int k, m;
    if (is_pow2(n)) {
        for (m = 0; m < n; m++) { w->fre[m] = x[m]; w->fim[m] = 0.0; }
        fft_pow2(s, w, n, -1);
        for (k = 0; k <= n / 2; k++) { w->wre[k] = w->fre[k]; w->wim[k] = w->fim[k]; }
    } else {
        for (k = 0; k <= n / 2; k++) {
            double re = 0.0, im = 0.0;
            for (m = 0; m < n; m++) {
                long km = ((long)k * m) % n;
                re += x[m] * s->twc[km];
                im -= x[m] * s->tws[km];
            }
            w->wre[k] = re; w->wim[k] = im;
        }
    }
    for (k = 0; k <= n / 2; k++) hc[k] = w->wre[k];
    for (k = 1; k < (n + 1) / 2; k++) hc[n - k] = w->wim[k];
}

/* Inverse of r2hc without normalisation (FFTW HC2R). */
static void hc2r(const pvo_state *s, struct pvo_ws *w, const double *hc, double *x, int n)
{
    int k, m;
    if (is_pow2(n)) {
        w->fre[0] = hc[0]; w->fim[0] = 0.0;
        w->fre[n / 2] = hc[n / 2]; w->fim[n / 2] = 0.0;
        for (k = 1; k < n / 2; k++) {
            w->fre[k] = hc[k]; w->fim[k] = hc[n - k];
            w->fre[n - k] = hc[k]; w->fim[n - k] = -hc[n - k];
        }
        fft_pow2(s, w, n, +1);
        for (m = 0; m < n; m++) x[m] = w->fre[m];
        return;
    }
    for (m = 0; m < n; m++) {
        double acc = hc[0];
        for (k = 1; k < (n + 1) / 2; k++) {
            long km = ((long)k * m) % n;
            acc += 2.0 * (hc[k] * s->twc[km] - hc[n - k] * s->tws[km]);
        }
        if (n % 2 == 0) acc += hc[n / 2] * ((m % 2) ? -1.0 : 1.0);
        x[m] = acc;
    }
}

/* lsolver/psolve_fft.c:141-163 */
static void tridiag(const double *a, const double *b, const double *c, const double *r,
                    double *u, double *gam, int N)
{
    double bet = b[0];
    int i;
    u[0] = r[0] / bet;
    for (i = 1; i < N; i++) {
        gam[i] = c[i - 1] / bet;
        bet = b[i] - a[i] * gam[i];
        u[i] = (r[i] - a[i] * u[i - 1]) / bet;
    }
    for (i = N - 2; i >= 0; i--) u[i] = u[i] - gam[i + 1] * u[i + 1];
}

/* psolve_fft.c:17-76 (wave[]), :467-591 (coefficients), :1536-1802
 * (x-periodic solve; z has length 1 so modwaveksq = 0).  out = L0^{-1} rhs
 * with L0 = Lx + Ly on cells i < NXM, j < NYM; mode (0,0) pinned at j = 0.
 * mean_free != 0 additionally removes the mean of the zero-wavenumber
 * column (used when the solve serves as a PCG preconditioner). */
static void fft_solve(pvo_state *s, const double *rhs, double *out, int mean_free)
{
    int P = s->P, NXM = s->NXM, NYM = s->NYM, NY = s->NY, i, j;
    double idx = MT(s->m_idx_u, 0), inx = 1.0 / (double)NXM;
    double *spec = s->zbuf;     /* NYM rows of NXM halfcomplex values */
    PVO_FOR
    for (j = 0; j < NYM; j++) {
        struct pvo_ws *w = &s->ws[PVO_TID];
        for (i = 0; i < NXM; i++) w->wrow[i] = AT(rhs, i, j);
        r2hc(s, w, w->wrow, &spec[(size_t)j * NXM], NXM);
        for (i = 0; i < NXM; i++) spec[(size_t)j * NXM + i] *= inx;
    }
    PVO_FOR_COLS
    for (i = 0; i < NXM; i++) {
        struct pvo_ws *w = &s->ws[PVO_TID];
        int wave = (i < NXM / 2) ? i : (NXM - i);
        double mw = 2.0 * (1 - cos(2. * PI_REF * wave / (NXM))) * (idx * idx);
        for (j = 0; j < NYM; j++) {
            double as = (j > 0) ? MT(s->m_idy_v, j) * MT(s->m_idy_c, j - 1) : 0.0;
            double an = (j < NY - 2) ? MT(s->m_idy_v, j) * MT(s->m_idy_c, j) : 0.0;
            w->tsa[j] = as; w->tsc[j] = an;
            w->tsb[j] = (-as - an) - 0.0 - mw;
            w->tsr[j] = spec[(size_t)j * NXM + i];
        }
        if (i == 0) { w->tsa[0] = 0.0; w->tsc[0] = 0.0; w->tsb[0] = 1.0; w->tsr[0] = 0.0; }
        tridiag(w->tsa, w->tsb, w->tsc, w->tsr, w->tsu, w->tsg, NYM);
        if (i == 0 && mean_free) {
            double m = 0.0;
            for (j = 0; j < NYM; j++) m += w->tsu[j];
            m /= (double)NYM;
            for (j = 0; j < NYM; j++) w->tsu[j] -= m;
        }
        for (j = 0; j < NYM; j++) spec[(size_t)j * NXM + i] = w->tsu[j];
    }
    PVO_FOR
    for (j = 0; j < NYM; j++) {
        struct pvo_ws *w = &s->ws[PVO_TID];
        hc2r(s, w, &spec[(size_t)j * NXM], w->wrow, NXM);
        for (i = 0; i < NXM; i++) AT(out, i, j) = w->wrow[i];
    }
}

void pvo_Pressure_solve_fft(pvo_state *s, int rhs_field, int out_field)
{
    double *tmp = (double *)calloc(s->nel, sizeof(double));
    fft_solve(s, s->f[rhs_field], tmp, 0);
    memcpy(s->f[out_field], tmp, s->nel * sizeof(double));
    free(tmp);
    halo_cell(s, s->f[out_field]);
}

/* PCG on the variable-coefficient operator preconditioned by the FFT-x /
 * Thomas-y solve of the unit-coefficient Laplacian; x0 = 0, right-hand side
 * projected to zero mean, stop on ||r||_2 / ||b||_2 <= p_rel_tol, then the
 * reference's mean-removal gauge (psolve_cg.c:255-279). */
static int p_solve_fft_pcg(pvo_state *s, double *final_res)
{
    int P = s->P, i, j, it = 0, maxit = s->prm.p_max_iter > 0 ? s->prm.p_max_iter : 1000;
    double *phi = s->f[PV_F_PHI], *rhs = s->f[PV_F_P_RHS], *res = s->f[PV_F_P_RES];
    double *d = s->f[PV_F_P_D], *Ad = s->f[PV_F_P_AD];
    double *z = (double *)calloc(s->nel, sizeof(double));
    double bnorm, gamma, rel = 0.0, mean = 0.0;
    long cnt = (long)(s->NX - 1) * (s->NY - 1);
    PVO_FOR_SUM(mean)
    for (j = 0; j < s->NY - 1; j++)
        for (i = 0; i < s->NX - 1; i++) mean += AT(rhs, i, j);
    mean /= (double)cnt;
    PVO_FOR
    for (j = 0; j < s->NY - 1; j++)
        for (i = 0; i < s->NX - 1; i++) {
            AT(res, i, j) = AT(rhs, i, j) - mean;
            AT(phi, i, j) = 0.0;
        }
    bnorm = sqrt(p_inner(s, res, res));
    if (bnorm == 0.0) bnorm = 1.0;
    rel = sqrt(p_inner(s, res, res)) / bnorm;
    if (rel > s->prm.p_rel_tol) {
        fft_solve(s, res, z, 1);
        PVO_FOR
        for (j = 0; j < s->NY - 1; j++)
            for (i = 0; i < s->NX - 1; i++) AT(d, i, j) = AT(z, i, j);
        gamma = p_inner(s, res, z);
        while (it < maxit) {
            double dAd, alpha, gnew, beta;
            halo_cell(s, d);
            p_bc(s, d);
            p_operator(s, Ad, d);
            dAd = p_inner(s, d, Ad);
            alpha = gamma / dAd;
            PVO_FOR
            for (j = 0; j < s->NY - 1; j++)
                for (i = 0; i < s->NX - 1; i++) {
                    AT(phi, i, j) += alpha * AT(d, i, j);
                    AT(res, i, j) -= alpha * AT(Ad, i, j);
                }
            it++;
            rel = sqrt(p_inner(s, res, res)) / bnorm;
            if (rel <= s->prm.p_rel_tol) break;
            fft_solve(s, res, z, 1);
            gnew = p_inner(s, res, z);
            beta = gnew / gamma;
            PVO_FOR
            for (j = 0; j < s->NY - 1; j++)
                for (i = 0; i < s->NX - 1; i++) AT(d, i, j) = AT(z, i, j) + beta * AT(d, i, j);
            gamma = gnew;
        }
    }
    {
        double sum = 0.0, avg;
        PVO_FOR_SUM(sum)
        for (j = 0; j < s->NY - 1; j++)
            for (i = 0; i < s->NX - 1; i++) sum += AT(phi, i, j);
        avg = sum / (double)cnt;
        PVO_FOR
        for (j = 0; j < s->NY - 1; j++)
            for (i = 0; i < s->NX - 1; i++) AT(phi, i, j) -= avg;
        halo_cell(s, phi);
    }
    free(z);
    *final_res = rel;
    return it;
}

int pvo_Pressure_solve_cg(pvo_state *s)
{
    double res = 0.0;
    int it;
    if (s->prm.pressure_solver == PV_PSOLVE_FFT_PCG) it = p_solve_fft_pcg(s, &res);
    else it = p_solve_jacobi(s, &res);
    s->st.n_p[s->stage] += it;
    s->st.p_res[s->stage] = res;
    return it;
}

/* Pressure.c:218-464 (VOF branch; w reset in 2-D) */
void pvo_Pressure_project_velocity(pvo_state *s)
{
    int P = s->P, i, j;
    double *u = s->f[PV_F_U], *v = s->f[PV_F_V], *p = s->f[PV_F_P], *dp = s->f[PV_F_PHI];
    double *rho = s->f[PV_F_RHO];
    double mdt = 2.0 * BET[s->stage] * s->dt;
    p_bc(s, dp);
    halo_cell(s, dp);
    PVO_FOR
    for (j = 0; j < s->NY - 1; j++)
        for (i = 0; i < s->NX - 1; i++) AT(p, i, j) += AT(dp, i, j);
    p_bc(s, p);
    halo_cell(s, p);
    PVO_FOR
    for (j = 0; j < s->NY - 1; j++)
        for (i = 1; i < s->NX; i++) {
            double inv_rho_face = 2.0 / (AT(rho, i, j) + AT(rho, i - 1, j));
            double idxdt = MT(s->m_idx_c, i - 1) * mdt;
            AT(u, i, j) -= (AT(dp, i, j) - AT(dp, i - 1, j)) * (idxdt * inv_rho_face);
        }
    PVO_FOR
    for (j = 1; j < s->NY - 1; j++)
        for (i = 0; i < s->NX - 1; i++) {
            double inv_rho_face = 2.0 / (AT(rho, i, j) + AT(rho, i, j - 1));
            double idydt = MT(s->m_idy_c, j - 1) * mdt;
            AT(v, i, j) -= (AT(dp, i, j) - AT(dp, i, j - 1)) * (idydt * inv_rho_face);
        }
}

/* Pressure.c:475-597 */
double pvo_Pressure_compute_velocity_divergence(pvo_state *s)
{
    int P = s->P, i, j;
    double *u = s->f[PV_F_U], *v = s->f[PV_F_V], div_max = 0.0;
    for (j = 0; j < s->NY - 1; j++)
        for (i = 0; i < s->NX - 1; i++) {
            double dudx = (AT(u, i + 1, j) - AT(u, i, j)) * MT(s->m_idx_u, i);
            double dvdy = (AT(v, i, j + 1) - AT(v, i, j)) * MT(s->m_idy_v, j);
            double dc = fabs(dudx + dvdy);
            if (dc >= div_max) div_max = dc;
        }
    return div_max;
}

/* ------------------------------------------------------------------ */
/* time step and substep                                               */
/* ------------------------------------------------------------------ */

/* Dtime.c:24-249 (FULLY_IMPLICIT: viscous = 0; SURFACE_TENSION cap) */
double pvo_Dtime_cfl(pvo_state *s)
{
    int P = s->P, i, j;
    double *ub = s->f[PV_F_U_BC], *vb = s->f[PV_F_V_BC];
    double max_idt = 0.0, dt_cfl, dt, viscous = 0.0, settling = 0.0;
    for (j = 0; j < s->NY; j++)
        for (i = 0; i < s->NX; i++) {
            double u_cfl = fabs(AT(ub, i, j));
            double v_cell = fabs(AT(vb, i, j));
            double v_part = fabs(AT(vb, i, j) - settling);
            double v_cfl = MAXD(v_cell, v_part);
            double convective = u_cfl * MT(s->m_idx_u, i) + v_cfl * MT(s->m_idy_v, j);
            max_idt = MAXD(max_idt, convective + viscous);
        }
    dt_cfl = s->prm.cfl / max_idt;
    {
        double dx_min = MIND(1.0 / MT(s->m_idx_u, 0), 1.0 / MT(s->m_idy_v, 0));
        double dt_sigma = s->prm.cfl * sqrt(((s->prm.rho1 + s->prm.rho2) * pow(dx_min, 3) * s->prm.We) /
                                            (4.0 * PI_REF));
        dt_cfl = MIND(dt_cfl, dt_sigma);
    }
    if (s->prm.constant_dt == 1) return s->prm.default_dt;
    dt = dt_cfl;
    if (dt > 1.05 * s->dt_old) dt = 1.05 * s->dt_old;
    if (dt < 0.95 * s->dt_old) dt = 0.95 * s->dt_old;
    else dt = MIND(dt, s->prm.max_dt);
    return dt;
}

/* Temporal_int.c:563-1112 (no IBM, no CONC/LES/RANS) */
int pvo_Temporal_int_all_the_equations(pvo_state *s)
{
    int passes = 0, st = s->stage;
    double div;
    if (st > 0) pvo_Velocity_cell_center(s);
    halo_cell(s, s->f[PV_F_U_BC]);
    halo_cell(s, s->f[PV_F_V_BC]);
    s->st.n_p[st] = 0;
    pvo_VOF_DIFFUSE_step(s);
    pvo_VOF_update_density_viscosity(s);
    pvo_Velocity_u_set_implicit_explicit(s);
    pvo_Velocity_v_set_implicit_explicit(s);
    pvo_Velocity_u_set_RHS(s);
    pvo_Velocity_v_set_RHS(s);
    pvo_VOF_DIFFUSE_compute_psi_LG(s);
    pvo_VOF_DIFFUSE_compute_f_sigma(s);
    pvo_VOF_apply_f_sigma_old(s);
    pvo_Velocity_add_gravity_2_RHS(s);
    pvo_Velocity_solve_cg(s, 'u');
    pvo_Velocity_solve_cg(s, 'v');
    vel_bc(s, s->f[PV_F_U], 'u', PV_VEL_TYPE_NORMAL);
    vel_bc(s, s->f[PV_F_V], 'v', PV_VEL_TYPE_NORMAL);
    do {
        memset(s->f[PV_F_PHI], 0, s->nel * sizeof(double));
        pvo_Pressure_set_RHS(s);
        pvo_Pressure_compute_preconditioner(s);
        pvo_Pressure_solve_cg(s);
        pvo_Pressure_project_velocity(s);
        vel_bc(s, s->f[PV_F_U], 'u', PV_VEL_TYPE_NORMAL);
        vel_bc(s, s->f[PV_F_V], 'v', PV_VEL_TYPE_NORMAL);
        div = pvo_Pressure_compute_velocity_divergence(s);
        passes++;
        if (passes > 10) { s->st.n_pass[st] = passes; return PV_ENOCONV; }
    } while (div > 1e-6);
    s->st.n_pass[st] = passes;
    s->st.div_max[st] = div;
    pvo_VOF_swap_f_sigma(s);
    return PV_OK;
}

/* Temporal_int.c:399-448 */
int pvo_Temporal_int_rk3(pvo_state *s, int nsteps)
{
    int n, rk, rc;
    for (n = 0; n < nsteps; n++) {
        for (rk = 0; rk < 3; rk++) {
            if (s->time != 0 && rk == 0) {
                pvo_Velocity_cell_center(s);
                s->dt_old = s->dt;
                s->dt = pvo_Dtime_cfl(s);
            }
            s->stage = rk;
            rc = pvo_Temporal_int_all_the_equations(s);
            if (rc) return rc;
            s->time += s->dt * BET2[rk];
        }
        s->ntime++;
    }
    return PV_OK;
}

/* VOF_DIFFUSE.c:524-568 without IBM: C_L = clamp01(F) over the L range,
 * then boundary values / phase cache / clamp / boundary values / phase
 * cache twice, and the boundary values of C_G; F ends as the mirror of C_L
 * that the phase cache keeps (:108-126). */
void pvo_VOF_DIFFUSE_init(pvo_state *s)
{
    double *C = s->f[PV_F_C_L], *F = s->f[PV_F_F];
    size_t n;
    for (n = 0; n < s->nel; n++) C[n] = (F[n] < 0.0) ? 0.0 : ((F[n] > 1.0) ? 1.0 : F[n]);
    bc_ch(s, C);
    pvo_VOF_DIFFUSE_update_phase_cache(s);
    pvo_VOF_DIFFUSE_bound_liquid_fraction(s);
    bc_ch(s, C);
    pvo_VOF_DIFFUSE_update_phase_cache(s);
    bc_ch(s, C);
    pvo_VOF_DIFFUSE_update_phase_cache(s);
    bc_ch(s, s->f[PV_F_C_G]);
    memcpy(F, C, s->nel * sizeof(double));          /* the phase cache's F mirror (:108-126) */
}

/* Cart3d.c:1094-1130: boundary values of F, VOF_DIFFUSE_init, density and
 * viscosity, psi_LG and f_sigma, new -> old; Temporal_int.c:163-199 start. */
int pvo_init_from_F(pvo_state *s, const double *F)
{
    int k;
    for (k = 0; k < PV_NFIELDS; k++) memset(s->f[k], 0, s->nel * sizeof(double));
    memcpy(s->f[PV_F_F], F, s->nel * sizeof(double));
    bc_ch(s, s->f[PV_F_F]);
    pvo_VOF_DIFFUSE_init(s);
    pvo_VOF_update_density_viscosity(s);
    pvo_VOF_DIFFUSE_compute_psi_LG(s);
    pvo_VOF_DIFFUSE_compute_f_sigma(s);
    memcpy(s->f[PV_F_FSX_OLD], s->f[PV_F_FSX_NEW], s->nel * sizeof(double));
    memcpy(s->f[PV_F_FSY_OLD], s->f[PV_F_FSY_NEW], s->nel * sizeof(double));
    s->dt_old = s->dt = s->prm.default_dt;
    s->time = 0.0;
    s->ntime = 1;
    s->stage = 0;
    memset(&s->st, 0, sizeof(s->st));
    pvo_Velocity_cell_center(s);
    halo_cell(s, s->f[PV_F_P]);
    return PV_OK;
}
