/*
 * pv_host.c -- C host layer of the MI355X PARTIES substep.
 *
 * Owns one pv_bag (one 2-D subdomain on one GPU): device allocation, field
 * I/O, and the sequencing of Temporal_int_all_the_equations
 * (PARTIES/src/Temporal_int.c:563-1112) on a single HIP stream.  All
 * arithmetic happens in csrc/pv_kernels.hip; solver scalars (alpha, beta,
 * stop flags, dt) live in device memory so the host only enqueues work and
 * reads a flag every few iterations.
 */
#include "pv_host.h"
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static __thread char g_err[512];

const char *pv_last_error(void) { return g_err; }

int pv_set_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(call)                                                                 \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess)                                                        \
            return pv_set_error(PV_EHIP, "%s failed: %s (%s:%d)", #call,             \
                                hipGetErrorString(e_), __FILE__, __LINE__);          \
    } while (0)

/* IO/default.inp defaults for the 2-D no-IBM path */
void pv_params_default(pv_params *p)
{
    memset(p, 0, sizeof *p);
    p->NXM = 64; p->NYM = 64;
    p->xmin = 0.0; p->xmax = 1.0; p->ymin = 0.0; p->ymax = 1.0;
    p->twod_slab_thickness = -1.0;
    p->Re = 100.0; p->dp_dx = 0.0;
    p->We = 1.0; p->rho1 = 1.0; p->rho2 = 1.0; p->mu1 = 1.0; p->mu2 = 1.0;
    p->Cn = -1.0; p->Pe_CH = -1.0;
    p->weno_order = 5; p->ch_iter_max = 300; p->ch_tol = 1e-9;
    p->cfl = 0.5; p->default_dt = 1e-3; p->max_dt = 0.05; p->constant_dt = 0;
    p->grav[0] = 0.0; p->grav[1] = -1.0; p->grav[2] = 0.0;
    p->richardson = 1.0;
    p->P_CG_MAXIT = 300; p->P_CG_ETOL = 1e-7;
    p->CG_MAXIT = 120; p->CG_ETOL = 1e-9;
    p->pressure_solver = PV_PSOLVE_JACOBI_PCG;
    p->p_rel_tol = 1e-10; p->p_max_iter = 2000;
    p->precision = PV_PREC_FP64;
}

/* IO/Input.c:20-127: slab thickness, Cn and Pe_CH defaults */
void pv_derive_params(pv_params *p, double *slab)
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

/* Grid.c:845-861 coordinates and Grid.c:385-420 metrics; arrays hold
 * entries for indices -3 .. N+3 at offset 3. */
void pv_host_metrics(const pv_params *p, pv_metrics *m)
{
    int NX = p->NXM + 1, NY = p->NYM + 1, i, j;
    double dx = (p->xmax - p->xmin) / (double)(NX - 1);
    double dy = (p->ymax - p->ymin) / (double)(NY - 1);
    memset(m, 0, sizeof *m);
    m->nx = p->NXM + 7; m->ny = p->NYM + 7;
    m->xu = calloc(m->nx, sizeof(double)); m->xc = calloc(m->nx, sizeof(double));
    m->yv = calloc(m->ny, sizeof(double)); m->yc = calloc(m->ny, sizeof(double));
    m->dx_u = calloc(m->nx, sizeof(double)); m->dx_c = calloc(m->nx, sizeof(double));
    m->idx_u = calloc(m->nx, sizeof(double)); m->idx_c = calloc(m->nx, sizeof(double));
    m->dy_v = calloc(m->ny, sizeof(double)); m->dy_c = calloc(m->ny, sizeof(double));
    m->idy_v = calloc(m->ny, sizeof(double)); m->idy_c = calloc(m->ny, sizeof(double));
    for (i = -3; i < NX + 3; i++) {
        m->xu[i + 3] = p->xmin + i * dx;
        m->xc[i + 3] = p->xmin + (0.5 + i) * dx;
    }
    for (j = -3; j < NY + 3; j++) {
        m->yv[j + 3] = p->ymin + j * dy;
        m->yc[j + 3] = p->ymin + (0.5 + j) * dy;
    }
    for (i = -3; i < NX + 2; i++) {
        m->dx_u[i + 3] = m->xu[i + 4] - m->xu[i + 3];
        m->dx_c[i + 3] = m->xc[i + 4] - m->xc[i + 3];
        m->idx_u[i + 3] = 1.0 / m->dx_u[i + 3];
        m->idx_c[i + 3] = 1.0 / m->dx_c[i + 3];
    }
    for (j = -3; j < NY + 2; j++) {
        m->dy_v[j + 3] = m->yv[j + 4] - m->yv[j + 3];
        m->dy_c[j + 3] = m->yc[j + 4] - m->yc[j + 3];
        m->idy_v[j + 3] = 1.0 / m->dy_v[j + 3];
        m->idy_c[j + 3] = 1.0 / m->dy_c[j + 3];
    }
}

void pv_host_metrics_free(pv_metrics *m)
{
    free(m->xu); free(m->xc); free(m->yv); free(m->yc);
    free(m->dx_u); free(m->dx_c); free(m->idx_u); free(m->idx_c);
    free(m->dy_v); free(m->dy_c); free(m->idy_v); free(m->idy_c);
    memset(m, 0, sizeof *m);
}

/* ------------------------------------------------------------------ */
/* handle                                                              */
/* ------------------------------------------------------------------ */

#define OFF0 13                     /* (OFF0 + 3) % 16 == 0: column 0 on 128 B */

static double *fld(pv_bag *b, int f) { return b->f[f]; }

/* doubles of one field allocation, and its cell (0,0) given the allocation base */
size_t pv_field_elems(const pv_bag *b) { return b->alloc_elems; }
/* offset of cell (0, 0) from the start of its field's allocation */
size_t pv_field_front(const pv_bag *b) { return (size_t)OFF0 + 3 * (size_t)b->P + 3; }
double *pv_field_origin(const pv_bag *b, double *base) { return base + OFF0 + 3 * (size_t)b->P + 3; }

static void fill_phys(pv_bag *b)
{
    pv_params *p = &b->prm;
    pvk_phys *ph = &b->ph;
    double gm;
    memset(ph, 0, sizeof *ph);
    ph->Re = p->Re; ph->iRe = 1.0 / p->Re; ph->Pe_CH = p->Pe_CH; ph->Cn = p->Cn; ph->We = p->We;
    ph->rho_g = p->rho2 / p->rho1; ph->mu_g = p->mu2 / p->mu1;
    ph->bih = p->Cn * p->Cn / p->Pe_CH;
    gm = sqrt(p->grav[0] * p->grav[0] + p->grav[1] * p->grav[1]);
    ph->gx = 0.0; ph->gy = 0.0;
    if (gm > 0.0) { ph->gx = p->grav[0] / gm; ph->gy = p->grav[1] / gm; }
    ph->Ri = p->richardson; ph->dp_dx = p->dp_dx;
    ph->rho1 = p->rho1; ph->rho2 = p->rho2; ph->cfl = p->cfl; ph->max_dt = p->max_dt;
    ph->default_dt = p->default_dt;
    ph->ch_tol = (p->ch_tol > 0.0) ? p->ch_tol : p->CG_ETOL;
    ph->CG_ETOL = p->CG_ETOL; ph->P_CG_ETOL = p->P_CG_ETOL; ph->p_rel_tol = p->p_rel_tol;
    ph->weno_order = p->weno_order; ph->constant_dt = p->constant_dt;
    ph->CG_MAXIT = p->CG_MAXIT; ph->P_CG_MAXIT = p->P_CG_MAXIT;
    ph->ch_iter_max = (p->ch_iter_max > 0) ? p->ch_iter_max : p->CG_MAXIT;
    ph->p_max_iter = (p->p_max_iter > 0) ? p->p_max_iter : 1000;
    ph->precision = p->precision;
}

static int create_on_device(pv_bag *b, const pv_params *prm);

int pv_create(const pv_params *prm, int device, pv_bag **out)
{
    pv_bag *b;
    int ndev = 0, rc;
    hipDeviceProp_t prop;
    if (!out) return pv_set_error(PV_EINVAL, "null output handle pointer");
    *out = NULL;
    if (!prm || prm->NXM < 8 || prm->NYM < 8)
        return pv_set_error(PV_EINVAL, "NXM and NYM must be >= 8");
    if (prm->precision != PV_PREC_FP64 && prm->precision != PV_PREC_MIXED)
        return pv_set_error(PV_EINVAL, "unknown precision %d", prm->precision);
    if (prm->pressure_solver < PV_PSOLVE_JACOBI_PCG || prm->pressure_solver > PV_PSOLVE_LEARNED_MG)
        return pv_set_error(PV_EINVAL, "unknown pressure_solver %d", prm->pressure_solver);
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0)
        return pv_set_error(PV_ENODEV, "no HIP device %d (count %d)", device, ndev);
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipGetDeviceProperties(&prop, device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return pv_set_error(PV_ENODEV, "device %d is %s; this library is built for gfx950", device,
                            prop.gcnArchName);
    b = (pv_bag *)calloc(1, sizeof *b);
    if (!b) return pv_set_error(PV_ENOMEM, "host allocation failed");
    b->device = device;
    /* every failure after this point releases what was made so far */
    if ((rc = create_on_device(b, prm)) != PV_OK) {
        pv_destroy(b);
        return rc;
    }
    *out = b;
    return PV_OK;
}

static int create_on_device(pv_bag *b, const pv_params *prm)
{
    int k;
    size_t bytes;
    pv_metrics m;
    b->prm = *prm;
    pv_derive_params(&b->prm, &b->slab);
    b->NXM = prm->NXM; b->NYM = prm->NYM;
    b->NXT = prm->NXM + 7; b->NYT = prm->NYM + 7;
    b->P = ((b->NXT + 15) / 16) * 16;
    b->alloc_elems = (size_t)OFF0 + (size_t)b->P * (size_t)b->NYT + 16;
    bytes = b->alloc_elems * sizeof(double) * PV_FI_END;
    if (hipMalloc((void **)&b->slabmem, bytes) != hipSuccess) {
        b->slabmem = NULL;
        return pv_set_error(PV_ENOMEM, "hipMalloc of %zu bytes failed", bytes);
    }
    HIPCHK(hipMemset(b->slabmem, 0, bytes));
    for (k = 0; k < PV_FI_END; k++)
        b->f[k] = b->slabmem + (size_t)k * b->alloc_elems + OFF0 + 3 * (size_t)b->P + 3;
    HIPCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));

    /* metrics */
    pv_host_metrics(&b->prm, &m);
    memcpy(&b->hm, &m, sizeof m);      /* owned by the handle from here on */
    {
        size_t nmx = m.nx, nmy = m.ny;
        double *dm;
        HIPCHK(hipMalloc((void **)&dm, sizeof(double) * (4 * nmx + 4 * nmy)));
        b->metricmem = dm;
        HIPCHK(hipMemcpy(dm + 0 * nmx, m.idx_u, nmx * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dm + 1 * nmx, m.idx_c, nmx * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dm + 2 * nmx, m.dx_u, nmx * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dm + 3 * nmx, m.dx_c, nmx * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dm + 4 * nmx + 0 * nmy, m.idy_v, nmy * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dm + 4 * nmx + 1 * nmy, m.idy_c, nmy * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dm + 4 * nmx + 2 * nmy, m.dy_v, nmy * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dm + 4 * nmx + 3 * nmy, m.dy_c, nmy * 8, hipMemcpyHostToDevice));
        b->g.idx_u = dm + 0 * nmx + 3;
        b->g.idx_c = dm + 1 * nmx + 3;
        b->g.dx_u = dm + 2 * nmx + 3;
        b->g.dx_c = dm + 3 * nmx + 3;
        b->g.idy_v = dm + 4 * nmx + 0 * nmy + 3;
        b->g.idy_c = dm + 4 * nmx + 1 * nmy + 3;
        b->g.dy_v = dm + 4 * nmx + 2 * nmy + 3;
        b->g.dy_c = dm + 4 * nmx + 3 * nmy + 3;
        b->dx_min = (1.0 / m.idx_u[3] > 1.0 / m.idy_v[3]) ? 1.0 / m.idy_v[3] : 1.0 / m.idx_u[3];
        b->h_idx_u0 = m.idx_u[3];
        b->h_idy_v0 = m.idy_v[3];
    }
    b->g.NXM = b->NXM; b->g.NYM = b->NYM; b->g.P = b->P;
    b->g.nbx = (b->NXM + 63) / 64; b->g.nby = (b->NYM + 3) / 4;
    b->g.slab = b->slab;
    fill_phys(b);

    /* reductions: up to 4 partial sets per launch, plus the CFL grid */
    b->rd.nblocks = b->g.nbx * b->g.nby;
    {
        int nb_cfl = ((b->NXM + 1 + 63) / 64) * ((b->NYM + 1 + 3) / 4);
        int nmax = (nb_cfl > b->rd.nblocks) ? nb_cfl : b->rd.nblocks;
        HIPCHK(hipMalloc((void **)&b->rd.part, sizeof(double) * 4 * (size_t)nmax));
    }
    HIPCHK(hipMalloc((void **)&b->d_sc, sizeof(pvk_scal)));
    HIPCHK(hipMemset(b->d_sc, 0, sizeof(pvk_scal)));
    HIPCHK(hipHostMalloc((void **)&b->h_sc, sizeof(pvk_scal), hipHostMallocDefault));
    memset(b->h_sc, 0, sizeof(pvk_scal));
    b->time = 0.0; b->dt = b->dt_old = b->prm.default_dt; b->stage = 0; b->ntime = 1; b->time_nonzero = 0;
    b->ch_chunk = 8; b->mom_chunk[0] = b->mom_chunk[1] = 4; b->p_chunk = 16;
    {
        int rc = pv_cg_create(b);
        if (rc) return rc;
        if ((rc = pv_fft_create(b))) return rc;
        if ((rc = pv_mg_create(b))) return rc;
        if ((b->prm.pressure_solver == PV_PSOLVE_MG_PCG || b->prm.pressure_solver == PV_PSOLVE_LEARNED_MG) &&
            !b->mg_ok)
            return pv_set_error(PV_EINVAL, "MG-PCG: %d x %d does not halve to <= %d cells and a last level "
                                "of <= %d cells and <= %d columns", prm->NXM, prm->NYM, PVK_MG_LDS_CELLS,
                                PVK_MG_BOTTOM, PVK_MG_BOTTOM_NX);
        return pv_sync_scalars_to_device(b);
    }
}

/* Work sets of the fused biharmonic and momentum CG loops: one band of rows
 * per 256-thread block (1 row for the biharmonic CG, 1, 2 or 4 rows for the
 * momentum CG: 2 unless pv_set_momentum_rows says otherwise), one
 * partial-sum slot per band and reduction (sized for one-row bands). */
int pv_cg_create(pv_bag *b)
{
    const int nx = b->NXM, ny = b->NYM;
    const int band = 1, mband = 2;
    int nbc = (ny + band - 1) / band, nbu = ny, nbv = ny - 1, c;
    int npap = ((ny + 7) / 8) * ((nx + 127) / 128);
    int nt0 = ((ny + 15) / 16) * ((nx + 63) / 64), nbs = (nt0 > nbc) ? nt0 : nbc;   /* start partials */
    size_t n = 2 * (size_t)nbs + (size_t)nbc + (size_t)npap + 4 * (size_t)nbu + 4 * (size_t)nbv;
    double *q;
    HIPCHK(hipMalloc((void **)&b->cgmem, sizeof(double) * n));
    HIPCHK(hipMemset(b->cgmem, 0, sizeof(double) * n));
    q = b->cgmem;
    memset(&b->ccg, 0, sizeof b->ccg);
    b->ccg.nb = nbc; b->ccg.band = band;
    b->ccg.p[0] = b->f[PV_F_CH_DIR]; b->ccg.p[1] = b->f[PV_FI_CHP2];
    b->ccg.bb = q; q += nbs;
    b->ccg.rr[0] = q; q += nbs;
    b->ccg.nt0 = nt0;
    b->ccg.rr[1] = q; q += nbc;
    b->ccg.pap = q; q += npap;
    b->ccg.npap = npap;
    memset(&b->mcg, 0, sizeof b->mcg);
    b->mcg.band = mband;
    b->mcg.rho = b->f[PV_F_RHO]; b->mcg.mu = b->f[PV_F_MU];
    b->mcg.maxit = b->ph.CG_MAXIT; b->mcg.tol = b->ph.CG_ETOL;
    /* RMS over the unknowns (msolve_cg.c:691): NXM * NYM for u, NXM * (NYM - 1) for v */
    b->mcg.in_tot[0] = 1.0 / (double)(nx * ny);
    b->mcg.in_tot[1] = 1.0 / (double)(nx * (ny - 1));
    for (c = 0; c < 2; c++) {
        int nbk = c ? nbv : nbu;              /* slots for one-row bands */
        b->mcg.x[c] = b->f[c ? PV_F_V : PV_F_U];
        b->mcg.r[c] = b->f[c ? PV_F_V_R : PV_F_U_R];
        b->mcg.dg[c] = b->f[c ? PV_FI_DG_V : PV_FI_DG_U];
        b->mcg.Ad[c] = b->f[c ? PV_F_V_AD : PV_F_U_AD];
        b->mcg.d[c][0] = b->f[c ? PV_F_V_D : PV_F_U_D];
        b->mcg.d[c][1] = b->f[c ? PV_FI_DV2 : PV_FI_DU2];
        b->mcg.z[c] = b->f[c ? PV_FI_MZV : PV_FI_MZU];
        b->mcg.r2[c] = b->f[c ? PV_FI_MRV2 : PV_FI_MRU2];
        b->mcg.rhs[c] = b->f[c ? PV_F_V_RHS : PV_F_U_RHS];
        b->mcg.rz[c][0] = q; q += nbk;
        b->mcg.rz[c][1] = q; q += nbk;
        b->mcg.dad[c] = q; q += nbk;
        b->mcg.rr[c] = q; q += nbk;
        b->mcg.nb[c] = (c ? ny - 1 + mband - 1 : ny + mband - 1) / mband;
        /* coefficient k of component c: half (k & 1) of field MCF(3c + k / 2),
         * each half NYM rows of pitch P floats from the field's cell (0, 0) */
        for (int k = 0; k < 5; k++)
            b->mcg.cf[c][k] = (float *)b->f[PV_FI_MCF0 + 3 * c + k / 2] + (size_t)(k & 1) * ny * b->g.P;
    }
    return PV_OK;
}

int pv_destroy(pv_bag *b)
{
    if (!b) return PV_OK;
    hipSetDevice(b->device);
    if (b->stream) hipStreamSynchronize(b->stream);
    if (b->timers_on) pv_timers_enable(b, 0);
    pv_fft_destroy(b);
    pv_mg_destroy(b);
    pv_learn_free(b);
    pv_graphs_free(b);
    pv_cnn_free(b);
    hipFree(b->cgmem);
    hipFree(b->slabmem);
    hipFree(b->metricmem);
    hipFree(b->rd.part);
    hipFree(b->d_sc);
    if (b->h_sc) hipHostFree(b->h_sc);
    if (b->stream) hipStreamDestroy(b->stream);
    pv_host_metrics_free(&b->hm);
    free(b);
    return PV_OK;
}

int pv_get_params(const pv_bag *b, pv_params *out)
{
    if (!b || !out) return pv_set_error(PV_EINVAL, "null argument");
    *out = b->prm;
    return PV_OK;
}

static int check_field(pv_bag *b, int f)
{
    if (!b) return pv_set_error(PV_EINVAL, "null handle");
    if (f < 0 || f >= PV_NFIELDS) return pv_set_error(PV_EINVAL, "bad field id %d", f);
    return PV_OK;
}

int pv_set_field(pv_bag *b, int f, const double *host)
{
    int rc = check_field(b, f);
    if (rc) return rc;
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipMemcpy2DAsync(b->f[f] - 3 * (size_t)b->P - 3, (size_t)b->P * 8, host, (size_t)b->NXT * 8,
                            (size_t)b->NXT * 8, b->NYT, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return PV_OK;
}

int pv_get_field(pv_bag *b, int f, double *host)
{
    int rc = check_field(b, f);
    if (rc) return rc;
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipMemcpy2DAsync(host, (size_t)b->NXT * 8, b->f[f] - 3 * (size_t)b->P - 3, (size_t)b->P * 8,
                            (size_t)b->NXT * 8, b->NYT, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return PV_OK;
}

int pv_fill_field(pv_bag *b, int f, double value)
{
    int rc = check_field(b, f);
    if (rc) return rc;
    pvk_fill(b->f[f] - 3 * (size_t)b->P - 3, (size_t)b->P * b->NYT, value, b->stream);
    HIPCHK(hipGetLastError());
    return PV_OK;
}

/* host mirror of time state <-> device scalars */
int pv_sync_scalars_to_device(pv_bag *b)
{
    b->h_sc->dt = b->dt; b->h_sc->dt_old = b->dt_old; b->h_sc->time = b->time;
    HIPCHK(hipMemcpyAsync(b->d_sc, b->h_sc, sizeof(pvk_scal), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return PV_OK;
}

int pv_read_scalars(pv_bag *b)
{
    HIPCHK(hipMemcpyAsync(b->h_sc, b->d_sc, sizeof(pvk_scal), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    if (b->timers_on) {
        int rc = pv_timers_flush(b);
        if (rc) return rc;
    }
    b->dt = b->h_sc->dt; b->dt_old = b->h_sc->dt_old; b->time = b->h_sc->time;
    return PV_OK;
}

int pv_set_time_state(pv_bag *b, double time, double dt, double dt_old, int which_stage)
{
    if (!b || which_stage < 0 || which_stage > 2) return pv_set_error(PV_EINVAL, "bad time state");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemcpy(b->h_sc, b->d_sc, sizeof(pvk_scal), hipMemcpyDeviceToHost));
    b->time = time; b->dt = dt; b->dt_old = dt_old; b->stage = which_stage;
    b->time_nonzero = (time != 0.0);
    b->have_mprev[0] = b->have_mprev[1] = b->have_mprev[2] = 0;   /* increments of another run */
    b->mg_bottom_age[0] = b->mg_bottom_age[1] = -1;                /* rebuilt at the next solve */
    return pv_sync_scalars_to_device(b);
}

int pv_set_step_count(pv_bag *b, int ntime)
{
    if (!b || ntime < 0) return pv_set_error(PV_EINVAL, "bad step count");
    b->ntime = ntime;
    return PV_OK;
}

int pv_get_stats(pv_bag *b, pv_stats *out)
{
    int rc;
    if (!b || !out) return pv_set_error(PV_EINVAL, "null argument");
    HIPCHK(hipSetDevice(b->device));
    rc = pv_read_scalars(b);
    if (rc) return rc;
    *out = b->st;
    out->time = b->time; out->dt = b->dt; out->dt_old = b->dt_old;
    out->ntime = b->ntime; out->which_stage = b->stage;
    return PV_OK;
}

/* device handles for other host files */
double *pv_dev_field(pv_bag *b, int f) { return fld(b, f); }

/* ------------------------------------------------------------------ */
/* FFT-x / DCT-y tables (csrc/pv_fft.hip)                              */
/* ------------------------------------------------------------------ */
static int ilog2_exact(int n)
{
    int l = 0;
    if (n <= 0 || (n & (n - 1))) return -1;
    while ((1 << l) < n) l++;
    return l;
}

/* NXM = 2^a 3^b with 16 | NXM: the row transforms are complex transforms of
 * length NXM / 2 with 8 | NXM / 2 (M / 8 threads each, mixed radix 8/4/2/3) */
static int xlen_ok(int n)
{
    if (n < 16 || n % 16) return 0;
    while (n % 2 == 0) n /= 2;
    while (n % 3 == 0) n /= 3;
    return n == 1;
}

/* Launch shape of the FFT preconditioner's y solves (k_tri_cols): segments
 * of rows rows per thread (at most 16, the registers a thread keeps per
 * row; 0 = 8) and cols columns per block (a power of two; 0 = 2, or more
 * where a block would hold less than a wavefront; halved while the block
 * would exceed 512 threads).  The kernel is latency-bound (a chain of FP64 divisions per
 * segment): short segments and many blocks (512 at 1024 columns) run
 * fastest -- 20.6 us at 1024^2 against 25.1 us for 16 rows and 8 columns.
 * The blocks of one XCD take neighbouring columns, so each row's cache lines
 * come from HBM once per XCD.  Nonzero if no shape fits. */
static int tri_shape(pv_bag *b, int rows, int cols)
{
    const int ny = b->NYM;
    int L = rows ? rows : 8, S, kw = cols ? cols : 2;
    if (L > ny) L = ny;
    S = (ny + L - 1) / L;
    L = (ny + S - 1) / S;
    S = (ny + L - 1) / L;
    if (!cols)
        while (kw < 64 && kw * S < 64) kw <<= 1;   /* at least one wavefront a block */
    while (kw > 1 && kw * S > 512) kw >>= 1;
    if (kw * S > 512 || L > 16) return 1;
    b->fft.tri_l = L; b->fft.tri_s = S; b->fft.tri_kw = kw;
    for (b->fft.tri_lkw = 0; (1 << b->fft.tri_lkw) < kw; b->fft.tri_lkw++) {}
    return 0;
}

int pv_set_fft_tri_shape(pv_bag *b, int rows, int cols)
{
    if (!b) return pv_set_error(PV_EINVAL, "null handle");
    if (!b->fft.ok) return pv_set_error(PV_EINVAL, "the FFT solve is not available for this grid size");
    if (rows < 0 || rows > 16) return pv_set_error(PV_EINVAL, "rows per thread must be in [0, 16] (got %d)", rows);
    if (cols < 0 || cols > 64 || (cols & (cols - 1)))
        return pv_set_error(PV_EINVAL, "columns per block must be 0 or a power of two <= 64 (got %d)", cols);
    {
        pvk_fft keep = b->fft;
        if (tri_shape(b, rows, cols)) {
            b->fft = keep;
            return pv_set_error(PV_EINVAL, "no y-solve shape of %d rows and %d columns fits 512 threads", rows, cols);
        }
    }
    pv_graphs_free(b);                      /* captured iterations hold the old shape */
    return PV_OK;
}

int pv_fft_create(pv_bag *b)
{
    const double PI_ = 3.14159265358979323846;
    int nx = b->NXM, ny = b->NYM, lx = ilog2_exact(nx), ly = ilog2_exact(ny), k;
    size_t n_rtx, n_rty, n_dct, n_wx, n_wy, total;
    double *h, *d, *q;
    double idx2, idy2;
    memset(&b->fft, 0, sizeof b->fft);
    b->fftmem = NULL;
    /* FFT-PCG unavailable for this size: one row of N/2 complex doubles and
     * its twiddles must fit the LDS budget of the row transforms (N <= 4096) */
    if (!xlen_ok(nx) || nx > 4096 || ny < 8 || ny > 8192) return PV_OK;
    b->fft.mx = nx / 2;
    b->fft.pow2x = (lx >= 4);
    b->fft.logmx = (lx >= 4) ? lx - 1 : 0;
    /* the y transforms of the Cahn-Hilliard spectral start: power-of-two NYM */
    b->fft.dct_ok = (ly >= 3 && ny <= 4096);
    b->fft.logmy = (ly >= 3) ? ly - 1 : 0;
    n_rtx = 2 * (size_t)(nx / 2 + 1); n_rty = 2 * (size_t)(ny / 2 + 1);
    n_dct = 2 * (size_t)ny;
    n_wx = 2 * (size_t)(nx / 2); n_wy = 2 * (size_t)(ny / 2);
    total = n_rtx + n_rty + n_dct + nx + ny + n_wx + n_wy + 2 * (size_t)ny + nx;
    h = (double *)calloc(total, sizeof(double));
    if (!h) return pv_set_error(PV_ENOMEM, "FFT table allocation failed");
    q = h;
    for (k = 0; k <= nx / 2; k++) { double a = 2.0 * PI_ * k / nx; q[2 * k] = cos(a); q[2 * k + 1] = sin(a); }
    q += n_rtx;
    for (k = 0; k <= ny / 2; k++) { double a = 2.0 * PI_ * k / ny; q[2 * k] = cos(a); q[2 * k + 1] = sin(a); }
    q += n_rty;
    for (k = 0; k < ny; k++) { double a = PI_ * k / (2.0 * ny); q[2 * k] = cos(a); q[2 * k + 1] = sin(a); }
    q += n_dct;
    /* psolve_fft.c:577-583 modified wavenumbers with idx = idx_u[0]; the y
     * coefficient idy_v * idy_c of the uniform Neumann tridiagonal. */
    idx2 = b->hm.idx_u[3] * b->hm.idx_u[3];
    idy2 = b->hm.idy_v[3] * b->hm.idy_c[3];
    for (k = 0; k < nx; k++) {
        int wave = (k < nx / 2) ? k : (nx - k);
        q[k] = 2.0 * (1 - cos(2. * PI_ * wave / nx)) * idx2;
    }
    q += nx;
    for (k = 0; k < ny; k++) q[k] = 2.0 * (1 - cos(PI_ * k / ny)) * idy2;
    q += ny;
    /* Stockham twiddles of the half-length complex transforms */
    for (k = 0; k < nx / 2; k++) { double a = 2.0 * PI_ * k / (nx / 2); q[2 * k] = cos(a); q[2 * k + 1] = -sin(a); }
    q += n_wx;
    for (k = 0; k < ny / 2; k++) { double a = 2.0 * PI_ * k / (ny / 2); q[2 * k] = cos(a); q[2 * k + 1] = -sin(a); }
    q += n_wy;
    /* y tridiagonals of psolve_fft.c:467-591 (Neumann walls: as_0 = 0,
     * an_{NYM-1} = 0) and the modified wavenumber of each halfcomplex column
     * (:577-583, idx = idx_u[0]) */
    for (k = 0; k < ny; k++) q[k] = (k > 0) ? b->hm.idy_v[k + 3] * b->hm.idy_c[k + 2] : 0.0;
    q += ny;
    for (k = 0; k < ny; k++) q[k] = (k < ny - 1) ? b->hm.idy_v[k + 3] * b->hm.idy_c[k + 3] : 0.0;
    q += ny;
    for (k = 0; k < nx; k++) {
        int wave = (k < nx / 2) ? k : (nx - k);
        q[k] = 2.0 * (1 - cos(2. * PI_ * wave / (nx))) * (b->hm.idx_u[3] * b->hm.idx_u[3]);
    }
    if (tri_shape(b, 0, 0)) { free(h); return PV_OK; }
    if (hipMalloc((void **)&d, sizeof(double) * (total + (size_t)nx * ny)) != hipSuccess) {
        free(h);
        return pv_set_error(PV_ENOMEM, "FFT table device allocation failed");
    }
    if (hipMemcpy(d, h, sizeof(double) * total, hipMemcpyHostToDevice) != hipSuccess) {
        free(h);
        return pv_set_error(PV_EHIP, "FFT table upload failed");
    }
    free(h);
    b->fftmem = d;
    q = d;
    /* fused PCG work set: spectrum transpose + five partial buffers */
    {
        /* row transforms use M/8 threads each (M = N/2, Stockham radix 8) in
         * 256-thread blocks: 256 / (M/8) rows per block, 32 KiB of LDS */
        int band = (ny >= 512) ? 2 : 1, rrow, rdct, nband, nrow, np;
        double *pm;
        rrow = 256 / ((nx / 2 >= 8) ? nx / 16 : 1);
        rdct = 256 / ((ny / 2 >= 8) ? ny / 16 : 1);
        nband = (ny + band - 1) / band;
        nrow = (ny + rrow - 1) / rrow;
        np = ny;                       /* the operator kernel stores one partial per row */
        if (nband > np) np = nband;
        if (nrow > np) np = nrow;
        if (hipMalloc((void **)&pm, sizeof(double) * ((size_t)nx * ny + 7 * (size_t)np)) != hipSuccess)
            return pv_set_error(PV_ENOMEM, "PCG work allocation failed");
        b->pcgmem = pm;
        memset(&b->pcg, 0, sizeof b->pcg);
        b->pcg.specT = pm;
        b->pcg.dad[0] = pm + (size_t)nx * ny;
        b->pcg.dad[1] = b->pcg.dad[0] + np;
        b->pcg.zad = b->pcg.dad[1] + np;
        b->pcg.rr = b->pcg.zad + np;
        b->pcg.rz[0] = b->pcg.rr + np;
        b->pcg.rz[1] = b->pcg.rz[0] + np;
        b->pcg.mean = b->pcg.rz[1] + np;
        b->pcg.ndad = ny; b->pcg.nmean = nband;
        b->pcg.nrr = nrow; b->pcg.nrz = nrow;
        b->pcg.band = band; b->pcg.rrow = rrow; b->pcg.rdct = rdct;
        b->pcg.phi = b->f[PV_F_PHI]; b->pcg.r = b->f[PV_F_P_RES]; b->pcg.z = b->f[PV_FI_Z];
        b->pcg.d[0] = b->f[PV_F_P_D]; b->pcg.d[1] = b->f[PV_FI_D2]; b->pcg.Ad = b->f[PV_F_P_AD];
        b->pcg.rhs = b->f[PV_F_P_RHS]; b->pcg.rho = b->f[PV_F_RHO];
        b->pcg.bx = b->f[PV_FI_PBX]; b->pcg.by = b->f[PV_FI_PBY];
    }
    b->fft.rtx = q; q += n_rtx;
    b->fft.rty = q; q += n_rty;
    b->fft.dcty = q; q += n_dct;
    b->fft.lamx = q; q += nx;
    b->fft.lamy = q; q += ny;
    b->fft.wx = q; q += n_wx;
    b->fft.wy = q; q += n_wy;
    b->fft.tri_as = q; q += ny;
    b->fft.tri_an = q; q += ny;
    b->fft.tri_mw = q; q += nx;
    b->fft.spec = q;
    b->fft.ok = 1;
    return PV_OK;
}

void pv_fft_destroy(pv_bag *b)
{
    if (b->pcgmem) hipFree(b->pcgmem);
    b->pcgmem = NULL;
    if (b->fftmem) hipFree(b->fftmem);
    b->fftmem = NULL;
    b->fft.ok = 0;
}

/* ------------------------------------------------------------------ */
/* kernel timers                                                       */
/* ------------------------------------------------------------------ */
void pv_tbegin(pv_bag *b, int t)
{
    if (!b->timers_on) return;
    b->tactive[t] = (b->tseen[t]++ % b->timers_on) == 0;
    if (!b->tactive[t]) return;
    if (b->tpending[t] >= PV_TPOOL) {
        hipStreamSynchronize(b->stream);
        pv_timers_flush(b);
    }
    hipEventRecord(b->tev[t][0][b->tpending[t]], b->stream);
}

void pv_tend(pv_bag *b, int t)
{
    if (!b->timers_on || !b->tactive[t]) return;
    hipEventRecord(b->tev[t][1][b->tpending[t]], b->stream);
    b->tpending[t]++;
}

/* caller has synchronised the stream */
int pv_timers_flush(pv_bag *b)
{
    int t, k;
    for (t = 0; t < PV_NTIMERS; t++) {
        for (k = 0; k < b->tpending[t]; k++) {
            float ms = 0.0f;
            HIPCHK(hipEventElapsedTime(&ms, b->tev[t][0][k], b->tev[t][1][k]));
            b->tms[t] += ms;
            b->tcount[t]++;
            if (b->tsamp && b->tnsamp[t] < PV_TSAMP) b->tsamp[(size_t)t * PV_TSAMP + b->tnsamp[t]++] = ms;
        }
        b->tpending[t] = 0;
    }
    return PV_OK;
}

int pv_timers_enable(pv_bag *b, int on)
{
    int t, k;
    if (!b) return pv_set_error(PV_EINVAL, "null handle");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    if (on && !b->timers_on) {
        b->tsamp = (float *)calloc((size_t)PV_NTIMERS * PV_TSAMP, sizeof(float));
        if (!b->tsamp) return pv_set_error(PV_ENOMEM, "timer sample allocation failed");
        memset(b->tnsamp, 0, sizeof b->tnsamp);
        for (t = 0; t < PV_NTIMERS; t++)
            for (k = 0; k < PV_TPOOL; k++) {
                /* timing-only events: no system-scope fence (cache writeback) per record */
                HIPCHK(hipEventCreateWithFlags(&b->tev[t][0][k], hipEventDisableSystemFence));
                HIPCHK(hipEventCreateWithFlags(&b->tev[t][1][k], hipEventDisableSystemFence));
            }
    } else if (!on && b->timers_on) {
        pv_timers_flush(b);
        for (t = 0; t < PV_NTIMERS; t++)
            for (k = 0; k < PV_TPOOL; k++) {
                hipEventDestroy(b->tev[t][0][k]);
                hipEventDestroy(b->tev[t][1][k]);
            }
        free(b->tsamp);
        b->tsamp = NULL;
    }
    b->timers_on = (on > 0) ? on : 0;
    memset(b->tseen, 0, sizeof b->tseen);
    return PV_OK;
}

int pv_timers_reset(pv_bag *b)
{
    if (!b) return pv_set_error(PV_EINVAL, "null handle");
    HIPCHK(hipStreamSynchronize(b->stream));
    if (b->timers_on) pv_timers_flush(b);
    memset(b->tms, 0, sizeof b->tms);
    memset(b->tcount, 0, sizeof b->tcount);
    memset(b->tnsamp, 0, sizeof b->tnsamp);
    return PV_OK;
}

/* Like pv_timers_read, over the working launches only: the sampled launches
 * (the first PV_TSAMP per timer since the last reset) that ran at least a
 * quarter of the timer's longest sampled launch.  A CG launch queued past
 * convergence exits on the device stop flag in a few microseconds and does
 * none of the work its class's bytes per launch count. */
int pv_timers_read_working(pv_bag *b, double *ms, long long *count)
{
    int t, k;
    if (!b) return pv_set_error(PV_EINVAL, "null handle");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    if (b->timers_on) pv_timers_flush(b);
    for (t = 0; t < PV_NTIMERS; t++) {
        const float *v = b->tsamp ? b->tsamp + (size_t)t * PV_TSAMP : NULL;
        float mx = 0.0f;
        double sum = 0.0;
        long long n = 0;
        for (k = 0; v && k < b->tnsamp[t]; k++) mx = (v[k] > mx) ? v[k] : mx;
        for (k = 0; v && k < b->tnsamp[t]; k++)
            if (v[k] >= 0.25f * mx) {
                sum += v[k];
                n++;
            }
        if (ms) ms[t] = sum;
        if (count) count[t] = n;
    }
    return PV_OK;
}

int pv_timers_read(pv_bag *b, double *ms, long long *count)
{
    int t;
    if (!b) return pv_set_error(PV_EINVAL, "null handle");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    if (b->timers_on) pv_timers_flush(b);
    for (t = 0; t < PV_NTIMERS; t++) {
        if (ms) ms[t] = b->tms[t];
        if (count) count[t] = b->tcount[t];
    }
    return PV_OK;
}

int pv_synchronize(pv_bag *b)
{
    if (!b) return pv_set_error(PV_EINVAL, "null handle");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    return PV_OK;
}
