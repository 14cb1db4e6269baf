// pv_cg.hip -- fused Krylov loops of the Cahn-Hilliard biharmonic solve
// (VOF_DIFFUSE.c:339-429) and of the momentum Jacobi-PCG (msolve_cg.c:525-722).
//
// Each iteration is two launches of 256-thread blocks, each block owning a
// band of rows (pv_block.h):
//   K1 apply : stop test and beta from the previous partials; the new search
//              direction p = z + beta p_old evaluated on the fly at every
//              stencil point through the ghost-fill index maps (so no ghost
//              fill kernel runs); A p; partial <p, A p>.
//   K2 update: alpha; x += alpha p; r -= alpha A p; partials of the next
//              residual products.
// A scalar that an earlier kernel already formed (||b||, the previous <r,z>)
// is stored by block 0 in a parity-indexed slot that only later kernels read,
// so a kernel re-adds at most two partial sets.  The arithmetic of every
// update is the reference's, in its order; only reduction order differs.
#include <hip/hip_runtime.h>
#include <math.h>
#include "pv_kernels.h"
#include "pv_stencil.h"
#include "pv_block.h"

#define CPT 4   // cells per thread per pass: loads of CPT cells issue before any store

// ---------------------------------------------------------------------------
// Cahn-Hilliard
// ---------------------------------------------------------------------------
// CH ghost fill as an index map: x periodic cell image, y zero gradient
// (VOF_DIFFUSE.c:571-683; all ghost layers equal the first interior value).
__device__ __forceinline__ int chy(int j, int n) { return j < 0 ? 0 : (j >= n ? n - 1 : j); }

// column index of the x-periodic CH map for i in [-NXM, 3 NXM)
__device__ __forceinline__ int chx(int i, int n)
{
    while (i < 0) i += n;
    while (i >= n) i -= n;
    return i;
}

// The CG start kernels work on tiles of CIT_X x CIT_Y cells: C on the tile
// with a three-cell halo through the CH ghost maps, then on the tile with a
// one-cell halo S(C) and the inner Laplacian of C (CH ghost rules: column
// NXM is column 0, rows -1 / NYM are rows 0 / NYM-1), and the WENO face
// fluxes of the tile's faces, each formed once in LDS instead of once per
// stencil use.  Every value is the arithmetic of the per-cell helpers
// (weno_rhs_at, diffusion_at, bilap_at) on the same operands, so the result
// is bitwise that of the untiled form.  One <b, b> and <r, r> partial per
// tile (w.nt0 of them).
#define CIT_X 64
#define CIT_Y 16
#define CIT_NX (CIT_X + 6)

template <class R> struct CitLds {
    double c[CIT_Y + 6][CIT_NX];           // C, halo 3
    double l[CIT_Y + 2][CIT_X + 2];        // inner Laplacian of C, halo 1
    R s[CIT_Y + 2][CIT_X + 2];             // S(C), halo 1
    double fx[CIT_Y][CIT_X + 1];           // u C_face at x face c0 + b of row j0 + a
    double fy[CIT_Y + 1][CIT_X];           // v C_face at y face j0 + a of column c0 + b
};

// tile of this block (XCD-aware order), its origin, C and the inner Laplacian
__device__ __forceinline__ int cit_tile(int &c0, int &j0)
{
    const int t = xcd_band(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int ty = t / gridDim.x, tx = t - ty * gridDim.x;
    c0 = tx * CIT_X;
    j0 = ty * CIT_Y;
    return t;
}
template <class R>
__device__ __forceinline__ void cit_load(const pvk_grid &g, const double *C, int c0, int j0, CitLds<R> &T, bool with_s)
{
    const int P = g.P, NXM = g.NXM, NYM = g.NYM;
    for (int q = threadIdx.x; q < (CIT_Y + 6) * CIT_NX; q += NTH) {
        const int a = q / CIT_NX, b = q - a * CIT_NX;
        T.c[a][b] = AT(C, chx(c0 + b - 3, NXM), chy(j0 + a - 3, NYM));
    }
    __syncthreads();
    for (int q = threadIdx.x; q < (CIT_Y + 2) * (CIT_X + 2); q += NTH) {
        const int a = q / (CIT_X + 2), b = q - a * (CIT_X + 2), jb = j0 + a - 1;
        if (with_s) T.s[a][b] = bulkS<R>(R(T.c[a + 2][b + 2]));
        if (jb >= 0 && jb < NYM)
            T.l[a][b] = lap_vals(g, chx(c0 + b - 1, NXM), jb, T.c[a + 2][b + 2], T.c[a + 2][b + 3], T.c[a + 2][b + 1],
                                 T.c[a + 3][b + 2], T.c[a + 1][b + 2]);
    }
}
// L(L C) at tile cell (tr, tc) = global (i, j)
template <class R>
__device__ __forceinline__ double cit_bilap(const pvk_grid &g, const CitLds<R> &T, int tr, int tc, int i, int j)
{
    const double L0 = T.l[tr + 1][tc + 1];
    const double LN = (j + 1 == g.NYM) ? L0 : T.l[tr + 2][tc + 1];
    const double LS = (j == 0) ? L0 : T.l[tr][tc + 1];
    return lap_vals(g, i, j, L0, T.l[tr + 1][tc + 2], T.l[tr + 1][tc], LN, LS);
}

// Explicit CH right-hand side and CG start: rc -> rn, r = p = b - A C;
// partials ||b||^2, ||r||^2.
template <class R>
__global__ void __launch_bounds__(NTH) k_ccg_init_rhs(pvk_grid g, pvk_phys ph, int stage, pvk_scal *sc,
                                                     const double *__restrict__ C, const double *__restrict__ u,
                                                     const double *__restrict__ v, double *__restrict__ rn,
                                                     double *__restrict__ r, double *__restrict__ p, pvk_ccg w,
                                                     double *__restrict__ bout)
{
    __shared__ double sh[NTH];
    __shared__ CitLds<R> T;
    const int P = g.P, NXM = g.NXM, NYM = g.NYM, wo = ph.weno_order;
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) { sc->ch_done = 0; sc->ch_iter = 0; }
    int c0, j0;
    const int t = cit_tile(c0, j0);
    cit_load<R>(g, C, c0, j0, T, true);
    // face fluxes (the C tile is complete after cit_load's first barrier)
    for (int q = threadIdx.x; q < CIT_Y * (CIT_X + 1); q += NTH) {
        const int a = q / (CIT_X + 1), b = q - a * (CIT_X + 1), i = c0 + b, j = j0 + a;
        if (i <= NXM && j < NYM) {
            const double vel = AT(u, i, j);
            const double *cr = &T.c[a + 3][b];                 // cr[k] = C at column i - 3 + k
            T.fx[a][b] = vel * (double)face_vals<R>(cr[0], cr[1], cr[2], cr[3], cr[4], cr[5], vel, wo);
        }
    }
    for (int q = threadIdx.x; q < (CIT_Y + 1) * CIT_X; q += NTH) {
        const int a = q / CIT_X, b = q - a * CIT_X, i = c0 + b, j = j0 + a;
        if (i < NXM && j <= NYM) {
            const double vel = AT(v, i, j);
            T.fy[a][b] = vel * (double)face_vals<R>(T.c[a][b + 3], T.c[a + 1][b + 3], T.c[a + 2][b + 3],
                                                    T.c[a + 3][b + 3], T.c[a + 4][b + 3], T.c[a + 5][b + 3], vel, wo);
        }
    }
    __syncthreads();
    const double a_dt = 1.0 / (sc->dt * c_BET[stage]);
    double sbb = 0.0, srr = 0.0;
    for (int q = threadIdx.x; q < CIT_X * CIT_Y; q += NTH) {
        const int tr = q / CIT_X, tc = q - tr * CIT_X, i = c0 + tc, j = j0 + tr;
        if (i < NXM && j < NYM) {
            double rc = weno_div(g, i, j, T.fx[tr][tc], T.fx[tr][tc + 1], T.fy[tr][tc], T.fy[tr + 1][tc]);
            rc += diffusion_vals<R>(g, ph, i, j, T.s[tr + 1][tc + 1], T.s[tr + 1][tc + 2], T.s[tr + 1][tc],
                                    T.s[tr + 2][tc + 1], T.s[tr][tc + 1]);
            double cc = T.c[tr + 3][tc + 3];
            double rv = (stage == 0) ? a_dt * cc + c_GAMB[0] * rc
                                     : a_dt * cc + c_GAMB[stage] * rc + c_ZETB[stage] * AT(rn, i, j);
            double L2 = cit_bilap<R>(g, T, tr, tc, i, j);
            double b = rv - ph.bih * L2;
            double Ap = a_dt * cc + ph.bih * L2;
            double rr = b - Ap;
            if (bout) AT(bout, i, j) = b;
            AT(rn, i, j) = rc;
            AT(r, i, j) = rr;
            AT(p, i, j) = rr;
            sbb += ch_w(g, i, j, b, b);
            srr += ch_w(g, i, j, rr, rr);
        }
    }
    block_store_partial(sbb, w.bb, t, sh);
    block_store_partial(srr, w.rr[0], t, sh);
}

// CG start from a given b (b already holds rhs - (Cn^2/Pe) L^2 C).
__global__ void __launch_bounds__(NTH) k_ccg_init_b(pvk_grid g, pvk_phys ph, int stage, pvk_scal *sc,
                                                   const double *__restrict__ C, const double *__restrict__ bv,
                                                   double *__restrict__ r, double *__restrict__ p, pvk_ccg w)
{
    __shared__ double sh[NTH];
    __shared__ CitLds<double> T;
    const int P = g.P, NXM = g.NXM, NYM = g.NYM;
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) { sc->ch_done = 0; sc->ch_iter = 0; }
    int c0, j0;
    const int t = cit_tile(c0, j0);
    cit_load<double>(g, C, c0, j0, T, false);
    __syncthreads();
    const double a_dt = 1.0 / (sc->dt * c_BET[stage]);
    double sbb = 0.0, srr = 0.0;
    for (int q = threadIdx.x; q < CIT_X * CIT_Y; q += NTH) {
        const int tr = q / CIT_X, tc = q - tr * CIT_X, i = c0 + tc, j = j0 + tr;
        if (i < NXM && j < NYM) {
            double cc = T.c[tr + 3][tc + 3];
            double Ap = a_dt * cc + ph.bih * cit_bilap<double>(g, T, tr, tc, i, j);
            double bb = AT(bv, i, j);
            double rr = bb - Ap;
            AT(r, i, j) = rr;
            AT(p, i, j) = rr;
            sbb += ch_w(g, i, j, bb, bb);
            srr += ch_w(g, i, j, rr, rr);
        }
    }
    block_store_partial(sbb, w.bb, t, sh);
    block_store_partial(srr, w.rr[0], t, sh);
}

// K1: stop test on ||r_n|| (VOF_DIFFUSE.c:403-414), beta, p_n = r + beta p_{n-1}
// (:416-422), Ap = a_dt p + (Cn^2/Pe) L^2 p (:318-337), partials of
// <p, Ap> (one per block).  A block owns a tile of
// CHT_X x CHT_Y cells: p on the tile with a two-cell halo (CH ghost maps) and
// the inner Laplacian with a one-cell halo are formed once in LDS, so each p
// value is formed once per tile instead of once per stencil use; every value is formed by the arithmetic of bilap_at
// (inner Laplacian with the CH ghost rules, then the outer one).
#define CHT_X 128
#define CHT_Y 8
__global__ void __launch_bounds__(NTH) k_ccg_apply(pvk_grid g, pvk_phys ph, int stage, pvk_scal *sc, pvk_ccg w,
                                                  const double *__restrict__ r, const double *__restrict__ pold,
                                                  double *__restrict__ pnew, double *__restrict__ Ap, int n,
                                                  int maxit, double tol)
{
    __shared__ double sh[NTH];
    __shared__ double sp[CHT_Y + 4][CHT_X + 4];
    __shared__ double sl[CHT_Y + 2][CHT_X + 2];
    if (sc->ch_done) return;
    const int P = g.P, NXM = g.NXM, NYM = g.NYM;
    const bool lead = (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0);
    const double rr = 2.0 * block_sum(w.rr[n & 1], (n == 0) ? w.nt0 : w.nb, sh);   // start: one per tile
    double bnorm;
    if (n == 0) {
        bnorm = sqrt(2.0 * block_sum(w.bb, w.nt0, sh));
        if (bnorm < 1e-30) bnorm = 1.0;
        if (lead) sc->ch_bnorm = bnorm;
    } else {
        bnorm = sc->ch_bnorm;
    }
    const double rel = sqrt(rr) / bnorm;
    double beta = 0.0;
    bool stop = (rel < tol) || (n == 0 && maxit <= 0);
    if (!stop && n > 0) {
        beta = rr / (sc->ch_rzb[(n + 1) & 1] + 1e-30);
        stop = (n >= maxit);
    }
    if (lead) {
        sc->ch_rzb[n & 1] = rr;
        if (stop) {
            sc->ch_rel = rel;
            sc->ch_iter = n;
            sc->ch_done = (rel < tol || maxit <= 0) ? 1 : 3;
        }
    }
    if (stop) return;

    const double a_dt = 1.0 / (sc->dt * c_BET[stage]);
    const int t = xcd_band(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int ty = t / gridDim.x, tx = t - ty * gridDim.x;
    const int c0 = tx * CHT_X, j0 = ty * CHT_Y;
    for (int q = threadIdx.x; q < (CHT_Y + 4) * (CHT_X + 4); q += NTH) {
        const int a = q / (CHT_X + 4), b = q - a * (CHT_X + 4);
        const int ia = chx(c0 + b - 2, NXM), jb = chy(j0 + a - 2, NYM);
        sp[a][b] = (n == 0) ? AT(r, ia, jb) : AT(r, ia, jb) + beta * AT(pold, ia, jb);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < (CHT_Y + 2) * (CHT_X + 2); q += NTH) {
        const int a = q / (CHT_X + 2), b = q - a * (CHT_X + 2), jb = j0 + a - 1;
        if (jb >= 0 && jb < NYM)
            sl[a][b] = lap_vals(g, chx(c0 + b - 1, NXM), jb, sp[a + 1][b + 1], sp[a + 1][b + 2], sp[a + 1][b],
                                sp[a + 2][b + 1], sp[a][b + 1]);
    }
    __syncthreads();
    double acc = 0.0;
    for (int q = threadIdx.x; q < CHT_X * CHT_Y; q += NTH) {
        const int tr = q / CHT_X, tc = q - tr * CHT_X, i = c0 + tc, j = j0 + tr;
        if (i < NXM && j < NYM) {
            const double L0 = sl[tr + 1][tc + 1];
            const double LN = (j + 1 == NYM) ? L0 : sl[tr + 2][tc + 1];
            const double LS = (j == 0) ? L0 : sl[tr][tc + 1];
            const double pc = sp[tr + 2][tc + 2];
            const double ap = a_dt * pc + ph.bih * lap_vals(g, i, j, L0, sl[tr + 1][tc + 2], sl[tr + 1][tc], LN, LS);
            AT(pnew, i, j) = pc;
            AT(Ap, i, j) = ap;
            acc += ch_w(g, i, j, pc, ap);
        }
    }
    block_store_partial(acc, w.pap, t, sh);
}

// K2: alpha = <r,r> / <p,Ap> (VOF_DIFFUSE.c:384-390); x += alpha p;
// r -= alpha Ap (:390-405); partial <r, r> of the next iteration.
__global__ void __launch_bounds__(NTH) k_ccg_update(pvk_grid g, pvk_scal *sc, pvk_ccg w, double *__restrict__ x,
                                                   const double *__restrict__ p, double *__restrict__ r,
                                                   const double *__restrict__ Ap, int n)
{
    __shared__ double sh[NTH];
    if (sc->ch_done) return;
    const int P = g.P, NXM = g.NXM;
    const int band = xcd_band(blockIdx.x, gridDim.x);
    const double pAp = 2.0 * block_sum(w.pap, w.npap, sh);
    if (fabs(pAp) < 1e-30) {
        if (blockIdx.x == 0 && threadIdx.x == 0) { sc->ch_done = 2; sc->ch_iter = n; }
        return;
    }
    const double alpha = sc->ch_rzb[n & 1] / pAp;
    const int j0 = band * w.band, j1 = min(g.NYM, j0 + w.band);
    double acc = 0.0;
    for (int j = j0; j < j1; j++)
    for (int i = threadIdx.x; i < NXM; i += NTH) {
        AT(x, i, j) += alpha * AT(p, i, j);
        double rr = AT(r, i, j) - alpha * AT(Ap, i, j);
        AT(r, i, j) = rr;
        acc += ch_w(g, i, j, rr, rr);
    }
    block_store_partial(acc, w.rr[(n + 1) & 1], band, sh);
}

extern "C" void pvk_ccg_init(const pvk_grid *g, const pvk_phys *ph, int stage, pvk_scal *sc, const pvk_ccg *w,
                             const double *C, const double *u, const double *v, double *rn, const double *b,
                             double *r, double *p, double *bout, hipStream_t s)
{
    const dim3 gr((g->NXM + CIT_X - 1) / CIT_X, (g->NYM + CIT_Y - 1) / CIT_Y);
    if (b)
        hipLaunchKernelGGL(k_ccg_init_b, gr, dim3(NTH), 0, s, *g, *ph, stage, sc, C, b, r, p, *w);
    else if (ph->precision)
        hipLaunchKernelGGL(k_ccg_init_rhs<float>, gr, dim3(NTH), 0, s, *g, *ph, stage, sc, C, u, v, rn, r, p, *w,
                           bout);
    else
        hipLaunchKernelGGL(k_ccg_init_rhs<double>, gr, dim3(NTH), 0, s, *g, *ph, stage, sc, C, u, v, rn, r, p, *w,
                           bout);
}

extern "C" void pvk_ccg_iter(const pvk_grid *g, const pvk_phys *ph, int stage, pvk_scal *sc, const pvk_ccg *w,
                             int half, int n, double *x, double *r, double *Ap, hipStream_t s)
{
    if (half == 0)
        hipLaunchKernelGGL(k_ccg_apply, dim3((g->NXM + CHT_X - 1) / CHT_X, (g->NYM + CHT_Y - 1) / CHT_Y), dim3(NTH),
                           0, s, *g, *ph, stage, sc, *w, (const double *)r,
                           (const double *)w->p[(n + 1) & 1], w->p[n & 1], Ap, n, ph->ch_iter_max, ph->ch_tol);
    else
        hipLaunchKernelGGL(k_ccg_update, dim3(w->nb), dim3(NTH), 0, s, *g, sc, *w, x, (const double *)w->p[n & 1], r,
                           (const double *)Ap, n);
}

// ---------------------------------------------------------------------------
// Momentum (u and v solved in the same launches; blocks [0, nb[0]) take u)
// ---------------------------------------------------------------------------
// Unknown ranges (msolve_cg.c:525-722): u at i = 1..NXM, j = 0..NYM-1; v at
// i = 0..NXM-1, j = 1..NYM-1.  Ghost maps of the CG-type boundary update
// (Velocity.c:2427-2805): u is x periodic with column 0 = column NXM and
// ghost rows -1 / NYM holding minus rows 0 / NYM-1; v is x periodic (cell
// image) and its wall rows 0 and NYM are never written by the solve.
struct mom_band {
    int comp, band, j0, j1, i0, ncell;
};
__device__ __forceinline__ mom_band mom_locate(const pvk_grid &g, const pvk_mcg &w)
{
    mom_band m;
    const int vb = xcd_band(blockIdx.x, gridDim.x);
    m.comp = (vb < w.nb[0]) ? 0 : 1;
    m.band = m.comp ? vb - w.nb[0] : vb;
    const int jlo = m.comp ? 1 : 0, jhi = g.NYM;     // rows [jlo, NYM)
    m.j0 = jlo + m.band * w.band;
    m.j1 = min(jhi, m.j0 + w.band);
    m.i0 = m.comp ? 0 : 1;
    m.ncell = (m.j1 - m.j0) * g.NXM;
    return m;
}

// r = rhs - A x; dg = diag + EPS; d = r / dg; partial <r, d> (msolve_cg.c:627-645)
template <class R>
__global__ void __launch_bounds__(NTH) k_mcg_init(pvk_grid g, pvk_phys ph, int stage, pvk_scal *sc, pvk_mcg w)
{
    __shared__ double sh[NTH];
    const int P = g.P, NXM = g.NXM;
    const mom_band m = mom_locate(g, w);
    const int comp = m.comp ? 'v' : 'u';
    if (m.band == 0 && threadIdx.x == 0) { sc->m_done[m.comp] = (w.maxit <= 0); sc->m_iter[m.comp] = 0; }
    const double alpha_k = c_BET[stage], dt = sc->dt;
    const double *x = w.x[m.comp], *rhs = w.rhs[m.comp];
    double *r = w.r[m.comp], *dg = w.dg[m.comp], *d = w.d[m.comp][0], *Ad = w.Ad[m.comp];
    double acc = 0.0;
    for (int j = m.j0; j < m.j1; j++)
    for (int i = m.i0 + (int)threadIdx.x; i < m.i0 + NXM; i += NTH) {
        double ad = (double)mom_op<R>(g, ph, alpha_k, dt, w.rho, w.mu, x, comp, i, j);
        double rr = AT(rhs, i, j) - ad;
        double dgv = (double)mom_diag<R>(g, ph, alpha_k, dt, w.rho, w.mu, comp, i, j) + 1e-14;
        if constexpr (sizeof(R) == sizeof(float)) {     // the FP32 iterations read these instead of rho, mu
            float cc[5];
            mom_coeffs(g, ph, alpha_k, dt, w.rho, w.mu, m.comp, i, j, cc);
#pragma unroll
            for (int k = 0; k < 5; k++) AT(w.cf[m.comp][k], i, j) = cc[k];
        }
        double dd = rr / dgv;
        AT(Ad, i, j) = ad;
        AT(r, i, j) = rr;
        AT(dg, i, j) = dgv;
        AT(d, i, j) = dd;
        acc += rr * dd;
    }
    block_store_partial(acc, w.rz[m.comp][0], m.band, sh);
}

// K1: stop test sqrt(r.r / N) < CG_ETOL (msolve_cg.c:691-704), beta =
// rz_new / (rz + EPS) (w.exact: rz_new / rz, 0 when rz is 0),
// d_n = r / dg + beta d_{n-1} (:705-713) evaluated at
// the five operator points through the ghost maps; Ad = A d_n; partial <d, Ad>.
#define MCH 256                                                  // columns per chunk of k_mcg_apply
#define MSQ (((BAND + 2) * (MCH + 2) + NTH - 1) / NTH)     // staged values per thread
template <class R, int BAND, bool PC>
__global__ void __launch_bounds__(NTH) k_mcg_apply(pvk_grid g, pvk_phys ph, int stage, pvk_scal *sc, pvk_mcg w,
                                                  int n)
{
    __shared__ double sh[NTH];
    __shared__ double sd[BAND + 2][MCH + 2];
    const mom_band m = mom_locate(g, w);
    const int c = m.comp;
    if (sc->m_done[c]) return;
    const int P = g.P, NXM = g.NXM, NYM = g.NYM;
    const double EPS = 1e-14;
    const bool lead = (m.band == 0 && threadIdx.x == 0);
    const double rz = block_sum(w.rz[c][n & 1], w.nb[c], sh);
    double beta = 0.0;
    if (n > 0) {
        const double rr = block_sum(w.rr[c], w.nb[c], sh);
        const double rms = sqrt(w.in_tot[c] * rr);
        bool stop = rms < w.tol;
        if (!stop) {
            const double rzo = sc->m_rzb[c][(n + 1) & 1];
            beta = w.exact ? (rzo != 0.0 ? rz / rzo : 0.0) : rz / (rzo + EPS);
            stop = (n >= w.maxit);
        }
        if (stop) {
            if (lead) {
                sc->m_rms[c] = rms;
                sc->m_iter[c] = n;
                sc->m_done[c] = (rms < w.tol) ? 1 : 3;
            }
            return;
        }
    }
    if (lead) sc->m_rzb[c][n & 1] = rz;
    const double *r = w.r[c], *dg = w.dg[c], *dold = w.d[c][(n + 1) & 1], *dwall = w.d[c][0], *zc = w.z[c];
    double *dnew = w.d[c][n & 1], *Ad = w.Ad[c];
    const double alpha_k = c_BET[stage], dt = sc->dt;
    // d at (a, b) after the CG-type ghost fill of the new direction
    auto dv = [&](int a, int b) -> double {
        if (c == 0) {
            const int ia = wrapu(a, NXM);
            const int jb = b < 0 ? 0 : (b >= NYM ? NYM - 1 : b);
            double val = (n == 0) ? AT(dwall, ia, jb)
                                  : (PC ? AT(zc, ia, jb) : AT(r, ia, jb) / AT(dg, ia, jb)) + beta * AT(dold, ia, jb);
            return (b < 0 || b >= NYM) ? -val : val;
        }
        const int ia = wrapc(a, NXM);
        if (b <= 0 || b >= NYM) return AT(dwall, ia, b);
        return (n == 0) ? AT(dwall, ia, b)
                        : (PC ? AT(zc, ia, b) : AT(r, ia, b) / AT(dg, ia, b)) + beta * AT(dold, ia, b);
    };
    // The band (w.band = BAND rows, the last one may have fewer) in chunks of MCH columns: d of the
    // chunk and a one-cell ring is formed once into LDS (every value as dv
    // forms it), then A d of the chunk reads its neighbours there.
    const int nrow = m.j1 - m.j0;
    double acc = 0.0;
    for (int cb = 0; cb < NXM; cb += MCH) {
        double sv[MSQ];
#pragma unroll
        for (int k = 0; k < MSQ; k++) {               // all loads of the stage in flight together
            const int q = threadIdx.x + k * NTH, a = q / (MCH + 2), bb = q - a * (MCH + 2);
            sv[k] = (a < nrow + 2) ? dv(m.i0 + cb - 1 + bb, m.j0 - 1 + a) : 0.0;
        }
#pragma unroll
        for (int k = 0; k < MSQ; k++) {
            const int q = threadIdx.x + k * NTH, a = q / (MCH + 2), bb = q - a * (MCH + 2);
            if (a < BAND + 2) sd[a][bb] = sv[k];
        }
        __syncthreads();
        for (int q = threadIdx.x; q < nrow * MCH; q += NTH) {
            const int a = q / MCH, bb = q - a * MCH, j = m.j0 + a, i = m.i0 + cb + bb;
            if (cb + bb < NXM) {
                const double dc = sd[a + 1][bb + 1];
                double ad;
                if constexpr (sizeof(R) == sizeof(float)) {
                    float cc[5];
#pragma unroll
                    for (int k = 0; k < 5; k++) cc[k] = AT(w.cf[c][k], i, j);
                    ad = (double)mom_op_coeffs(cc, float(dc), float(sd[a + 1][bb + 2]), float(sd[a + 1][bb]),
                                               float(sd[a + 2][bb + 1]), float(sd[a][bb + 1]));
                } else {
                    ad = (double)mom_op_vals<R>(g, ph, alpha_k, dt, w.rho, w.mu, c, i, j, R(dc), R(sd[a + 1][bb + 2]),
                                                R(sd[a + 1][bb]), R(sd[a + 2][bb + 1]), R(sd[a][bb + 1]));
                }
                if (n > 0) AT(dnew, i, j) = dc;
                AT(Ad, i, j) = ad;
                acc += dc * ad;
            }
        }
        __syncthreads();
    }
    block_store_partial(acc, w.dad[c], m.band, sh);
}

// K2: alpha = rz / (dAd + EPS) (msolve_cg.c:674; w.exact: rz / dAd, 0 when
// dAd is 0); x += alpha d; r -= alpha Ad;
// partials of r^2 / dg and r^2 (:676-690).
__global__ void __launch_bounds__(NTH) k_mcg_update(pvk_grid g, pvk_scal *sc, pvk_mcg w, int n)
{
    __shared__ double sh[NTH];
    const mom_band m = mom_locate(g, w);
    const int c = m.comp;
    if (sc->m_done[c]) return;
    const int P = g.P, NXM = g.NXM;
    const double EPS = 1e-14;
    const double dAd = block_sum(w.dad[c], w.nb[c], sh);
    const double rzn = sc->m_rzb[c][n & 1];
    const double alpha = w.exact ? (dAd != 0.0 ? rzn / dAd : 0.0) : rzn / (dAd + EPS);
    if (m.band == 0 && threadIdx.x == 0) sc->m_iter[c] = n + 1;
    double *x = w.x[c], *r = w.r[c];
    const double *d = w.d[c][n & 1], *Ad = w.Ad[c], *dg = w.dg[c];
    double a0 = 0.0, a1 = 0.0;
    for (int j = m.j0; j < m.j1; j++)
    for (int i = m.i0 + (int)threadIdx.x; i < m.i0 + NXM; i += NTH) {
        AT(x, i, j) += alpha * AT(d, i, j);
        double ri = AT(r, i, j) - alpha * AT(Ad, i, j);
        AT(r, i, j) = ri;
        a0 += ri * ri / AT(dg, i, j);
        a1 += ri * ri;
    }
    block_store_partial(a0, w.rz[c][(n + 1) & 1], m.band, sh);
    block_store_partial(a1, w.rr[c], m.band, sh);
}

// Momentum preconditioner pc = 1, the first-order Neumann polynomial of
// the Jacobi splitting A = D - N: z = D^-1 (r + N D^-1 r), i.e.
// z_p = (r_p + sum_q c_pq r_q / D_q) / D_p over the unknown neighbours q of
// p (periodic in x, no coupling across the walls), c_pq the FP32 operator
// coefficients (E, W, N, S) the start kernel formed and D = dg.  On a
// uniform grid c_pq = c_qp, so M = D^-1 + D^-1 N D^-1 is symmetric, and it
// is positive definite where A is diagonally dominant (the spectral radius
// of D^-1/2 N D^-1/2 below one).  It takes about half the Jacobi-PCG's
// iterations for one more stencil of r per iteration, which the update
// kernel forms from r_n and A d_n at the neighbours (r is double-buffered:
// r_{n+1} goes to the other buffer while neighbours read r_n).
template <class RF>
__device__ __forceinline__ double mom_pc_z(const pvk_grid &g, const pvk_mcg &w, int c, int i, int j, RF rv)
{
    const int P = g.P, NXM = g.NXM, NYM = g.NYM;
    const double *dg = w.dg[c];
    const int ie = c ? wrapc(i + 1, NXM) : wrapu(i + 1, NXM);
    const int iw = c ? wrapc(i - 1, NXM) : wrapu(i - 1, NXM);
    const int jlo = c ? 1 : 0;
    double s = rv(i, j);
    s += (double)AT(w.cf[c][1], i, j) * (rv(ie, j) / AT(dg, ie, j));
    s += (double)AT(w.cf[c][2], i, j) * (rv(iw, j) / AT(dg, iw, j));
    if (j + 1 < NYM) s += (double)AT(w.cf[c][3], i, j) * (rv(i, j + 1) / AT(dg, i, j + 1));
    if (j - 1 >= jlo) s += (double)AT(w.cf[c][4], i, j) * (rv(i, j - 1) / AT(dg, i, j - 1));
    return s / AT(dg, i, j);
}

// pc = 1 start: d_0 = z_0 = M r_0, partial <r_0, z_0> (replacing the Jacobi
// ones of the start kernel)
__global__ void __launch_bounds__(NTH) k_mcg_pc_start(pvk_grid g, pvk_mcg w)
{
    __shared__ double sh[NTH];
    const int P = g.P, NXM = g.NXM;
    const mom_band m = mom_locate(g, w);
    const int c = m.comp;
    const double *r = w.r[c];
    double *d = w.d[c][0];
    auto rv = [&](int a, int b) { return AT(r, a, b); };
    double acc = 0.0;
    for (int j = m.j0; j < m.j1; j++)
    for (int i = m.i0 + (int)threadIdx.x; i < m.i0 + NXM; i += NTH) {
        const double zz = mom_pc_z(g, w, c, i, j, rv);
        AT(d, i, j) = zz;
        acc += AT(r, i, j) * zz;
    }
    block_store_partial(acc, w.rz[c][0], m.band, sh);
}

// pc = 1 update: alpha as k_mcg_update; x += alpha d; r_{n+1} = r_n - alpha
// A d_n into the other r buffer; z_{n+1} = M r_{n+1}; partials <r, z>, <r, r>
__global__ void __launch_bounds__(NTH) k_mcg_update_pc(pvk_grid g, pvk_scal *sc, pvk_mcg w, int n)
{
    __shared__ double sh[NTH];
    const mom_band m = mom_locate(g, w);
    const int c = m.comp;
    if (sc->m_done[c]) return;
    const int P = g.P, NXM = g.NXM;
    const double EPS = 1e-14;
    const double dAd = block_sum(w.dad[c], w.nb[c], sh);
    const double rzn = sc->m_rzb[c][n & 1];
    const double alpha = w.exact ? (dAd != 0.0 ? rzn / dAd : 0.0) : rzn / (dAd + EPS);
    if (m.band == 0 && threadIdx.x == 0) sc->m_iter[c] = n + 1;
    double *x = w.x[c], *z = w.z[c];
    const double *rold = (n & 1) ? w.r2[c] : w.r[c];
    double *rnew = (n & 1) ? w.r[c] : w.r2[c];
    const double *d = w.d[c][n & 1], *Ad = w.Ad[c];
    auto rv = [&](int a, int b) { return AT(rold, a, b) - alpha * AT(Ad, a, b); };
    double a0 = 0.0, a1 = 0.0;
    for (int j = m.j0; j < m.j1; j++)
    for (int i = m.i0 + (int)threadIdx.x; i < m.i0 + NXM; i += NTH) {
        AT(x, i, j) += alpha * AT(d, i, j);
        const double ri = rv(i, j);
        const double zz = mom_pc_z(g, w, c, i, j, rv);
        AT(rnew, i, j) = ri;
        AT(z, i, j) = zz;
        a0 += ri * zz;
        a1 += ri * ri;
    }
    block_store_partial(a0, w.rz[c][(n + 1) & 1], m.band, sh);
    block_store_partial(a1, w.rr[c], m.band, sh);
}

// pc = 1, after the loop: a component that stopped after an odd number of
// updates holds its residual in r2: copy it to r (the public field)
__global__ void __launch_bounds__(NTH) k_mcg_pc_finish(pvk_grid g, const pvk_scal *sc, pvk_mcg w)
{
    const int P = g.P, NXM = g.NXM;
    const mom_band m = mom_locate(g, w);
    const int c = m.comp;
    if (!(sc->m_iter[c] & 1)) return;
    for (int j = m.j0; j < m.j1; j++)
    for (int i = m.i0 + (int)threadIdx.x; i < m.i0 + NXM; i += NTH) AT(w.r[c], i, j) = AT(w.r2[c], i, j);
}

extern "C" void pvk_mcg_pc_start(const pvk_grid *g, const pvk_mcg *w, hipStream_t s)
{
    hipLaunchKernelGGL(k_mcg_pc_start, dim3(w->nb[0] + w->nb[1]), dim3(NTH), 0, s, *g, *w);
}

extern "C" void pvk_mcg_pc_finish(const pvk_grid *g, const pvk_scal *sc, const pvk_mcg *w, hipStream_t s)
{
    hipLaunchKernelGGL(k_mcg_pc_finish, dim3(w->nb[0] + w->nb[1]), dim3(NTH), 0, s, *g, sc, *w);
}

extern "C" void pvk_mcg_init(const pvk_grid *g, const pvk_phys *ph, int stage, pvk_scal *sc, const pvk_mcg *w,
                             hipStream_t s)
{
    if (ph->precision)
        hipLaunchKernelGGL(k_mcg_init<float>, dim3(w->nb[0] + w->nb[1]), dim3(NTH), 0, s, *g, *ph, stage, sc, *w);
    else
        hipLaunchKernelGGL(k_mcg_init<double>, dim3(w->nb[0] + w->nb[1]), dim3(NTH), 0, s, *g, *ph, stage, sc, *w);
}

extern "C" void pvk_mcg_iter(const pvk_grid *g, const pvk_phys *ph, int stage, pvk_scal *sc, const pvk_mcg *w,
                             int half, int n, hipStream_t s)
{
#define MCG_APPLY(RT, B, PC) \
    hipLaunchKernelGGL((k_mcg_apply<RT, B, PC>), dim3(w->nb[0] + w->nb[1]), dim3(NTH), 0, s, *g, *ph, stage, sc, *w, n)
    if (half == 0 && ph->precision && w->pc) {      // the polynomial preconditioner: FP32 form only
        if (w->band == 1) MCG_APPLY(float, 1, true);
        else if (w->band == 2) MCG_APPLY(float, 2, true);
        else MCG_APPLY(float, 4, true);
    } else if (half == 0 && ph->precision) {
        if (w->band == 1) MCG_APPLY(float, 1, false);
        else if (w->band == 2) MCG_APPLY(float, 2, false);
        else MCG_APPLY(float, 4, false);
    } else if (half == 0) {
        if (w->band == 1) MCG_APPLY(double, 1, false);
        else if (w->band == 2) MCG_APPLY(double, 2, false);
        else MCG_APPLY(double, 4, false);
    } else if (w->pc)
#undef MCG_APPLY
        hipLaunchKernelGGL(k_mcg_update_pc, dim3(w->nb[0] + w->nb[1]), dim3(NTH), 0, s, *g, sc, *w, n);
    else
        hipLaunchKernelGGL(k_mcg_update, dim3(w->nb[0] + w->nb[1]), dim3(NTH), 0, s, *g, sc, *w, n);
}
