// pv_mg.hip -- multigrid-preconditioned CG for the variable-density pressure
// equation: the same operator, start, gauge and stopping rule as the
// FFT-preconditioned CG (csrc/pv_pcg.hip, psolve_cg.c:255-279), with one
// multigrid V-cycle as the preconditioner in place of the constant-density
// FFT solve.  With a density ratio of 1e-3 the FFT preconditioner leaves a
// spectrum spread over three decades (M_fft A ~ 1 in the liquid, ~1e3 in the
// gas); the V-cycle, built on the variable face coefficients, does not.
//
// Operator, as k_pcg_apply: (A x)_ij = bx_i (x_E - x) - bx_{i-1} (x - x_W)
//   + by_j (x_N - x) - by_{j-1} (x - x_S), periodic in x, zero-gradient walls.
// V-cycle (cell-centred, 2 x 2 aggregation; symmetric, so CG stays CG):
//   smoother      red-black Gauss-Seidel, nu sweeps: (red, black) before the
//                 coarse correction, (black, red) after it;
//   restriction   sum of the four fine residuals;
//   prolongation  piecewise constant;
//   coarse faces  half the sum of the two fine faces across the coarse face;
//   levels of at most PVK_MG_LDS_CELLS cells run in one workgroup from LDS;
//   the last level (<= PVK_MG_BOTTOM cells) is solved with the inverse of its
//   operator grounded at cell 0, both sides projected to zero mean.
// A Gauss-Seidel value depends only on the neighbours, so the first red sweep
// starts from x = 0 without clearing x, and the coarse correction is added on
// the fly where the first black post-sweep reads the red neighbours.
#include <hip/hip_runtime.h>
#include <math.h>
#include "pv_kernels.h"
#include "pv_block.h"

#define AT(a, i, j) (a)[(ptrdiff_t)(j) * P + (i)]
#define TXP 64      // cell pairs per block row
#define TYR 4       // rows per block; TXP * TYR = NTH

__device__ __forceinline__ double ld(const double *a, int pitch, int i, int j) { return a[(ptrdiff_t)j * pitch + i]; }

// Gauss-Seidel value of cell (i, j): (sum_k c_k x_k - f) / sum_k c_k
__device__ __forceinline__ double gs_value(const pvk_mglev &L, int i, int j, double xe, double xw, double xn,
                                           double xs, double f)
{
    const int p = L.pitch, iw = (i == 0) ? L.nx - 1 : i - 1;
    const double ce = ld(L.bx, p, i, j), cw = ld(L.bx, p, iw, j);
    const double cn = (j + 1 < L.ny) ? ld(L.by, p, i, j) : 0.0;
    const double cs = (j > 0) ? ld(L.by, p, i, j - 1) : 0.0;
    return (ce * xe + cw * xw + cn * xn + cs * xs - f) / (ce + cw + cn + cs);
}

// (A x)(i, j) on a level
__device__ __forceinline__ double a_x(const pvk_mglev &L, int i, int j)
{
    const int p = L.pitch, ie = (i + 1 == L.nx) ? 0 : i + 1, iw = (i == 0) ? L.nx - 1 : i - 1;
    const double x0 = ld(L.x, p, i, j);
    double v = ld(L.bx, p, i, j) * (ld(L.x, p, ie, j) - x0) - ld(L.bx, p, iw, j) * (x0 - ld(L.x, p, iw, j));
    if (j + 1 < L.ny) v += ld(L.by, p, i, j) * (ld(L.x, p, i, j + 1) - x0);
    if (j > 0) v -= ld(L.by, p, i, j - 1) * (x0 - ld(L.x, p, i, j - 1));
    return v;
}

// stop test of CG iteration n on the <r, r> partials (as k_transpose):
// n = -1 sets the reference norm, n = -2 keeps it (warm restart)
__device__ void mg_stop_test(pvk_scal *sc, const pvk_pcg &w, int n, double tol, int maxit, double *sh)
{
    const double rr = block_sum(w.rr, w.nrr, sh);
    const double bn = (n == -1) ? ((sqrt(rr) == 0.0) ? 1.0 : sqrt(rr)) : sc->p_bnorm;
    const double rel = sqrt(rr) / bn;
    const bool done = (rel <= tol) || (n + 1 >= maxit && n >= 0);
    if (threadIdx.x == 0) {
        if (n == -1) sc->p_bnorm = bn;
        sc->p_res = rel;
        sc->p_iter = (n < 0) ? 0 : n + 1;
        if (done) sc->p_done = (rel <= tol) ? 1 : 3;
    }
}

// One half-sweep on the cells of colour col ((i + j) & 1 == col), one thread
// per horizontal cell pair.  MODE 0: first red sweep from x = 0; MODE 1:
// Gauss-Seidel; MODE 2: first black sweep after the coarse correction (red
// neighbours read as x + xc(parent)).  check: block (0, 0) runs the stop test
// of iteration n first (the other blocks go on; later kernels see the flag).
// dot: partial <f, x> of the level after the sweep into w.rz[slot].
template <int MODE>
__global__ void __launch_bounds__(NTH) k_mg_smooth(pvk_mglev L, pvk_mglev C, int col, pvk_scal *sc, pvk_pcg w,
                                                  int n, double tol, int maxit, int check, int dot, int slot)
{
    __shared__ double sh[NTH];
    if (sc->p_done) return;
    if (check && n != -3 && blockIdx.x == 0 && blockIdx.y == 0) mg_stop_test(sc, w, n, tol, maxit, sh);
    const int ip = blockIdx.x * TXP + (threadIdx.x & (TXP - 1));
    const int j = blockIdx.y * TYR + (threadIdx.x >> 6);
    const int p = L.pitch;
    double acc = 0.0;
    if (2 * ip < L.nx && j < L.ny) {
        const int i = 2 * ip + ((j + col) & 1);
        const double f = ld(L.f, p, i, j);
        double v;
        if (MODE == 0) {
            v = gs_value(L, i, j, 0.0, 0.0, 0.0, 0.0, f);
        } else {
            const int ie = (i + 1 == L.nx) ? 0 : i + 1, iw = (i == 0) ? L.nx - 1 : i - 1;
            double xe = ld(L.x, p, ie, j), xw = ld(L.x, p, iw, j);
            double xn = (j + 1 < L.ny) ? ld(L.x, p, i, j + 1) : 0.0;
            double xs = (j > 0) ? ld(L.x, p, i, j - 1) : 0.0;
            if (MODE == 2) {
                const int cp = C.pitch;
                xe += ld(C.x, cp, ie >> 1, j >> 1);
                xw += ld(C.x, cp, iw >> 1, j >> 1);
                if (j + 1 < L.ny) xn += ld(C.x, cp, i >> 1, (j + 1) >> 1);
                if (j > 0) xs += ld(C.x, cp, i >> 1, (j - 1) >> 1);
            }
            v = gs_value(L, i, j, xe, xw, xn, xs, f);
        }
        L.x[(ptrdiff_t)j * p + i] = v;
        if (dot) acc = f * v + ld(L.f, p, i ^ 1, j) * ld(L.x, p, i ^ 1, j);
    }
    if (dot) block_store_partial(acc, w.rz[slot], blockIdx.y * gridDim.x + blockIdx.x, sh);
}

// coarse right-hand side: sum of the four fine residuals f - A x
__global__ void __launch_bounds__(NTH) k_mg_restrict(pvk_mglev L, pvk_mglev C, const pvk_scal *sc)
{
    if (sc->p_done) return;
    const int I = blockIdx.x * TXP + (threadIdx.x & (TXP - 1)), J = blockIdx.y * TYR + (threadIdx.x >> 6);
    if (I >= C.nx || J >= C.ny) return;
    double s = 0.0;
#pragma unroll
    for (int dj = 0; dj < 2; dj++)
#pragma unroll
        for (int di = 0; di < 2; di++) {
            const int i = 2 * I + di, j = 2 * J + dj;
            s += ld(L.f, L.pitch, i, j) - a_x(L, i, j);
        }
    C.f[(ptrdiff_t)J * C.pitch + I] = s;
}

// coarse face coefficients from the fine ones
__global__ void __launch_bounds__(NTH) k_mg_faces(pvk_mglev L, pvk_mglev C)
{
    const int I = blockIdx.x * TXP + (threadIdx.x & (TXP - 1)), J = blockIdx.y * TYR + (threadIdx.x >> 6);
    if (I >= C.nx || J >= C.ny) return;
    const int p = L.pitch;
    C.bx[(ptrdiff_t)J * C.pitch + I] = 0.5 * (ld(L.bx, p, 2 * I + 1, 2 * J) + ld(L.bx, p, 2 * I + 1, 2 * J + 1));
    C.by[(ptrdiff_t)J * C.pitch + I] = 0.5 * (ld(L.by, p, 2 * I, 2 * J + 1) + ld(L.by, p, 2 * I + 1, 2 * J + 1));
}

// inverse of the bottom operator -A_B + s e0 e0^T (s = its (0,0) entry):
// symmetric positive definite, in-place Gauss-Jordan without pivoting
__global__ void __launch_bounds__(NTH) k_mg_bottom(pvk_mglev B, double *G)
{
    __shared__ double M[PVK_MG_BOTTOM][PVK_MG_BOTTOM + 1];
    __shared__ double fac[PVK_MG_BOTTOM];
    __shared__ double piv;
    const int nx = B.nx, ny = B.ny, nb = nx * ny, p = B.pitch;
    for (int e = threadIdx.x; e < nb * nb; e += NTH) M[e / nb][e % nb] = 0.0;
    __syncthreads();
    if ((int)threadIdx.x < nb) {
        const int c = threadIdx.x, j = c / nx, i = c - j * nx;
        const int ie = (i + 1 == nx) ? 0 : i + 1, iw = (i == 0) ? nx - 1 : i - 1;
        const double ce = ld(B.bx, p, i, j), cw = ld(B.bx, p, iw, j);
        const double cn = (j + 1 < ny) ? ld(B.by, p, i, j) : 0.0, cs = (j > 0) ? ld(B.by, p, i, j - 1) : 0.0;
        M[c][c] += ce + cw + cn + cs;
        M[c][j * nx + ie] -= ce;
        M[c][j * nx + iw] -= cw;
        if (j + 1 < ny) M[c][(j + 1) * nx + i] -= cn;
        if (j > 0) M[c][(j - 1) * nx + i] -= cs;
    }
    __syncthreads();
    if (threadIdx.x == 0) M[0][0] += M[0][0];
    __syncthreads();
    for (int k = 0; k < nb; k++) {
        if (threadIdx.x == 0) piv = 1.0 / M[k][k];
        __syncthreads();
        for (int c = threadIdx.x; c < nb; c += NTH) M[k][c] = ((c == k) ? 1.0 : M[k][c]) * piv;
        __syncthreads();
        for (int r = threadIdx.x; r < nb; r += NTH) fac[r] = (r == k) ? 0.0 : M[r][k];
        __syncthreads();
        for (int e = threadIdx.x; e < nb * nb; e += NTH) {
            const int r = e / nb, c = e - r * nb;
            if (r != k) M[r][c] = ((c == k) ? 0.0 : M[r][c]) - fac[r] * M[k][c];
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < nb * nb; e += NTH) G[e] = M[e / nb][e % nb];
}

// ---------------------------------------------------------------------------
// Levels c0 .. nlev-1 in one workgroup of 1024 threads, from LDS.
// ---------------------------------------------------------------------------
#define CT 1024

struct LdsLev {
    int nx, ny;
    double *x, *f, *bx, *by;     // compact (pitch nx), in LDS
};

__device__ __forceinline__ double lgs(const LdsLev &L, int i, int j, double xe, double xw, double xn, double xs)
{
    const int nx = L.nx, iw = (i == 0) ? nx - 1 : i - 1;
    const double ce = L.bx[j * nx + i], cw = L.bx[j * nx + iw];
    const double cn = (j + 1 < L.ny) ? L.by[j * nx + i] : 0.0;
    const double cs = (j > 0) ? L.by[(j - 1) * nx + i] : 0.0;
    return (ce * xe + cw * xw + cn * xn + cs * xs - L.f[j * nx + i]) / (ce + cw + cn + cs);
}

// half-sweep of colour col; mode 0: from x = 0, 1: Gauss-Seidel, 2: first
// post-sweep with the red neighbours read as x + xc(parent)
__device__ void lds_sweep(const LdsLev &L, const LdsLev &C, int col, int mode)
{
    const int hx = L.nx >> 1, npair = hx * L.ny, nx = L.nx;
    for (int q = threadIdx.x; q < npair; q += CT) {
        const int j = q / hx, i = 2 * (q - j * hx) + ((j + col) & 1);
        double v;
        if (mode == 0) {
            v = lgs(L, i, j, 0.0, 0.0, 0.0, 0.0);
        } else {
            const int ie = (i + 1 == nx) ? 0 : i + 1, iw = (i == 0) ? nx - 1 : i - 1;
            double xe = L.x[j * nx + ie], xw = L.x[j * nx + iw];
            double xn = (j + 1 < L.ny) ? L.x[(j + 1) * nx + i] : 0.0;
            double xs = (j > 0) ? L.x[(j - 1) * nx + i] : 0.0;
            if (mode == 2) {
                const int cx = C.nx;
                xe += C.x[(j >> 1) * cx + (ie >> 1)];
                xw += C.x[(j >> 1) * cx + (iw >> 1)];
                if (j + 1 < L.ny) xn += C.x[((j + 1) >> 1) * cx + (i >> 1)];
                if (j > 0) xs += C.x[((j - 1) >> 1) * cx + (i >> 1)];
            }
            v = lgs(L, i, j, xe, xw, xn, xs);
        }
        L.x[j * nx + i] = v;
    }
    __syncthreads();
}

__device__ __forceinline__ double lds_ax(const LdsLev &L, int i, int j)
{
    const int nx = L.nx, ie = (i + 1 == nx) ? 0 : i + 1, iw = (i == 0) ? nx - 1 : i - 1;
    const double x0 = L.x[j * nx + i];
    double v = L.bx[j * nx + i] * (L.x[j * nx + ie] - x0) - L.bx[j * nx + iw] * (x0 - L.x[j * nx + iw]);
    if (j + 1 < L.ny) v += L.by[j * nx + i] * (L.x[(j + 1) * nx + i] - x0);
    if (j > 0) v -= L.by[(j - 1) * nx + i] * (x0 - L.x[(j - 1) * nx + i]);
    return v;
}

__device__ void lds_restrict(const LdsLev &L, const LdsLev &C)
{
    for (int q = threadIdx.x; q < C.nx * C.ny; q += CT) {
        const int J = q / C.nx, I = q - J * C.nx;
        double s = 0.0;
        for (int dj = 0; dj < 2; dj++)
            for (int di = 0; di < 2; di++) {
                const int i = 2 * I + di, j = 2 * J + dj;
                s += L.f[j * L.nx + i] - lds_ax(L, i, j);
            }
        C.f[q] = s;
    }
    __syncthreads();
}

__global__ void __launch_bounds__(CT) k_mg_coarse(pvk_mg m, pvk_scal *sc)
{
    extern __shared__ double lds[];
    __shared__ double red[CT / 64];
    __shared__ double mean;
    if (sc->p_done) return;
    LdsLev lv[PVK_MG_MAXLEV];
    {
        double *q = lds;
        for (int l = m.c0; l < m.nlev; l++) {
            const int nc = m.L[l].nx * m.L[l].ny;
            lv[l].nx = m.L[l].nx; lv[l].ny = m.L[l].ny;
            lv[l].x = q; lv[l].f = q + nc; lv[l].bx = q + 2 * nc; lv[l].by = q + 3 * nc;
            q += 4 * nc;
        }
    }
    // faces of every in-LDS level and the right-hand side of the first
    for (int l = m.c0; l < m.nlev; l++) {
        const pvk_mglev &G = m.L[l];
        const int hx = G.nx;
        for (int q = threadIdx.x; q < G.nx * G.ny; q += CT) {
            const int j = q / hx, i = q - j * hx;
            lv[l].bx[q] = ld(G.bx, G.pitch, i, j);
            lv[l].by[q] = ld(G.by, G.pitch, i, j);
            if (l == m.c0) lv[l].f[q] = ld(G.f, G.pitch, i, j);
        }
    }
    __syncthreads();
    const int nb = m.nlev - 1;
    for (int l = m.c0; l < nb; l++) {
        lds_sweep(lv[l], lv[l], 0, 0);
        lds_sweep(lv[l], lv[l], 1, 1);
        for (int s = 1; s < m.nu; s++) {
            lds_sweep(lv[l], lv[l], 0, 1);
            lds_sweep(lv[l], lv[l], 1, 1);
        }
        lds_restrict(lv[l], lv[l + 1]);
    }
    // bottom: x = -P G P f, P the projection to zero mean
    {
        const LdsLev &B = lv[nb];
        const int n = B.nx * B.ny;
        if (threadIdx.x == 0) {
            double s = 0.0;
            for (int k = 0; k < n; k++) s += B.f[k];
            mean = s / n;
        }
        __syncthreads();
        double v = 0.0;
        if ((int)threadIdx.x < n) {
            const double *g = m.G + (size_t)threadIdx.x * n;
            for (int k = 0; k < n; k++) v -= g[k] * (B.f[k] - mean);
        }
        __syncthreads();
        if ((int)threadIdx.x < n) B.x[threadIdx.x] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            double s = 0.0;
            for (int k = 0; k < n; k++) s += B.x[k];
            mean = s / n;
        }
        __syncthreads();
        if ((int)threadIdx.x < n) B.x[threadIdx.x] -= mean;
        __syncthreads();
    }
    for (int l = nb - 1; l >= m.c0; l--) {
        lds_sweep(lv[l], lv[l + 1], 1, 2);
        lds_sweep(lv[l], lv[l], 0, 1);
        for (int s = 1; s < m.nu; s++) {
            lds_sweep(lv[l], lv[l], 1, 1);
            lds_sweep(lv[l], lv[l], 0, 1);
        }
    }
    {
        const pvk_mglev &G = m.L[m.c0];
        for (int q = threadIdx.x; q < G.nx * G.ny; q += CT) {
            const int j = q / G.nx, i = q - j * G.nx;
            G.x[(ptrdiff_t)j * G.pitch + i] = lv[m.c0].x[q];
        }
    }
    (void)red;
}

// stop test alone, and <r, z> alone (grids small enough for one workgroup)
__global__ void __launch_bounds__(NTH) k_mg_check(pvk_scal *sc, pvk_pcg w, int n, double tol, int maxit)
{
    __shared__ double sh[NTH];
    if (sc->p_done || n == -3) return;
    mg_stop_test(sc, w, n, tol, maxit, sh);
}
__global__ void __launch_bounds__(NTH) k_mg_dot(pvk_mglev L, const pvk_scal *sc, pvk_pcg w, int slot)
{
    __shared__ double sh[NTH];
    if (sc->p_done) return;
    double acc = 0.0;
    for (int q = threadIdx.x; q < L.nx * L.ny; q += NTH) {
        const int j = q / L.nx, i = q - j * L.nx;
        acc += ld(L.f, L.pitch, i, j) * ld(L.x, L.pitch, i, j);
    }
    block_store_partial(acc, w.rz[slot], 0, sh);
}

// ---------------------------------------------------------------------------
// CG vector kernels, one block per row
// ---------------------------------------------------------------------------
// start: r = rhs - mean (phi = 0), or with warm != 0 r = rhs - mean - Ad
// (Ad = A phi0, phi kept); partial <r, r>
__global__ void __launch_bounds__(NTH) k_mg_start(pvk_grid g, const pvk_scal *sc, pvk_pcg w, int warm)
{
    __shared__ double sh[NTH];
    const int P = g.P, j = blockIdx.x;
    const double mean = block_sum(w.mean, w.nmean, sh) / ((double)g.NXM * (double)g.NYM);
    double acc = 0.0;
    for (int i = threadIdx.x; i < g.NXM; i += NTH) {
        double r = AT(w.rhs, i, j) - mean;
        if (warm) r -= AT(w.Ad, i, j);
        else AT(w.phi, i, j) = 0.0;
        AT(w.r, i, j) = r;
        acc += r * r;
    }
    block_store_partial(acc, w.rr, j, sh);
    (void)sc;
}

// alpha = <r, z> / <d, A d>; phi += alpha d; r -= alpha A d; partial <r, r>
__global__ void __launch_bounds__(NTH) k_mg_update(pvk_grid g, const pvk_scal *sc, pvk_pcg w, int n)
{
    __shared__ double sh[NTH];
    if (sc->p_done) return;
    const int P = g.P, j = blockIdx.x;
    const double alpha = block_sum(w.rz[n & 1], w.nrz, sh) / block_sum(w.dad[n & 1], w.ndad, sh);
    const double *d = w.d[n & 1];
    double acc = 0.0;
    for (int i = threadIdx.x; i < g.NXM; i += NTH) {
        AT(w.phi, i, j) += alpha * AT(d, i, j);
        const double r = AT(w.r, i, j) - alpha * AT(w.Ad, i, j);
        AT(w.r, i, j) = r;
        acc += r * r;
    }
    block_store_partial(acc, w.rr, j, sh);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static dim3 pair_grid(const pvk_mglev &L) { return dim3((L.nx / 2 + TXP - 1) / TXP, (L.ny + TYR - 1) / TYR); }
static dim3 cell_grid(const pvk_mglev &L) { return dim3((L.nx + TXP - 1) / TXP, (L.ny + TYR - 1) / TYR); }
static size_t coarse_lds(const pvk_mg *m) { return (size_t)m->ldsx * sizeof(double); }

extern "C" int pvk_mg_nrz(const pvk_mg *m)
{
    if (m->c0 == 0) return 1;
    const dim3 gr = pair_grid(m->L[0]);
    return (int)(gr.x * gr.y);
}

// per solve: coarse faces of every level and the bottom inverse
extern "C" void pvk_mg_setup(const pvk_mg *m, hipStream_t s)
{
    for (int l = 0; l + 1 < m->nlev; l++)
        hipLaunchKernelGGL(k_mg_faces, cell_grid(m->L[l + 1]), dim3(NTH), 0, s, m->L[l], m->L[l + 1]);
    hipLaunchKernelGGL(k_mg_bottom, dim3(1), dim3(NTH), 0, s, m->L[m->nlev - 1], m->G);
}

// z = V(r) for CG iteration n (n < 0: the start; n = -3: no stop test),
// with the stop test of n first and <r, z> partials into w->rz[(n < 0) ? 0 : (n + 1) & 1]
extern "C" void pvk_mg_vcycle(const pvk_mg *m, pvk_scal *sc, const pvk_pcg *w, int n, double tol, int maxit,
                              hipStream_t s)
{
    const int slot = (n < 0) ? 0 : ((n + 1) & 1);
    if (m->c0 == 0) {
        hipLaunchKernelGGL(k_mg_check, dim3(1), dim3(NTH), 0, s, sc, *w, n, tol, maxit);
        hipLaunchKernelGGL(k_mg_coarse, dim3(1), dim3(CT), coarse_lds(m), s, *m, sc);
        hipLaunchKernelGGL(k_mg_dot, dim3(1), dim3(NTH), 0, s, m->L[0], sc, *w, slot);
        return;
    }
    for (int l = 0; l < m->c0; l++) {
        const pvk_mglev &L = m->L[l], &C = m->L[l + 1];
        const dim3 gr = pair_grid(L);
        hipLaunchKernelGGL(k_mg_smooth<0>, gr, dim3(NTH), 0, s, L, C, 0, sc, *w, n, tol, maxit, l == 0, 0, 0);
        hipLaunchKernelGGL(k_mg_smooth<1>, gr, dim3(NTH), 0, s, L, C, 1, sc, *w, n, tol, maxit, 0, 0, 0);
        for (int k = 1; k < m->nu; k++) {
            hipLaunchKernelGGL(k_mg_smooth<1>, gr, dim3(NTH), 0, s, L, C, 0, sc, *w, n, tol, maxit, 0, 0, 0);
            hipLaunchKernelGGL(k_mg_smooth<1>, gr, dim3(NTH), 0, s, L, C, 1, sc, *w, n, tol, maxit, 0, 0, 0);
        }
        hipLaunchKernelGGL(k_mg_restrict, cell_grid(C), dim3(NTH), 0, s, L, C, (const pvk_scal *)sc);
    }
    hipLaunchKernelGGL(k_mg_coarse, dim3(1), dim3(CT), coarse_lds(m), s, *m, sc);
    for (int l = m->c0 - 1; l >= 0; l--) {
        const pvk_mglev &L = m->L[l], &C = m->L[l + 1];
        const dim3 gr = pair_grid(L);
        const int last = (m->nu == 1) && l == 0;
        hipLaunchKernelGGL(k_mg_smooth<2>, gr, dim3(NTH), 0, s, L, C, 1, sc, *w, n, tol, maxit, 0, 0, 0);
        hipLaunchKernelGGL(k_mg_smooth<1>, gr, dim3(NTH), 0, s, L, C, 0, sc, *w, n, tol, maxit, 0, last, slot);
        for (int k = 1; k < m->nu; k++) {
            const int dl = (k + 1 == m->nu) && l == 0;
            hipLaunchKernelGGL(k_mg_smooth<1>, gr, dim3(NTH), 0, s, L, C, 1, sc, *w, n, tol, maxit, 0, 0, 0);
            hipLaunchKernelGGL(k_mg_smooth<1>, gr, dim3(NTH), 0, s, L, C, 0, sc, *w, n, tol, maxit, 0, dl, slot);
        }
    }
}

extern "C" void pvk_mg_start(const pvk_grid *g, pvk_scal *sc, const pvk_pcg *w, int warm, hipStream_t s)
{
    hipLaunchKernelGGL(k_mg_start, dim3(g->NYM), dim3(NTH), 0, s, *g, (const pvk_scal *)sc, *w, warm);
}

extern "C" void pvk_mg_update(const pvk_grid *g, pvk_scal *sc, const pvk_pcg *w, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_mg_update, dim3(g->NYM), dim3(NTH), 0, s, *g, (const pvk_scal *)sc, *w, n);
}
