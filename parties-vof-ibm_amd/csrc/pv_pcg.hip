// pv_pcg.hip -- fused FFT-preconditioned PCG for the variable-coefficient
// pressure equation div(beta grad phi) = rhs, beta = 2 / (rho_i + rho_nb)
// (operator of lsolver/psolve_cg.c:302-344), preconditioned by the
// constant-coefficient FFT-x / DCT-y solve.
//
// One iteration = six launches and no one-block kernels:
//   K1 k_pcg_apply      d_n = z + beta d_{n-1} on the fly (double-buffered d),
//                       Ad = A d_n, partials <d, Ad>           (row bands)
//   K2 k_pcg_update_fwd alpha; phi += alpha d; r -= alpha Ad; partials <r,r>;
//                       forward real FFT of the new r rows    (R rows per block)
//   K3 k_transpose      stop test from <r,r>; spec -> specT   (32 x 32 tiles)
//   K4 k_dct_rows       DCT-II, divide by lambda, DCT-III along y (rows of specT)
//   K5 k_transpose      specT -> spec
//   K6 k_ifft_rows_dot  z = inverse real FFT; partials <r, z>
// Scalars are never stored by one block and read by others in the same
// kernel: every block re-adds the few hundred partials of the previous
// kernel in the same fixed order, so all blocks see bit-identical alpha,
// beta and stop decisions and the result does not depend on scheduling.
// <r, z> partials alternate between two buffers by iteration parity so beta
// can be formed from gamma_n and gamma_{n-1}.
#include <hip/hip_runtime.h>
#include <math.h>
#include "pv_kernels.h"
#include "pv_block.h"
#include "pv_fftk.h"

#define AT(a, i, j) (a)[(ptrdiff_t)(j) * P + (i)]

struct cplx2 { double x, y; };

// threads per transform of length M in the Stockham FFT (pv_fftk.h)
__device__ __forceinline__ int sk_threads(int M) { return M >= 8 ? M / 8 : 1; }

// X_k, k = 0..M, of a real length-2M sequence from Z = FFT_M(x_{2m} + i x_{2m+1})
__device__ __forceinline__ cplx2 rpost(const double *sr, const double *si, int base, int M, int k, const double *rt)
{
    int ka = (k == M) ? 0 : k, kb = (M - k) & (M - 1);
    double zr = sr[base + ka], zi = si[base + ka];
    double cr = sr[base + kb], ci = -si[base + kb];
    double er = 0.5 * (zr + cr), ei = 0.5 * (zi + ci);
    double dr = zr - cr, di = zi - ci;
    double or_ = 0.5 * di, oi = -0.5 * dr;
    double c = rt[2 * k], s = rt[2 * k + 1];
    cplx2 X;
    X.x = er + (c * or_ + s * oi);
    X.y = ei + (c * oi - s * or_);
    return X;
}

// Y_k from X_k and X_{M-k} for the half-length inverse
__device__ __forceinline__ cplx2 rpre(cplx2 Xk, cplx2 Xmk, int k, const double *rt)
{
    double ar = Xk.x + Xmk.x, ai = Xk.y - Xmk.y;
    double dr = Xk.x - Xmk.x, di = Xk.y + Xmk.y;
    double c = rt[2 * k], s = rt[2 * k + 1];
    double pr = dr * c - di * s, pi_ = dr * s + di * c;
    cplx2 Y;
    Y.x = ar - pi_;
    Y.y = ai + pr;
    return Y;
}

__device__ __forceinline__ cplx2 hcget(const double *row, int N, int k)
{
    cplx2 X;
    X.x = row[k];
    X.y = (k == 0 || k == N / 2) ? 0.0 : row[N - k];
    return X;
}

// ---------------------------------------------------------------------------
// K1: d_n = z + beta d_{n-1}, Ad = A d_n, partial <d_n, Ad>
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(NTH) k_pcg_apply(pvk_grid g, pvk_scal *sc, pvk_pcg w, int n)
{
    __shared__ double sh[NTH];
    if (sc->p_done) return;
    const int P = g.P, NXM = g.NXM, NYM = g.NYM;
    const double gam = block_sum(w.rz[n & 1], w.nrz, sh);
    const double beta = (n == 0) ? 0.0 : gam / block_sum(w.rz[(n + 1) & 1], w.nrz, sh);
    const double *dold = w.d[(n + 1) & 1];
    double *dnew = w.d[n & 1];
    const double idx2 = g.idx_c[1] * g.idx_c[1], idy2 = g.idy_c[1] * g.idy_c[1];
    const int j0 = blockIdx.x * w.band, j1 = min(NYM, j0 + w.band);
    double acc = 0.0;
    for (int c = threadIdx.x; c < (j1 - j0) * NXM; c += NTH) {
        int j = j0 + c / NXM, i = c - (c / NXM) * NXM;
        int ie = (i + 1 == NXM) ? 0 : i + 1, iw = (i == 0) ? NXM - 1 : i - 1;
        int jn = (j + 1 == NYM) ? j : j + 1, js = (j == 0) ? 0 : j - 1;
#define DN(ii, jj) (n == 0 ? AT(w.z, ii, jj) : AT(w.z, ii, jj) + beta * AT(dold, ii, jj))
        double d0 = DN(i, j), dE = DN(ie, j), dW = DN(iw, j), dN = DN(i, jn), dS = DN(i, js);
#undef DN
        double r0 = AT(w.rho, i, j);
        double cxp = (2.0 / (r0 + AT(w.rho, i + 1, j))) * idx2;
        double cxm = (2.0 / (r0 + AT(w.rho, i - 1, j))) * idx2;
        double cyp = (2.0 / (r0 + AT(w.rho, i, j + 1))) * idy2;
        double cym = (2.0 / (r0 + AT(w.rho, i, j - 1))) * idy2;
        double ad = cxp * (dE - d0) - cxm * (d0 - dW) + cyp * (dN - d0) - cym * (d0 - dS);
        AT(dnew, i, j) = d0;
        AT(w.Ad, i, j) = ad;
        acc += d0 * ad;
    }
    block_store_partial(acc, w.dad, blockIdx.x, sh);
}

// ---------------------------------------------------------------------------
// K2: alpha; phi, r update; <r,r>; forward real FFT of the r rows -> spec
// (also the start kernel when n < 0: r = rhs - mean, phi = 0)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(NTH) k_pcg_update_fwd(pvk_grid g, pvk_fft F, pvk_scal *sc, pvk_pcg w, int n)
{
    extern __shared__ double lds[];
    __shared__ double sh[NTH];
    if (sc->p_done) return;
    const int P = g.P, N = g.NXM, M = N / 2, logM = F.logmx, R = w.rrow;
    double *sr = lds, *si = lds + R * M;
    double alpha = 0.0, mean = 0.0;
    if (n >= 0) alpha = block_sum(w.rz[n & 1], w.nrz, sh) / block_sum(w.dad, w.ndad, sh);
    else mean = block_sum(w.mean, w.nmean, sh) / ((double)g.NXM * (double)g.NYM);
    const double *d = (n >= 0) ? w.d[n & 1] : (const double *)0;
    const int j0 = blockIdx.x * R;
    double acc = 0.0;
    for (int t = threadIdx.x; t < R * M; t += NTH) {
        int q = t / M, m = t - q * M, j = j0 + q;
        double a = 0.0, b = 0.0;
        if (j < g.NYM) {
            int i = 2 * m;
            if (n >= 0) {
                AT(w.phi, i, j) += alpha * AT(d, i, j);
                AT(w.phi, i + 1, j) += alpha * AT(d, i + 1, j);
                a = AT(w.r, i, j) - alpha * AT(w.Ad, i, j);
                b = AT(w.r, i + 1, j) - alpha * AT(w.Ad, i + 1, j);
            } else {
                AT(w.phi, i, j) = 0.0;
                AT(w.phi, i + 1, j) = 0.0;
                a = AT(w.rhs, i, j) - mean;
                b = AT(w.rhs, i + 1, j) - mean;
            }
            AT(w.r, i, j) = a;
            AT(w.r, i + 1, j) = b;
            acc += a * a + b * b;
        }
        sr[q * M + m] = a; si[q * M + m] = b;
    }
    block_store_partial(acc, w.rr, blockIdx.x, sh);
    {
        const int T = sk_threads(M), q = threadIdx.x / T;
        sk_fft<false>(sr + q * M, si + q * M, M, logM, F.wx, threadIdx.x - q * T, T);
    }
    const double scale = 1.0 / (double)N;
    for (int t = threadIdx.x; t < R * (M + 1); t += NTH) {
        int q = t / (M + 1), k = t - q * (M + 1), j = j0 + q;
        if (j >= g.NYM) continue;
        cplx2 X = rpost(sr, si, q * M, M, k, F.rtx);
        double *row = F.spec + (size_t)j * N;
        row[k] = X.x * scale;
        if (k > 0 && k < M) row[N - k] = X.y * scale;
    }
}

// ---------------------------------------------------------------------------
// K3 / K5: 32 x 32 tiled transpose of an (rows x cols) row-major array;
// with check != 0 the stop test of iteration n runs first (all blocks).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(NTH) k_transpose(pvk_scal *sc, pvk_pcg w, const double *in, double *out, int rows,
                                                  int cols, int check, int n, double tol, int maxit)
{
    __shared__ double tile[32][33];
    __shared__ double sh[NTH];
    // a stop flag set by an earlier iteration leaves <r,r> stale: no recount
    if (sc->p_done) return;
    if (check) {
        double rr = block_sum(w.rr, w.nrr, sh);
        double bn = (n < 0) ? ((sqrt(rr) == 0.0) ? 1.0 : sqrt(rr)) : sc->p_bnorm;
        double rel = sqrt(rr) / bn;
        bool done = (rel <= tol) || (n + 1 >= maxit && n >= 0);
        if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
            if (n < 0) sc->p_bnorm = bn;
            sc->p_res = rel;
            sc->p_iter = n + 1;
            if (done) sc->p_done = (rel <= tol) ? 1 : 3;
        }
        if (done) return;
    }
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    for (int k = 0; k < 32; k += 8) {
        int r = r0 + ty + k, c = c0 + tx;
        if (r < rows && c < cols) tile[ty + k][tx] = in[(size_t)r * cols + c];
    }
    __syncthreads();
    for (int k = 0; k < 32; k += 8) {
        int r = c0 + ty + k, c = r0 + tx;      // out is cols x rows
        if (r < cols && c < rows) out[(size_t)r * rows + c] = tile[tx][ty + k];
    }
}

// ---------------------------------------------------------------------------
// K4: per row of specT (one halfcomplex x-index i, length NYM along y):
// DCT-II, multiply by -1/(lamx_i + lamy_k) ((0,0) -> 0), DCT-III
// ---------------------------------------------------------------------------
__device__ __forceinline__ int vmap2(int nn, int N) { return (nn < N / 2) ? 2 * nn : 2 * (N - 1 - nn) + 1; }

#define KST 8
__global__ void __launch_bounds__(NTH) k_dct_rows(pvk_grid g, pvk_fft F, pvk_scal *sc, pvk_pcg w, double *specT)
{
    extern __shared__ double lds[];
    if (sc && sc->p_done) return;
    const int N = g.NYM, M = N / 2, logM = F.logmy, R = w.rdct, NX = g.NXM;
    double *sr = lds, *si = lds + R * M;
    const int i0 = blockIdx.x * R;
    for (int t = threadIdx.x; t < R * M; t += NTH) {
        int q = t / M, m = t - q * M, i = i0 + q;
        double a = 0.0, b = 0.0;
        if (i < NX) {
            const double *row = specT + (size_t)i * N;
            a = row[vmap2(2 * m, N)];
            b = row[vmap2(2 * m + 1, N)];
        }
        sr[q * M + m] = a; si[q * M + m] = b;
    }
    __syncthreads();
    const int T = sk_threads(M), qq = threadIdx.x / T, lane = threadIdx.x - qq * T;
    sk_fft<false>(sr + qq * M, si + qq * M, M, logM, F.wy, lane, T);
    const int npair = M / 2 + 1;
    cplx2 Ya[KST], Yb[KST];
    int cnt = 0;
    for (int t = threadIdx.x; t < R * npair && cnt < KST; t += NTH, cnt++) {
        int q = t / npair, k = t - q * npair, i = i0 + q;
        int kk[2] = {k, M - k};
        cplx2 Xp[2];
        for (int h = 0; h < 2; h++) {
            int kv = kk[h];
            cplx2 V = rpost(sr, si, q * M, M, kv, F.rty);
            double c = F.dcty[2 * kv], s = F.dcty[2 * kv + 1];
            double Ck = c * V.x + s * V.y;
            double Sk = c * V.y - s * V.x;
            double CNk = -Sk;
            double lx = (i < NX) ? F.lamx[i] : 1.0;
            double gk = (kv == 0 && i == 0) ? 0.0 : -1.0 / (lx + F.lamy[kv]);
            double gN = (kv == 0) ? 0.0 : -1.0 / (lx + F.lamy[N - kv]);
            double Cp = Ck * gk, CNp = (kv == 0) ? 0.0 : CNk * gN;
            Xp[h].x = c * Cp + s * CNp;
            Xp[h].y = s * Cp - c * CNp;
        }
        Ya[cnt] = rpre(Xp[0], Xp[1], k, F.rty);
        Yb[cnt] = rpre(Xp[1], Xp[0], M - k, F.rty);
    }
    __syncthreads();
    cnt = 0;
    for (int t = threadIdx.x; t < R * npair && cnt < KST; t += NTH, cnt++) {
        int q = t / npair, k = t - q * npair;
        if (k < M) { sr[q * M + k] = Ya[cnt].x; si[q * M + k] = Ya[cnt].y; }
        int k2 = M - k;
        if (k2 < M && k2 != k) { sr[q * M + k2] = Yb[cnt].x; si[q * M + k2] = Yb[cnt].y; }
    }
    __syncthreads();
    sk_fft<true>(sr + qq * M, si + qq * M, M, logM, F.wy, lane, T);
    const double invN = 1.0 / (double)N;
    for (int t = threadIdx.x; t < R * M; t += NTH) {
        int q = t / M, m = t - q * M, i = i0 + q;
        if (i >= NX) continue;
        double *row = specT + (size_t)i * N;
        row[vmap2(2 * m, N)] = sr[q * M + m] * invN;
        row[vmap2(2 * m + 1, N)] = si[q * M + m] * invN;
    }
}

// ---------------------------------------------------------------------------
// K6: z = inverse real FFT of spec rows; partials <r, z> into rz[(n+1)&1]
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(NTH) k_ifft_rows_dot(pvk_grid g, pvk_fft F, pvk_scal *sc, pvk_pcg w, int n)
{
    extern __shared__ double lds[];
    __shared__ double sh[NTH];
    if (sc->p_done) return;
    const int P = g.P, N = g.NXM, M = N / 2, logM = F.logmx, R = w.rrow;
    double *sr = lds, *si = lds + R * M;
    const int j0 = blockIdx.x * R;
    for (int t = threadIdx.x; t < R * M; t += NTH) {
        int q = t / M, k = t - q * M, j = j0 + q;
        cplx2 Y = {0.0, 0.0};
        if (j < g.NYM) {
            const double *row = F.spec + (size_t)j * N;
            Y = rpre(hcget(row, N, k), hcget(row, N, M - k), k, F.rtx);
        }
        sr[q * M + k] = Y.x; si[q * M + k] = Y.y;
    }
    __syncthreads();
    {
        const int T = sk_threads(M), q = threadIdx.x / T;
        sk_fft<true>(sr + q * M, si + q * M, M, logM, F.wx, threadIdx.x - q * T, T);
    }
    double acc = 0.0;
    for (int t = threadIdx.x; t < R * M; t += NTH) {
        int q = t / M, m = t - q * M, j = j0 + q;
        if (j >= g.NYM) continue;
        double a = sr[q * M + m], b = si[q * M + m];
        AT(w.z, 2 * m, j) = a;
        AT(w.z, 2 * m + 1, j) = b;
        acc += AT(w.r, 2 * m, j) * a + AT(w.r, 2 * m + 1, j) * b;
    }
    block_store_partial(acc, w.rz[(n + 1) & 1], blockIdx.x, sh);
}

// partial sums of rhs (mean removal before the solve)
__global__ void __launch_bounds__(NTH) k_pcg_mean(pvk_grid g, pvk_scal *sc, pvk_pcg w)
{
    __shared__ double sh[NTH];
    if (blockIdx.x == 0 && threadIdx.x == 0) { sc->p_done = 0; sc->p_iter = 0; }
    const int P = g.P, NXM = g.NXM;
    const int j0 = blockIdx.x * w.band, j1 = min(g.NYM, j0 + w.band);
    double acc = 0.0;
    for (int c = threadIdx.x; c < (j1 - j0) * NXM; c += NTH) {
        int j = j0 + c / NXM, i = c - (c / NXM) * NXM;
        acc += AT(w.rhs, i, j);
    }
    block_store_partial(acc, w.mean, blockIdx.x, sh);
}

static size_t lds_rows(int R, int N) { return (size_t)R * (size_t)(N / 2) * 2 * sizeof(double); }

extern "C" void pvk_pcg_start(const pvk_grid *g, const pvk_fft *F, pvk_scal *sc, const pvk_pcg *w, double tol,
                              hipStream_t s)
{
    int nband = (g->NYM + w->band - 1) / w->band, nrow = (g->NYM + w->rrow - 1) / w->rrow;
    hipLaunchKernelGGL(k_pcg_mean, dim3(nband), dim3(NTH), 0, s, *g, sc, *w);
    hipLaunchKernelGGL(k_pcg_update_fwd, dim3(nrow), dim3(NTH), lds_rows(w->rrow, g->NXM), s, *g, *F, sc, *w, -1);
    (void)tol;
}

// stage: 0 = K1 apply, 1 = K2 update + forward rows, 2 = K3 check + transpose,
// 3 = K4 column DCTs, 4 = K5 transpose back, 5 = K6 inverse rows + <r, z>.
// n = -1 runs stages 2..5 on the start residual.
extern "C" void pvk_pcg_stage(const pvk_grid *g, const pvk_fft *F, pvk_scal *sc, const pvk_pcg *w, int stage, int n,
                              double tol, int maxit, hipStream_t s)
{
    int nband = (g->NYM + w->band - 1) / w->band, nrow = (g->NYM + w->rrow - 1) / w->rrow;
    switch (stage) {
    case 0:
        hipLaunchKernelGGL(k_pcg_apply, dim3(nband), dim3(NTH), 0, s, *g, sc, *w, n);
        break;
    case 1:
        hipLaunchKernelGGL(k_pcg_update_fwd, dim3(nrow), dim3(NTH), lds_rows(w->rrow, g->NXM), s, *g, *F, sc, *w, n);
        break;
    case 2:
        hipLaunchKernelGGL(k_transpose, dim3((g->NXM + 31) / 32, (g->NYM + 31) / 32), dim3(NTH), 0, s, sc, *w,
                           (const double *)F->spec, w->specT, g->NYM, g->NXM, 1, n, tol, maxit);
        break;
    case 3:
        hipLaunchKernelGGL(k_dct_rows, dim3((g->NXM + w->rdct - 1) / w->rdct), dim3(NTH), lds_rows(w->rdct, g->NYM),
                           s, *g, *F, sc, *w, w->specT);
        break;
    case 4:
        hipLaunchKernelGGL(k_transpose, dim3((g->NYM + 31) / 32, (g->NXM + 31) / 32), dim3(NTH), 0, s, sc, *w,
                           (const double *)w->specT, F->spec, g->NXM, g->NYM, 0, n, tol, maxit);
        break;
    default:
        hipLaunchKernelGGL(k_ifft_rows_dot, dim3(nrow), dim3(NTH), lds_rows(w->rrow, g->NXM), s, *g, *F, sc, *w, n);
        break;
    }
}
