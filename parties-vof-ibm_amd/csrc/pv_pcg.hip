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
#include "pv_stencil.h"

#define AT(a, i, j) (a)[(ptrdiff_t)(j) * P + (i)]

struct cplx2 { double x, y; };

// log2 of the threads per transform of length 2^logM in the Stockham FFT
// (pv_fftk.h): M / 8 threads, at least one
__device__ __forceinline__ int sk_logT(int logM) { return logM >= 3 ? logM - 3 : 0; }

// X_k, k = 0..M, of a real length-2M sequence from Z = FFT_M(x_{2m} + i x_{2m+1})
__device__ __forceinline__ cplx2 rpost(const double *sr, const double *si, int base, int M, int k, const double *rt)
{
    int ka = (k == M) ? 0 : k, kb = (k == 0 || k == M) ? 0 : M - k;
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

// row q and entry m of item t of a block's rows of M values
__device__ __forceinline__ void row_split(int t, int M, int logM, bool p2, int &q, int &m)
{
    if (p2) { q = t >> logM; m = t & (M - 1); }
    else { q = t / M; m = t - q * M; }
}

// the block's transforms of its R rows: power-of-two lengths by the
// shift-indexed passes, 2^a 3^b lengths by the mixed-radix ones
template <bool INV>
__device__ __forceinline__ void rows_fft(double *sr, double *si, int M, int logM, bool p2, int R, const double *tw)
{
    if (p2) {
        const int lT = sk_logT(logM), q = threadIdx.x >> lT;
        sk_fft<INV>(sr + q * M, si + q * M, logM, tw, threadIdx.x & ((1 << lT) - 1), 1 << lT);
    } else {
        const int T = M >> 3, q = threadIdx.x / T;
        sg_fft<INV>(sr + q * M, si + q * M, M, tw, threadIdx.x - q * T, T, q < R);
    }
}

// Every FFT kernel handles R = 2048 / M rows of M complex points per block
// (T = M / 8 threads per transform, 256 threads), so each thread owns
// IT = R * M / 256 = 8 items of the row sweeps; the sweeps load all of a
// thread's items before storing any, so the loads are in flight together.
#define IT 8
#define CPT 4

// Stockham twiddles (M complex) staged in LDS behind the 2 R M row values.
__device__ __forceinline__ double *stage_twiddles(double *lds, int RM, int M, const double *tw)
{
    double *t = lds + 2 * RM;
    for (int k = threadIdx.x; k < 2 * M; k += NTH) t[k] = tw[k];
    return t;
}

// ---------------------------------------------------------------------------
// Face coefficients of the operator, formed once per solve (rho is fixed):
// bx(i,j) = 2 / (rho_i + rho_{i+1}) idx^2 for i in [-1, NXM), by(i,j) the
// same along y for j in [-1, NYM).  Addition commutes exactly, so
// bx(i-1, j) is bit for bit the west coefficient of cell (i, j).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(NTH) k_pcg_faces(pvk_grid g, pvk_pcg w)
{
    const int P = g.P, NXM = g.NXM, NYM = g.NYM;
    const int j = (int)blockIdx.x - 1;
    const double idx2 = g.idx_c[1] * g.idx_c[1], idy2 = g.idy_c[1] * g.idy_c[1];
    const double *__restrict__ rho = w.rho;
    for (int i = (int)threadIdx.x - 1; i < NXM; i += NTH) {
        const double r0 = AT(rho, i, j);
        AT(w.bx, i, j) = (2.0 / (r0 + AT(rho, i + 1, j))) * idx2;
        AT(w.by, i, j) = (2.0 / (r0 + AT(rho, i, j + 1))) * idy2;
    }
}

// ---------------------------------------------------------------------------
// K1: d_n = z + beta d_{n-1}, Ad = A d_n, partial <d_n, Ad>.  One block per
// row j: the y neighbours and row clamps are block-uniform.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(NTH) k_pcg_apply(pvk_grid g, pvk_scal *sc, pvk_pcg w, int n)
{
    __shared__ double sh[2 * NTH];
    if (sc->p_done) return;
    const int P = g.P, NXM = g.NXM, NYM = g.NYM;
    double beta = 0.0, sa, sb;
    if (n > 0 && w.flex) {    // flexible CG: A-orthogonal to the previous direction only
        block_sum2(w.zad, w.nrz, w.dad[(n + 1) & 1], w.ndad, sh, sa, sb);
        beta = -sa / sb;
    } else if (n > 0) {       // Fletcher-Reeves: gamma_n / gamma_{n-1}
        block_sum2(w.rz[n & 1], w.nrz, w.rz[(n + 1) & 1], w.nrz, sh, sa, sb);
        beta = sa / sb;
    }
    const double *__restrict__ z = w.z;
    const double *__restrict__ dold = w.d[(n + 1) & 1];
    const double *__restrict__ bx = w.bx, *__restrict__ by = w.by;
    double *__restrict__ dnew = w.d[n & 1];
    double *__restrict__ Ad = w.Ad;
    const int j = xcd_band(blockIdx.x, gridDim.x);
    const int jn = (j + 1 == NYM) ? j : j + 1, js = (j == 0) ? 0 : j - 1;
    double acc = 0.0, sacc = 0.0;
    for (int base = 0; base < NXM; base += NTH * CPT) {
        double dc[CPT], ad[CPT];
#pragma unroll
        for (int u = 0; u < CPT; u++) {
            const int i = base + u * NTH + (int)threadIdx.x;
            dc[u] = 0.0; ad[u] = 0.0;
            if (i < NXM) {
                const int ie = (i + 1 == NXM) ? 0 : i + 1, iw = (i == 0) ? NXM - 1 : i - 1;
#define DN(ii, jj) (n == 0 ? AT(z, ii, jj) : AT(z, ii, jj) + beta * AT(dold, ii, jj))
                const double d0 = DN(i, j), dE = DN(ie, j), dW = DN(iw, j), dN = DN(i, jn), dS = DN(i, js);
#undef DN
                const double cxp = AT(bx, i, j), cxm = AT(bx, i - 1, j);
                const double cyp = AT(by, i, j), cym = AT(by, i, j - 1);
                dc[u] = d0;
                ad[u] = cxp * (dE - d0) - cxm * (d0 - dW) + cyp * (dN - d0) - cym * (d0 - dS);
            }
        }
#pragma unroll
        for (int u = 0; u < CPT; u++) {
            const int i = base + u * NTH + (int)threadIdx.x;
            if (i < NXM) {
                AT(dnew, i, j) = dc[u];
                AT(Ad, i, j) = ad[u];
                acc += dc[u] * ad[u];
                sacc += ad[u];
            }
        }
    }
    block_store_partial(acc, w.dad[n & 1], j, sh);
    if (w.sad) block_store_partial(sacc, w.sad, j, sh);
}

// ---------------------------------------------------------------------------
// K2: alpha; phi, r update; <r,r>; forward real FFT of the r rows -> spec
// (also the start kernel when n < 0: r = rhs - mean, phi = 0)
// ---------------------------------------------------------------------------
template <bool P2>
__global__ void __launch_bounds__(NTH) k_pcg_update_fwd(pvk_grid g, pvk_fft F, pvk_scal *sc, pvk_pcg w, int n)
{
    extern __shared__ double lds[];
    __shared__ double sh[2 * NTH];
    if (w.chb == 0.0 && sc->p_done) return;
    const int P = g.P, N = g.NXM, logM = F.logmx, M = F.mx, R = w.rrow, RM = R * M;
    constexpr bool p2 = P2;     // power-of-two rows: shift-indexed passes
    double *sr = lds, *si = lds + RM;
    double *tw = stage_twiddles(lds, RM, M, F.wx);
    const double *__restrict__ d = (n >= 0) ? w.d[n & 1] : w.rhs;
    const double *__restrict__ Ad = w.Ad, *__restrict__ rhs = w.rhs;
    double *__restrict__ phi = w.phi, *__restrict__ r = w.r;
    const int j0 = blockIdx.x * R;
    // the rows' loads go out before the scalar sums, whose partials and
    // barriers would otherwise hold them back; the arithmetic is unchanged
    double xa[IT], xb[IT], ea[IT], eb[IT], ra[IT], rb[IT], qa[IT], qb[IT];
#pragma unroll
    for (int u = 0; u < IT; u++) {
        const int t = threadIdx.x + u * NTH;
        int q, m;
        row_split(t, M, logM, p2, q, m);
        const int j = j0 + q, i = 2 * m;
        xa[u] = xb[u] = ea[u] = eb[u] = ra[u] = rb[u] = qa[u] = qb[u] = 0.0;
        if (t < RM && j < g.NYM) {
            if (n >= 0) {
                xa[u] = AT(phi, i, j); xb[u] = AT(phi, i + 1, j);
                ea[u] = AT(d, i, j);   eb[u] = AT(d, i + 1, j);
                ra[u] = AT(r, i, j);   rb[u] = AT(r, i + 1, j);
                qa[u] = AT(Ad, i, j);  qb[u] = AT(Ad, i + 1, j);
            } else if (n == -3) {               // M_fft of the current r, no update
                ra[u] = AT(r, i, j);
                rb[u] = AT(r, i + 1, j);
            } else if (n == -2) {               // warm restart: Ad holds A phi0
                xa[u] = AT(phi, i, j); xb[u] = AT(phi, i + 1, j);
                ra[u] = AT(rhs, i, j);     rb[u] = AT(rhs, i + 1, j);
                qa[u] = AT(Ad, i, j);      qb[u] = AT(Ad, i + 1, j);
            } else {
                ra[u] = AT(rhs, i, j);
                rb[u] = AT(rhs, i + 1, j);
            }
        }
    }
    double alpha = 0.0, mean = 0.0;
    if (n >= 0) {
        double srz, sdad;
        block_sum2(w.rz[n & 1], w.nrz, w.dad[n & 1], w.ndad, sh, srz, sdad);
        alpha = srz / sdad;
    } else if (n != -3)
        mean = block_sum(w.mean, w.nmean, sh) / ((double)g.NXM * (double)g.NYM);
#pragma unroll
    for (int u = 0; u < IT; u++) {
        const int t = threadIdx.x + u * NTH;
        int q, m;
        row_split(t, M, logM, p2, q, m);
        if (!(t < RM && j0 + q < g.NYM)) continue;
        if (n >= 0) {
            ra[u] = ra[u] - alpha * qa[u];
            rb[u] = rb[u] - alpha * qb[u];
        } else if (n == -2) {
            ra[u] = ra[u] - mean - qa[u];
            rb[u] = rb[u] - mean - qb[u];
        } else if (n != -3) {
            ra[u] = ra[u] - mean;
            rb[u] = rb[u] - mean;
        }
    }
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < IT; u++) {
        const int t = threadIdx.x + u * NTH;
        int q, m;
        row_split(t, M, logM, p2, q, m);
        const int j = j0 + q, i = 2 * m;
        if (t < RM && j < g.NYM && n != -3) {
            AT(phi, i, j) = (n >= 0) ? xa[u] + alpha * ea[u] : (n == -2 ? xa[u] : 0.0);
            AT(phi, i + 1, j) = (n >= 0) ? xb[u] + alpha * eb[u] : (n == -2 ? xb[u] : 0.0);
            AT(r, i, j) = ra[u];
            AT(r, i + 1, j) = rb[u];
        }
        if (t < RM && j < g.NYM) acc += ra[u] * ra[u] + rb[u] * rb[u];
        if (t < RM) { sr[t] = ra[u]; si[t] = rb[u]; }
    }
    block_store_partial(acc, w.rr, blockIdx.x, sh);
    rows_fft<false>(sr, si, M, logM, p2, R, tw);
    const double scale = 1.0 / (double)N;
    // X_k, k < M (and X_M for each row below): halfcomplex row of spec
#pragma unroll
    for (int u = 0; u < IT; u++) {
        const int t = threadIdx.x + u * NTH;
        int q, k;
        row_split(t, M, logM, p2, q, k);
        const int j = j0 + q;
        if (t < RM && j < g.NYM) {
            cplx2 X = rpost(sr, si, q * M, M, k, F.rtx);
            double *row = F.spec + (size_t)j * N;
            row[k] = X.x * scale;
            if (k > 0) row[N - k] = X.y * scale;
        }
    }
    if ((int)threadIdx.x < R && j0 + (int)threadIdx.x < g.NYM) {
        const int q = threadIdx.x;
        cplx2 X = rpost(sr, si, q * M, M, M, F.rtx);
        F.spec[(size_t)(j0 + q) * N + M] = X.x * scale;
    }
}

// ---------------------------------------------------------------------------
// K3 / K5: 32 x 32 tiled transpose of an (rows x cols) row-major array;
// with check != 0 the stop test of iteration n runs first (all blocks).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(NTH) k_transpose(pvk_scal *sc, pvk_pcg w, const double *__restrict__ in,
                                                  double *__restrict__ out, int rows, int cols, int check, int n,
                                                  double tol, int maxit)
{
    __shared__ double tile[32][33];
    __shared__ double sh[NTH];
    // a stop flag set by an earlier iteration leaves <r,r> stale: no recount
    if (w.chb == 0.0 && sc->p_done) return;
    if (check && n != -3) {                 // n = -3: a plain M_fft solve, no stop test
        double rr = block_sum(w.rr, w.nrr, sh);
        // n = -1: cold start, ||r0|| = ||b - mean|| is the reference norm;
        // n = -2: warm restart, the cold start's norm is kept
        double bn = (n == -1) ? ((sqrt(rr) == 0.0) ? 1.0 : sqrt(rr)) : sc->p_bnorm;
        double rel = sqrt(rr) / bn;
        bool done = (rel <= tol) || (n + 1 >= maxit && n >= 0);
        if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
            if (n == -1) sc->p_bnorm = bn;
            sc->p_res = rel;
            sc->p_iter = (n < 0) ? 0 : n + 1;
            if (done) sc->p_done = (rel <= tol) ? 1 : 3;
        }
        if (done) return;
    }
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    double v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int r = r0 + ty + 8 * k, c = c0 + tx;
        v[k] = (r < rows && c < cols) ? in[(size_t)r * cols + c] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) tile[ty + 8 * k][tx] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int r = c0 + ty + 8 * k, c = r0 + tx;      // out is cols x rows
        if (r < cols && c < rows) out[(size_t)r * rows + c] = tile[tx][ty + 8 * k];
    }
}

// ---------------------------------------------------------------------------
// K4: per row of specT (one halfcomplex x-index i, length NYM along y):
// DCT-II, multiply by -1/(lamx_i + lamy_k) ((0,0) -> 0), DCT-III.
// DCT-II of x_n as a real FFT of v_nn = x_vmap(nn) (evens ascending, odds
// descending); the row is read and written contiguously and permuted in LDS.
// ---------------------------------------------------------------------------
// LDS slot (sr or si, index) of row element n under the even/odd reordering
__device__ __forceinline__ int dct_slot(int n, int N)
{
    const int nn = (n & 1) ? N - 1 - (n >> 1) : (n >> 1);
    return nn;                       // complex index nn >> 1, part nn & 1
}

__global__ void __launch_bounds__(NTH) k_dct_rows(pvk_grid g, pvk_fft F, pvk_scal *sc, pvk_pcg w, double *specT)
{
    extern __shared__ double lds[];
    if (w.chb == 0.0 && sc && sc->p_done) return;
    const double a_ch = (w.chb > 0.0) ? 1.0 / (sc->dt * c_BET[w.chstage]) : 0.0;
    const int logM = F.logmy, M = 1 << logM, logN = logM + 1, N = 2 * M, R = w.rdct, NX = g.NXM, RM = R * M;
    double *sr = lds, *si = lds + RM;
    double *tw = stage_twiddles(lds, RM, M, F.wy);
    const int i0 = blockIdx.x * R;
    // R rows of N reals = 2 RM values: 2 IT per thread, contiguous per row
    {
        double v[2 * IT];
#pragma unroll
        for (int u = 0; u < 2 * IT; u++) {
            const int t = threadIdx.x + u * NTH, q = t >> logN, nidx = t & (N - 1), i = i0 + q;
            v[u] = (t < 2 * RM && i < NX) ? specT[(size_t)i * N + nidx] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 2 * IT; u++) {
            const int t = threadIdx.x + u * NTH, q = t >> logN, nidx = t & (N - 1);
            if (t < 2 * RM) {
                const int nn = dct_slot(nidx, N);
                ((nn & 1) ? si : sr)[q * M + (nn >> 1)] = v[u];
            }
        }
    }
    __syncthreads();
    const int lT = sk_logT(logM), T = 1 << lT, qq = threadIdx.x >> lT, lane = threadIdx.x & (T - 1);
    sk_fft<false>(sr + qq * M, si + qq * M, logM, tw, lane, T);
    // spectral step on the pairs (k, M - k): k < M/2 as IT/2 items per
    // thread, k = M/2 by threads t < R
    const int logH = logM - 1, H = M >> 1;
    cplx2 Ya[IT / 2 + 1], Yb[IT / 2 + 1];
    auto pair = [&](int q, int k, cplx2 &ya, cplx2 &yb) {
        const int i = i0 + q;
        const int kk[2] = {k, M - k};
        cplx2 Xp[2];
        const double lx = (i < NX) ? F.lamx[i] : 1.0;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int kv = kk[h];
            cplx2 V = rpost(sr, si, q * M, M, kv, F.rty);
            double c = F.dcty[2 * kv], s = F.dcty[2 * kv + 1];
            double Ck = c * V.x + s * V.y;
            double Sk = c * V.y - s * V.x;
            double CNk = -Sk;
            double gk, gN;
            if (w.chb > 0.0) {          // (a_dt + bih lambda^2)^-1, no null mode
                const double lk = lx + F.lamy[kv];
                gk = 1.0 / (a_ch + w.chb * lk * lk);
                gN = 0.0;
                if (kv != 0) {
                    const double lN = lx + F.lamy[N - kv];
                    gN = 1.0 / (a_ch + w.chb * lN * lN);
                }
            } else {
                gk = (kv == 0 && i == 0) ? 0.0 : -1.0 / (lx + F.lamy[kv]);
                gN = (kv == 0) ? 0.0 : -1.0 / (lx + F.lamy[N - kv]);
            }
            double Cp = Ck * gk, CNp = (kv == 0) ? 0.0 : CNk * gN;
            Xp[h].x = c * Cp + s * CNp;
            Xp[h].y = s * Cp - c * CNp;
        }
        ya = rpre(Xp[0], Xp[1], k, F.rty);
        yb = rpre(Xp[1], Xp[0], M - k, F.rty);
    };
#pragma unroll
    for (int u = 0; u < IT / 2; u++) {
        const int t = threadIdx.x + u * NTH;
        if (t < R * H) pair(t >> logH, t & (H - 1), Ya[u], Yb[u]);
    }
    if ((int)threadIdx.x < R) pair(threadIdx.x, H, Ya[IT / 2], Yb[IT / 2]);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < IT / 2; u++) {
        const int t = threadIdx.x + u * NTH;
        if (t < R * H) {
            const int q = t >> logH, k = t & (H - 1);
            sr[q * M + k] = Ya[u].x; si[q * M + k] = Ya[u].y;
            if (k > 0) { sr[q * M + M - k] = Yb[u].x; si[q * M + M - k] = Yb[u].y; }
        }
    }
    if ((int)threadIdx.x < R) {
        const int q = threadIdx.x;
        sr[q * M + H] = Ya[IT / 2].x; si[q * M + H] = Ya[IT / 2].y;
    }
    __syncthreads();
    sk_fft<true>(sr + qq * M, si + qq * M, logM, tw, lane, T);
    const double invN = 1.0 / (double)N;
#pragma unroll
    for (int u = 0; u < 2 * IT; u++) {
        const int t = threadIdx.x + u * NTH, q = t >> logN, nidx = t & (N - 1), i = i0 + q;
        if (t < 2 * RM && i < NX) {
            const int nn = dct_slot(nidx, N);
            specT[(size_t)i * N + nidx] = ((nn & 1) ? si : sr)[q * M + (nn >> 1)] * invN;
        }
    }
}

// ---------------------------------------------------------------------------
// K4': the y solve of the FFT preconditioner as the reference does it: one
// tridiagonal system per halfcomplex column i of spec (psolve_fft.c:1536-1802,
// tridiag :141-163), a_j = as_j, c_j = an_j, b_j = (-as_j - an_j) - mw_i,
// with row 0 of column 0 replaced by x_0 = 0 (the pin of mode (0, 0)), solved
// in place on spec[j][i] -- no transposes.
//
// Batched partitioned Thomas mapped to the workgroup: lane kk of a row group
// owns column i0 + kk (consecutive lanes read consecutive doubles of a row),
// and the tri_s segments of tri_l rows of a column are owned by tri_s
// threads.  The last row of every segment but the last is a separator.  Each
// thread eliminates its segment's interior rows (held in registers) with the
// couplings to the separators above and below removed, and keeps the
// forward-swept values, the Thomas multipliers and the forward sweep of the
// first unit vector; the separators' Schur complement is tridiagonal (one
// equation per separator: its own row with the neighbouring interior rows
// written in terms of the separators), solved by parallel cyclic reduction
// in LDS; then each thread back-substitutes its interior rows with the
// separator values in place.  With mean removal (the PCG's pseudo-inverse)
// column 0 is shifted to zero mean afterwards.  In the PCG this kernel also
// runs the stop test of iteration n (every block, from the <r, r> partials).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double tri_sum(const double *part, int n, double *sh, int nt)
{
    const int t = threadIdx.x;
    double s = 0.0;
    for (int b = t; b < n; b += nt) s += part[b];
    sh[t] = s;
    __syncthreads();
    for (int w = 256; w > 0; w >>= 1) {
        if (t < w && t + w < nt) sh[t] += sh[t + w];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

template <int TRI_LMAX>
__global__ void __launch_bounds__(512) k_tri_cols(pvk_grid g, pvk_fft F, pvk_scal *sc, pvk_pcg w, int n,
                                                   double tol, int maxit)
{
    extern __shared__ double lds[];
    if (sc->p_done) return;
    const int nt = F.tri_kw * F.tri_s;
    if (n != -3) {                          // stop test of iteration n (as the transpose stage does)
        double rr = tri_sum(w.rr, w.nrr, lds, nt);
        double bn = (n == -1) ? ((sqrt(rr) == 0.0) ? 1.0 : sqrt(rr)) : sc->p_bnorm;
        double rel = sqrt(rr) / bn;
        bool done = (rel <= tol) || (n + 1 >= maxit && n >= 0);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            if (n == -1) sc->p_bnorm = bn;
            sc->p_res = rel;
            sc->p_iter = (n < 0) ? 0 : n + 1;
            if (done) sc->p_done = (rel <= tol) ? 1 : 3;
        }
        if (done) return;
    }
    const int NX = g.NXM, NY = g.NYM, KW = F.tri_kw, S = F.tri_s, L = F.tri_l;
    const int kk = threadIdx.x & (KW - 1), s = threadIdx.x >> F.tri_lkw;
    const int grp = xcd_band(blockIdx.x, gridDim.x);     // neighbouring column groups share an XCD's L2
    const int col = grp * KW + kk;
    const bool act = col < NX;
    const int j0 = s * L, has_sep = s < S - 1;
    const int nrow = has_sep ? L : NY - (S - 1) * L, nint = has_sep ? nrow - 1 : nrow;
    const double mw = act ? F.tri_mw[col] : 0.0;
    double *spec = F.spec;
    // the segment's row coefficients are loaded with its right-hand side,
    // before the sweeps: loads inside the sweeps would put their latency on
    // the chain of dependent divisions
    double y[TRI_LMAX], gg[TRI_LMAX], ra[TRI_LMAX], rc[TRI_LMAX];
#pragma unroll
    for (int l = 0; l < TRI_LMAX; l++) {
        double v = 0.0, a = 0.0, c = 0.0;
        if (l < nrow) {
            a = F.tri_as[j0 + l];
            c = F.tri_an[j0 + l];
            if (act) v = spec[(size_t)(j0 + l) * NX + col];
        }
        if (col == 0 && j0 + l == 0) v = 0.0;
        y[l] = v; ra[l] = a; rc[l] = c;
    }
    // A, C, B of the segment's row l (l a compile-time index after unrolling)
    auto A = [&](int l) { return (col == 0 && j0 + l == 0) ? 0.0 : ra[l]; };
    auto Cc = [&](int l) { return (col == 0 && j0 + l == 0) ? 0.0 : rc[l]; };
    auto B = [&](int l) { return (col == 0 && j0 + l == 0) ? 1.0 : (-ra[l] - rc[l]) - 0.0 - mw; };
    // the separator (the segment's last row) and the last interior row's C
    double dsep = 0.0, as_ = 0.0, bs_ = 0.0, cs_ = 0.0, cbot = 0.0;
#pragma unroll
    for (int l = 0; l < TRI_LMAX; l++) {
        if (has_sep && l == nrow - 1) { dsep = y[l]; as_ = A(l); bs_ = B(l); cs_ = Cc(l); }
        if (has_sep && l == nint - 1) cbot = Cc(l);
    }
    // forward sweep over the interior rows; e1 = forward sweep of the first
    // unit vector, kept per row (the one division per row is the pivot's)
    double e1s[TRI_LMAX];
    double inv = 1.0 / B(0), e1 = inv, Yl = 0.0, Vl = 0.0;
    y[0] *= inv;
    gg[0] = 0.0;
    e1s[0] = e1;
    if (nint == 1) { Yl = y[0]; Vl = e1; }
#pragma unroll
    for (int l = 1; l < TRI_LMAX; l++) {
        gg[l] = 0.0;
        e1s[l] = 0.0;
        if (l < nint) {
            const double a = A(l);
            gg[l] = Cc(l - 1) * inv;
            inv = 1.0 / (B(l) - a * gg[l]);
            y[l] = (y[l] - a * y[l - 1]) * inv;
            e1 = -a * e1 * inv;
            e1s[l] = e1;
            if (l == nint - 1) { Yl = y[l]; Vl = e1; }
        }
    }
    const double Wl = inv;                  // 1 / last pivot: forward sweep of the last unit vector
    // first-row values of the local solutions for d, the first and the last unit vector
    double Yf = 0.0, Vf = 0.0, Wf = 0.0;
#pragma unroll
    for (int l = TRI_LMAX - 1; l >= 0; l--) {
        if (l == nint - 1) { Yf = y[l]; Vf = e1s[l]; Wf = Wl; }
        else if (l < nint - 1) {
            Yf = y[l] - gg[l + 1] * Yf;
            Vf = e1s[l] - gg[l + 1] * Vf;
            Wf = -gg[l + 1] * Wf;
        }
    }
    const double atop = (s > 0) ? A(0) : 0.0;
    // separators' Schur complement: row s couples z_{s-1}, z_s, z_{s+1}
    const int nsep = S - 1, SK = S * KW, me = s * KW + kk;
    double *RA0 = lds, *RB0 = lds + SK, *RC0 = lds + 2 * SK, *RD0 = lds + 3 * SK;
    double *RA1 = lds + 4 * SK, *RB1 = lds + 5 * SK, *RC1 = lds + 6 * SK, *RD1 = lds + 7 * SK;
    RB1[me] = Yf; RC1[me] = atop * Vf; RD1[me] = cbot * Wf;      // for the separator above this segment
    __syncthreads();
    if (s < nsep) {
        const int nb = me + KW;
        RA0[me] = -as_ * atop * Vl;
        RB0[me] = bs_ - as_ * cbot * Wl - cs_ * RC1[nb];
        RC0[me] = -cs_ * RD1[nb];
        RD0[me] = dsep - as_ * Yl - cs_ * RB1[nb];
    } else {
        RA0[me] = 0.0; RB0[me] = 1.0; RC0[me] = 0.0; RD0[me] = 0.0;
    }
    __syncthreads();
    double *ca = RA0, *cb = RB0, *cc = RC0, *cd = RD0, *na = RA1, *nb_ = RB1, *nc = RC1, *nd = RD1;
    for (int h = 1; h < nsep; h <<= 1) {
        if (s < nsep) {
            double a2 = 0.0, b2 = cb[me], c2 = 0.0, d2 = cd[me];
            if (s - h >= 0) {
                const int q = me - h * KW;
                const double al = -ca[me] / cb[q];
                a2 = al * ca[q]; b2 += al * cc[q]; d2 += al * cd[q];
            }
            if (s + h < nsep) {
                const int q = me + h * KW;
                const double ga = -cc[me] / cb[q];
                c2 = ga * cc[q]; b2 += ga * ca[q]; d2 += ga * cd[q];
            }
            na[me] = a2; nb_[me] = b2; nc[me] = c2; nd[me] = d2;
        }
        __syncthreads();
        double *t;
        t = ca; ca = na; na = t;
        t = cb; cb = nb_; nb_ = t;
        t = cc; cc = nc; nc = t;
        t = cd; cd = nd; nd = t;
    }
    double *z = na;                          // the other buffer is free now
    if (s < nsep) z[me] = cd[me] / cb[me];
    __syncthreads();
    const double zprev = (s > 0) ? z[me - KW] : 0.0, zsep = has_sep ? z[me] : 0.0;
    // interior rows from the separators (back substitution in place)
    double x = 0.0, colsum = 0.0;
#pragma unroll
    for (int l = TRI_LMAX - 1; l >= 0; l--) {
        if (l < nint) {
            double t = y[l] - atop * zprev * e1s[l];
            if (l == nint - 1) { t -= cbot * zsep * Wl; x = t; }
            else x = t - gg[l + 1] * x;
            y[l] = x;
            colsum += x;
        }
    }
    if (has_sep) colsum += zsep;
    double shift = 0.0;
    if (!w.keep_pin && grp == 0) {            // block-uniform: column 0 lives in group 0
        __syncthreads();
        if (kk == 0) na[s] = colsum;
        __syncthreads();
        if (threadIdx.x == 0) {
            double m = 0.0;
            for (int q = 0; q < S; q++) m += na[q];
            nb_[0] = m / (double)NY;
        }
        __syncthreads();
        if (col == 0) shift = nb_[0];
    }
    if (act) {
#pragma unroll
        for (int l = 0; l < TRI_LMAX; l++)
            if (l < nint) spec[(size_t)(j0 + l) * NX + col] = y[l] - shift;
        if (has_sep) spec[(size_t)(j0 + nint) * NX + col] = zsep - shift;
    }
}

// ---------------------------------------------------------------------------
// K6: z = inverse real FFT of spec rows; partials <r, z> into rz[(n+1)&1]
// ---------------------------------------------------------------------------
template <bool P2>
__global__ void __launch_bounds__(NTH) k_ifft_rows_dot(pvk_grid g, pvk_fft F, pvk_scal *sc, pvk_pcg w, int n)
{
    extern __shared__ double lds[];
    __shared__ double sh[NTH];
    if (w.chb == 0.0 && sc->p_done) return;
    const int P = g.P, N = g.NXM, logM = F.logmx, M = F.mx, R = w.rrow, RM = R * M;
    constexpr bool p2 = P2;     // power-of-two rows: shift-indexed passes
    double *sr = lds, *si = lds + RM;
    double *tw = stage_twiddles(lds, RM, M, F.wx);
    const double *__restrict__ r = w.r;
    double *__restrict__ z = w.z;
    const int j0 = blockIdx.x * R;
    {
        cplx2 Y[IT];
#pragma unroll
        for (int u = 0; u < IT; u++) {
            const int t = threadIdx.x + u * NTH;
        int q, k;
        row_split(t, M, logM, p2, q, k);
        const int j = j0 + q;
            Y[u].x = 0.0; Y[u].y = 0.0;
            if (t < RM && j < g.NYM) {
                const double *row = F.spec + (size_t)j * N;
                Y[u] = rpre(hcget(row, N, k), hcget(row, N, M - k), k, F.rtx);
            }
        }
#pragma unroll
        for (int u = 0; u < IT; u++) {
            const int t = threadIdx.x + u * NTH;
            if (t < RM) { sr[t] = Y[u].x; si[t] = Y[u].y; }
        }
    }
    __syncthreads();
    rows_fft<true>(sr, si, M, logM, p2, R, tw);
    double ra[IT], rb[IT];
#pragma unroll
    for (int u = 0; u < IT; u++) {
        const int t = threadIdx.x + u * NTH;
        int q, m;
        row_split(t, M, logM, p2, q, m);
        const int j = j0 + q;
        ra[u] = rb[u] = 0.0;
        if (t < RM && j < g.NYM) { ra[u] = AT(r, 2 * m, j); rb[u] = AT(r, 2 * m + 1, j); }
    }
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < IT; u++) {
        const int t = threadIdx.x + u * NTH;
        int q, m;
        row_split(t, M, logM, p2, q, m);
        const int j = j0 + q;
        if (t < RM && j < g.NYM) {
            const double a = sr[t], b = si[t];
            AT(z, 2 * m, j) = a;
            AT(z, 2 * m + 1, j) = b;
            acc += ra[u] * a + rb[u] * b;
        }
    }
    block_store_partial(acc, w.rz[(n < 0) ? 0 : ((n + 1) & 1)], blockIdx.x, sh);
}

__global__ void k_pcg_clear_done(pvk_scal *sc) { sc->p_done = 0; }
extern "C" void pvk_pcg_clear_done(pvk_scal *sc, hipStream_t s)
{
    hipLaunchKernelGGL(k_pcg_clear_done, dim3(1), dim3(1), 0, s, sc);
}

// sc->red[slot] = sum of the <r, z> partials of slot rzs (the energy-norm
// proxy <r, M r> that judges a warm start)
__global__ void __launch_bounds__(NTH) k_pcg_gamma(pvk_scal *sc, pvk_pcg w, int rzs, int slot)
{
    __shared__ double sh[NTH];
    const double v = block_sum(w.rz[rzs], w.nrz, sh);
    if (threadIdx.x == 0) sc->red[slot] = v;
}
extern "C" void pvk_pcg_gamma(pvk_scal *sc, const pvk_pcg *w, int rzs, int slot, hipStream_t s)
{
    hipLaunchKernelGGL(k_pcg_gamma, dim3(1), dim3(NTH), 0, s, sc, *w, rzs, slot);
}

// out = A x (the operator of k_pcg_apply on a given field; faces formed by
// the start of the solve)
__global__ void __launch_bounds__(NTH) k_pcg_opx(pvk_grid g, pvk_pcg w, const double *__restrict__ x,
                                                double *__restrict__ out)
{
    __shared__ double sh[NTH];
    const int P = g.P, NXM = g.NXM, NYM = g.NYM, j = xcd_band(blockIdx.x, gridDim.x);
    const int jn = (j + 1 == NYM) ? j : j + 1, js = (j == 0) ? 0 : j - 1;
    double sacc = 0.0;
    for (int i = threadIdx.x; i < NXM; i += NTH) {
        const int ie = (i + 1 == NXM) ? 0 : i + 1, iw = (i == 0) ? NXM - 1 : i - 1;
        const double d0 = AT(x, i, j);
        const double v = AT(w.bx, i, j) * (AT(x, ie, j) - d0) - AT(w.bx, i - 1, j) * (d0 - AT(x, iw, j)) +
                         AT(w.by, i, j) * (AT(x, i, jn) - d0) - AT(w.by, i, j - 1) * (d0 - AT(x, i, js));
        AT(out, i, j) = v;
        sacc += v;
    }
    if (w.sad) block_store_partial(sacc, w.sad, j, sh);
}
extern "C" void pvk_pcg_opx(const pvk_grid *g, const pvk_pcg *w, const double *x, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_pcg_opx, dim3(g->NYM), dim3(NTH), 0, s, *g, *w, x, out);
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

// R rows of N/2 complex values + the N/2 complex twiddles
static size_t lds_rows(int R, int N) { return ((size_t)R * (size_t)(N / 2) * 2 + (size_t)N) * sizeof(double); }

extern "C" void pvk_pcg_start(const pvk_grid *g, const pvk_fft *F, pvk_scal *sc, const pvk_pcg *w, double tol,
                              hipStream_t s)
{
    int nband = (g->NYM + w->band - 1) / w->band, nrow = (g->NYM + w->rrow - 1) / w->rrow;
    hipLaunchKernelGGL(k_pcg_faces, dim3(g->NYM + 1), dim3(NTH), 0, s, *g, *w);
    hipLaunchKernelGGL(k_pcg_mean, dim3(nband), dim3(NTH), 0, s, *g, sc, *w);
    if (F->pow2x) hipLaunchKernelGGL(k_pcg_update_fwd<true>, dim3(nrow), dim3(NTH), lds_rows(w->rrow, g->NXM), s, *g, *F, sc, *w, -1);
    else hipLaunchKernelGGL(k_pcg_update_fwd<false>, dim3(nrow), dim3(NTH), lds_rows(w->rrow, g->NXM), s, *g, *F, sc, *w, -1);
    (void)tol;
}

// face coefficients + partial sums of rhs (resets the stop flag): the start
// of a solve for preconditioners other than the FFT one
extern "C" void pvk_pcg_prep(const pvk_grid *g, pvk_scal *sc, const pvk_pcg *w, hipStream_t s)
{
    hipLaunchKernelGGL(k_pcg_faces, dim3(g->NYM + 1), dim3(NTH), 0, s, *g, *w);
    hipLaunchKernelGGL(k_pcg_mean, dim3((g->NYM + w->band - 1) / w->band), dim3(NTH), 0, s, *g, sc, *w);
}

// K1 alone (d_n = z + beta d_{n-1}, Ad, <d_n, Ad>), for any preconditioner
extern "C" void pvk_pcg_apply(const pvk_grid *g, pvk_scal *sc, const pvk_pcg *w, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_pcg_apply, dim3(g->NYM), dim3(NTH), 0, s, *g, sc, *w, n);
}

// stage: 0 = K1 apply, 1 = K2 update + forward rows, 2 = K3 check + transpose,
// 3 = K4 column DCTs, 4 = K5 transpose back, 5 = K6 inverse rows + <r, z>,
// 6 = K4' check + tridiagonal y solves in place (replaces 2..4 for the
// pressure).  n = -1 runs stages 2..5 (or 6, 5) on the start residual.
extern "C" void pvk_pcg_stage(const pvk_grid *g, const pvk_fft *F, pvk_scal *sc, const pvk_pcg *w, int stage, int n,
                              double tol, int maxit, hipStream_t s)
{
    int nband = (g->NYM + w->band - 1) / w->band, nrow = (g->NYM + w->rrow - 1) / w->rrow;
    switch (stage) {
    case 0:
        hipLaunchKernelGGL(k_pcg_apply, dim3(g->NYM), dim3(NTH), 0, s, *g, sc, *w, n);
        break;
    case 1:
        if (F->pow2x)
            hipLaunchKernelGGL(k_pcg_update_fwd<true>, dim3(nrow), dim3(NTH), lds_rows(w->rrow, g->NXM), s, *g, *F, sc, *w, n);
        else
            hipLaunchKernelGGL(k_pcg_update_fwd<false>, dim3(nrow), dim3(NTH), lds_rows(w->rrow, g->NXM), s, *g, *F, sc, *w, n);
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
    case 6: {
        const int nt = F->tri_kw * F->tri_s;
        if (F->tri_l <= 8)                   // registers for 8 or 16 rows a thread
            hipLaunchKernelGGL(k_tri_cols<8>, dim3((g->NXM + F->tri_kw - 1) / F->tri_kw), dim3(nt),
                               (size_t)8 * nt * sizeof(double), s, *g, *F, sc, *w, n, tol, maxit);
        else
            hipLaunchKernelGGL(k_tri_cols<16>, dim3((g->NXM + F->tri_kw - 1) / F->tri_kw), dim3(nt),
                               (size_t)8 * nt * sizeof(double), s, *g, *F, sc, *w, n, tol, maxit);
    } break;
    default:
        if (F->pow2x)
            hipLaunchKernelGGL(k_ifft_rows_dot<true>, dim3(nrow), dim3(NTH), lds_rows(w->rrow, g->NXM), s, *g, *F, sc, *w, n);
        else
            hipLaunchKernelGGL(k_ifft_rows_dot<false>, dim3(nrow), dim3(NTH), lds_rows(w->rrow, g->NXM), s, *g, *F, sc, *w, n);
        break;
    }
}
