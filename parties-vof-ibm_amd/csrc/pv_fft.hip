// pv_fft.hip -- constant-coefficient Poisson solve, FFT in x and cosine
// transform in y, used as the PCG preconditioner and as the standalone
// x-periodic Poisson solve (maths of lsolver/psolve_fft.c:17-207, :467-591,
// :1536-1802 with z of length 1).
//
// On a uniform grid the Neumann tridiagonal in y (as_j = an_j = idy^2,
// as_0 = an_{N-1} = 0) is diagonalised exactly by the DCT-II, so the
// per-wavenumber Thomas sweeps become one more batched transform:
//   z = IFFT_x( IDCT_y( DCT_y( FFT_x(r) / NXM ) / lambda(kx, ky) ) ),
//   lambda = -(2(1 - cos(2 pi kx / NXM)) idx^2 + 2(1 - cos(pi ky / NYM)) idy^2).
// Three kernels per application, each a batch of length-N real transforms
// done as length-N/2 complex radix-2 FFTs in LDS (power-of-two N):
//   k_fft_rows_fwd   rows j:      r -> spec (halfcomplex along i, scaled 1/NXM)
//   k_dct_cols       column strips: DCT-II, divide by lambda, DCT-III
//   k_fft_rows_inv   rows j:      spec -> z, plus the partial <r, z>
// Data stay row-major (no transposes): column strips read 8 adjacent
// columns, i.e. 64-byte segments per row.
#include <hip/hip_runtime.h>
#include <math.h>
#include "pv_kernels.h"

#define AT(a, i, j) (a)[(ptrdiff_t)(j) * P + (i)]
#define FT 256          // threads per block (row kernels)
#define FTC 512         // threads per block (column kernel)
#define KSTASH 8        // pairs stashed per thread in the column kernel

struct cplx { double x, y; };

__device__ __forceinline__ unsigned brev(unsigned m, int logM)
{
    return logM ? (__brev(m) >> (32 - logM)) : 0u;
}

// In-place radix-2 DIT on R sequences of length M (bit-reversed input) held
// in LDS as sr[q*M + m], si[q*M + m].  tw[t] = (cos, sin)(2 pi t / M), t < M/2.
__device__ void lds_fft(double *sr, double *si, int R, int M, int logM, const double *tw, bool inverse)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    const int nb = R * (M >> 1);
    for (int s = 0; s < logM; s++) {
        const int half = 1 << s;
        const int tstep = (M >> 1) >> s;
        for (int b = tid; b < nb; b += nt) {
            int q = b / (M >> 1), bb = b - q * (M >> 1);
            int grp = bb >> s, pos = bb & (half - 1);
            int i0 = q * M + grp * 2 * half + pos, i1 = i0 + half;
            double c = tw[2 * (pos * tstep)], sn = tw[2 * (pos * tstep) + 1];
            double wr = c, wi = inverse ? sn : -sn;
            double xr = sr[i1], xi = si[i1];
            double tr = wr * xr - wi * xi, ti = wr * xi + wi * xr;
            double ar = sr[i0], ai = si[i0];
            sr[i1] = ar - tr; si[i1] = ai - ti;
            sr[i0] = ar + tr; si[i0] = ai + ti;
        }
        __syncthreads();
    }
}

// X_k (k = 0..M) of a real sequence of length N = 2M from Z = FFT_M(z),
// z_m = x_{2m} + i x_{2m+1}.  rt[k] = (cos, sin)(2 pi k / N).
__device__ __forceinline__ cplx real_post(const double *sr, const double *si, int base, int M, int k,
                                          const double *rt)
{
    int ka = (k == M) ? 0 : k, kb = (M - k) & (M - 1);
    if (k == 0) kb = 0;
    double zr = sr[base + ka], zi = si[base + ka];
    double cr = sr[base + kb], ci = -si[base + kb];       // conj(Z_{M-k})
    double er = 0.5 * (zr + cr), ei = 0.5 * (zi + ci);   // E_k
    double dr = zr - cr, di = zi - ci;
    double or_ = 0.5 * di, oi = -0.5 * dr;               // O_k = -i/2 (Z_k - conj Z_{M-k})
    double c = rt[2 * k], s = rt[2 * k + 1];              // W^k = c - i s
    cplx X;
    X.x = er + (c * or_ + s * oi);
    X.y = ei + (c * oi - s * or_);
    return X;
}

// Y_k = (X_k + conj X_{M-k}) + i (X_k - conj X_{M-k}) W^{-k}; the inverse
// complex FFT of Y gives x_{2m} + i x_{2m+1} of the unnormalised HC2R.
__device__ __forceinline__ cplx real_pre(cplx Xk, cplx Xmk, int k, const double *rt)
{
    double ar = Xk.x + Xmk.x, ai = Xk.y - Xmk.y;         // X_k + conj X_{M-k}
    double dr = Xk.x - Xmk.x, di = Xk.y + Xmk.y;         // X_k - conj X_{M-k}
    double c = rt[2 * k], s = rt[2 * k + 1];              // W^{-k} = c + i s
    double pr = dr * c - di * s, pi_ = dr * s + di * c;
    cplx Y;
    Y.x = ar - pi_;
    Y.y = ai + pr;
    return Y;
}

__device__ __forceinline__ cplx hc_get(const double *row, int N, int k)
{
    cplx X;
    X.x = row[k];
    X.y = (k == 0 || k == N / 2) ? 0.0 : row[N - k];
    return X;
}

// ---------------------------------------------------------------------------
// rows: forward real FFT (R2HC, scaled by 1/N) of r -> spec
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(FT) k_fft_rows_fwd(pvk_grid g, pvk_fft F, const pvk_scal *sc, const double *in,
                                                     double *spec, double scale)
{
    extern __shared__ double lds[];
    if (sc && sc->p_done) return;
    const int N = g.NXM, M = N / 2, logM = F.logmx, R = F.rx, P = g.P;
    double *sr = lds, *si = lds + R * M;
    const int j0 = blockIdx.x * R;
    for (int t = threadIdx.x; t < R * M; t += FT) {
        int q = t / M, m = t - q * M, j = j0 + q;
        double a = 0.0, b = 0.0;
        if (j < g.NYM) { a = AT(in, 2 * m, j); b = AT(in, 2 * m + 1, j); }
        int d = q * M + brev(m, logM);
        sr[d] = a; si[d] = b;
    }
    __syncthreads();
    lds_fft(sr, si, R, M, logM, F.twx, false);
    for (int t = threadIdx.x; t < R * (M + 1); t += FT) {
        int q = t / (M + 1), k = t - q * (M + 1), j = j0 + q;
        if (j >= g.NYM) continue;
        cplx X = real_post(sr, si, q * M, M, k, F.rtx);
        double *row = spec + (size_t)j * N;
        row[k] = X.x * scale;
        if (k > 0 && k < M) row[N - k] = X.y * scale;
    }
}

// ---------------------------------------------------------------------------
// rows: inverse real FFT (HC2R, unnormalised) spec -> out, partial <r, out>
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(FT) k_fft_rows_inv(pvk_grid g, pvk_fft F, const pvk_scal *sc, const double *spec,
                                                     double *out, const double *rdot, double *part, int nbpart)
{
    extern __shared__ double lds[];
    if (sc && sc->p_done) return;
    const int N = g.NXM, M = N / 2, logM = F.logmx, R = F.rx, P = g.P;
    double *sr = lds, *si = lds + R * M;
    const int j0 = blockIdx.x * R;
    for (int t = threadIdx.x; t < R * M; t += FT) {
        int q = t / M, k = t - q * M, j = j0 + q;
        cplx Y = {0.0, 0.0};
        if (j < g.NYM) {
            const double *row = spec + (size_t)j * N;
            Y = real_pre(hc_get(row, N, k), hc_get(row, N, M - k), k, F.rtx);
        }
        int d = q * M + brev(k, logM);
        sr[d] = Y.x; si[d] = Y.y;
    }
    __syncthreads();
    lds_fft(sr, si, R, M, logM, F.twx, true);
    double acc = 0.0;
    for (int t = threadIdx.x; t < R * M; t += FT) {
        int q = t / M, m = t - q * M, j = j0 + q;
        if (j >= g.NYM) continue;
        double a = sr[q * M + m], b = si[q * M + m];
        AT(out, 2 * m, j) = a;
        AT(out, 2 * m + 1, j) = b;
        if (rdot) acc += AT(rdot, 2 * m, j) * a + AT(rdot, 2 * m + 1, j) * b;
    }
    if (part) {
        __shared__ double sh[FT / 64];
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            double s = sh[0];
            for (int w = 1; w < FT / 64; w++) s += sh[w];
            part[blockIdx.x] = s;
        }
    }
}

// ---------------------------------------------------------------------------
// columns: DCT-II along j, divide by lambda, DCT-III along j (in place)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int vmap(int n, int N) { return (n < N / 2) ? 2 * n : 2 * (N - 1 - n) + 1; }

__global__ void __launch_bounds__(FTC) k_dct_cols(pvk_grid g, pvk_fft F, const pvk_scal *sc, double *spec,
                                                 int mean_free)
{
    extern __shared__ double lds[];
    if (sc && sc->p_done) return;
    const int N = g.NYM, M = N / 2, logM = F.logmy, R = F.ry, NX = g.NXM;
    double *sr = lds, *si = lds + R * M;
    const int i0 = blockIdx.x * R;
    // v_n = x_{vmap(n)}; z_m = v_{2m} + i v_{2m+1}, loaded bit-reversed
    for (int t = threadIdx.x; t < R * M; t += blockDim.x) {
        int m = t / R, q = t - m * R, i = i0 + q;
        double a = 0.0, b = 0.0;
        if (i < NX) {
            a = spec[(size_t)vmap(2 * m, N) * NX + i];
            b = spec[(size_t)vmap(2 * m + 1, N) * NX + i];
        }
        int d = q * M + brev(m, logM);
        sr[d] = a; si[d] = b;
    }
    __syncthreads();
    lds_fft(sr, si, R, M, logM, F.twy, false);
    // pairs (k, M-k), k = 0..M/2: V_k -> C_k, C_{N-k} -> scale -> V'_k -> Y_k
    const int npair = M / 2 + 1;
    cplx Ya[KSTASH], Yb[KSTASH];
    int cnt = 0;
    for (int t = threadIdx.x; t < R * npair; t += FT, cnt++) {
        int q = t / npair, k = t - q * npair, i = i0 + q;
        int kk[2] = {k, M - k};
        cplx Xp[2];
        for (int h = 0; h < 2; h++) {
            int kv = kk[h];
            cplx V = real_post(sr, si, q * M, M, kv, F.rty);
            // Q_k = e^{-i pi k / 2N} V_k = C_k + i S_k;  C_{N-k} = -S_k
            double c = F.dcty[2 * kv], s = F.dcty[2 * kv + 1];
            double Ck = c * V.x + s * V.y;
            double Sk = c * V.y - s * V.x;
            double CNk = -Sk;
            double lx = (i < NX) ? F.lamx[i] : 1.0;
            double gk = -1.0 / (lx + F.lamy[kv]);
            double gN = (kv == 0) ? 0.0 : -1.0 / (lx + F.lamy[N - kv]);
            if (kv == 0 && mean_free && i == 0) gk = 0.0;
            if (kv == 0 && i == 0 && !mean_free) gk = 0.0;
            double Cp = Ck * gk, CNp = (kv == 0) ? 0.0 : CNk * gN;
            // V'_k = e^{+i pi k / 2N} (C'_k - i C'_{N-k})
            Xp[h].x = c * Cp + s * CNp;
            Xp[h].y = s * Cp - c * CNp;
        }
        cplx Y0 = real_pre(Xp[0], Xp[1], k, F.rty);
        cplx Y1 = real_pre(Xp[1], Xp[0], M - k, F.rty);
        Ya[cnt] = Y0;
        Yb[cnt] = Y1;
    }
    __syncthreads();
    cnt = 0;
    for (int t = threadIdx.x; t < R * npair; t += FT, cnt++) {
        int q = t / npair, k = t - q * npair;
        if (k < M) { int d = q * M + brev(k, logM); sr[d] = Ya[cnt].x; si[d] = Ya[cnt].y; }
        int k2 = M - k;
        if (k2 < M && k2 != k) { int d = q * M + brev(k2, logM); sr[d] = Yb[cnt].x; si[d] = Yb[cnt].y; }
    }
    __syncthreads();
    lds_fft(sr, si, R, M, logM, F.twy, true);
    const double invN = 1.0 / (double)N;
    for (int t = threadIdx.x; t < R * M; t += blockDim.x) {
        int m = t / R, q = t - m * R, i = i0 + q;
        if (i >= NX) continue;
        double a = sr[q * M + m] * invN, b = si[q * M + m] * invN;   // v_{2m}, v_{2m+1}
        spec[(size_t)vmap(2 * m, N) * NX + i] = a;
        spec[(size_t)vmap(2 * m + 1, N) * NX + i] = b;
    }
}

static size_t lds_bytes(int R, int N) { return (size_t)R * (size_t)(N / 2) * 2 * sizeof(double); }

// mean_free = 1: pseudo-inverse (the (0,0) mode set to zero).
extern "C" void pvk_fft_solve(const pvk_grid *g, const pvk_fft *F, const double *in, double *out, int mean_free,
                              hipStream_t s)
{
    int nbr = (g->NYM + F->rx - 1) / F->rx, nbc = (g->NXM + F->ry - 1) / F->ry;
    hipLaunchKernelGGL(k_fft_rows_fwd, dim3(nbr), dim3(FT), lds_bytes(F->rx, g->NXM), s, *g, *F,
                       (const pvk_scal *)0, in, F->spec, 1.0 / (double)g->NXM);
    hipLaunchKernelGGL(k_dct_cols, dim3(nbc), dim3(FTC), lds_bytes(F->ry, g->NYM), s, *g, *F, (const pvk_scal *)0,
                       F->spec, mean_free);
    hipLaunchKernelGGL(k_fft_rows_inv, dim3(nbr), dim3(FT), lds_bytes(F->rx, g->NXM), s, *g, *F,
                       (const pvk_scal *)0, (const double *)F->spec, out, (const double *)0, (double *)0, 0);
}

// PCG variant: skipped once the solve has converged; partials of <r, z>
// land in rd->part (one per row block) for the scalar kernel.
extern "C" void pvk_fft_solve_sc(const pvk_grid *g, const pvk_fft *F, const pvk_scal *sc, const double *in,
                                 double *out, int mean_free, hipStream_t s)
{
    int nbr = (g->NYM + F->rx - 1) / F->rx, nbc = (g->NXM + F->ry - 1) / F->ry;
    hipLaunchKernelGGL(k_fft_rows_fwd, dim3(nbr), dim3(FT), lds_bytes(F->rx, g->NXM), s, *g, *F, sc, in, F->spec,
                       1.0 / (double)g->NXM);
    hipLaunchKernelGGL(k_dct_cols, dim3(nbc), dim3(FTC), lds_bytes(F->ry, g->NYM), s, *g, *F, sc, F->spec,
                       mean_free);
    hipLaunchKernelGGL(k_fft_rows_inv, dim3(nbr), dim3(FT), lds_bytes(F->rx, g->NXM), s, *g, *F, sc,
                       (const double *)F->spec, out, (const double *)0, (double *)0, 0);
}

// Same, with the <r, z> partials fused into the inverse row transform.
extern "C" void pvk_fft_solve_dot(const pvk_grid *g, const pvk_fft *F, const pvk_scal *sc, const double *in,
                                  double *out, int mean_free, double *part, hipStream_t s)
{
    int nbr = (g->NYM + F->rx - 1) / F->rx, nbc = (g->NXM + F->ry - 1) / F->ry;
    hipLaunchKernelGGL(k_fft_rows_fwd, dim3(nbr), dim3(FT), lds_bytes(F->rx, g->NXM), s, *g, *F, sc, in, F->spec,
                       1.0 / (double)g->NXM);
    hipLaunchKernelGGL(k_dct_cols, dim3(nbc), dim3(FTC), lds_bytes(F->ry, g->NYM), s, *g, *F, sc, F->spec,
                       mean_free);
    hipLaunchKernelGGL(k_fft_rows_inv, dim3(nbr), dim3(FT), lds_bytes(F->rx, g->NXM), s, *g, *F, sc,
                       (const double *)F->spec, out, in, part, nbr);
}

extern "C" int pvk_fft_rows_blocks(const pvk_grid *g, const pvk_fft *F) { return (g->NYM + F->rx - 1) / F->rx; }

// The three stages as separate launchers (the host brackets each with timers).
extern "C" void pvk_fft_stage(const pvk_grid *g, const pvk_fft *F, const pvk_scal *sc, int stage, const double *in,
                              double *out, int mean_free, double *part, hipStream_t s)
{
    int nbr = (g->NYM + F->rx - 1) / F->rx, nbc = (g->NXM + F->ry - 1) / F->ry;
    if (stage == 0)
        hipLaunchKernelGGL(k_fft_rows_fwd, dim3(nbr), dim3(FT), lds_bytes(F->rx, g->NXM), s, *g, *F, sc, in, F->spec,
                           1.0 / (double)g->NXM);
    else if (stage == 1)
        hipLaunchKernelGGL(k_dct_cols, dim3(nbc), dim3(FTC), lds_bytes(F->ry, g->NYM), s, *g, *F, sc, F->spec,
                           mean_free);
    else
        hipLaunchKernelGGL(k_fft_rows_inv, dim3(nbr), dim3(FT), lds_bytes(F->rx, g->NXM), s, *g, *F, sc,
                           (const double *)F->spec, out, in, part, nbr);
}
