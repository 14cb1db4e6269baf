// pv_fftk.h -- radix-8/4/2 Stockham FFT of length M (a power of two, 4..2048)
// held in LDS as separate real and imaginary arrays.
//
// Pass with radix R and sub-transform size Ns (Ns = 1, R, R*R', ...): the
// butterfly bb < M/R reads x[bb + r M/R], r < R, multiplies element r by
// W^(k r), k = bb mod Ns, W = exp(-+2 pi i / (Ns R)), takes an R-point DFT in
// registers and writes y[(bb / Ns) Ns R + k + r Ns].  Input and output are
// in natural order (no bit reversal).  Passes: radix 8 while at least three
// factors of two remain, then one radix-4 or radix-2 pass.  With T = M/8
// threads per transform every thread holds 8 complex values in every pass.
//
// The twiddle table tw holds exp(-2 pi i t / M) for t < M as (cos, sin)
// pairs; W^(k r) of a pass is entry k r M / (Ns R).  The inverse transform
// uses conjugate twiddles and is unscaled.
//
// The load and store halves are __host__ __device__ so the same code is
// checked on the CPU against a direct DFT.
#ifndef PV_FFTK_H
#define PV_FFTK_H
#include <hip/hip_runtime.h>

#define SK_SQRT_HALF 0.70710678118654752440

// (x + i y) times -i (forward) or +i (inverse)
template <bool INV>
__host__ __device__ __forceinline__ void sk_rot(double &x, double &y)
{
    double t = x;
    if (INV) { x = -y; y = t; }
    else     { x = y;  y = -t; }
}

template <bool INV>
__host__ __device__ __forceinline__ void sk_dft2(double *vr, double *vi, int a, int b)
{
    double tr = vr[a] - vr[b], ti = vi[a] - vi[b];
    vr[a] = vr[a] + vr[b]; vi[a] = vi[a] + vi[b];
    vr[b] = tr; vi[b] = ti;
}

// 4-point DFT of v[o], v[o+s], v[o+2s], v[o+3s], results in natural order in place
template <bool INV>
__host__ __device__ __forceinline__ void sk_dft4(double *vr, double *vi, int o, int s)
{
    const int i0 = o, i1 = o + s, i2 = o + 2 * s, i3 = o + 3 * s;
    double ar = vr[i0] + vr[i2], ai = vi[i0] + vi[i2];
    double br = vr[i0] - vr[i2], bi = vi[i0] - vi[i2];
    double cr = vr[i1] + vr[i3], ci = vi[i1] + vi[i3];
    double dr = vr[i1] - vr[i3], di = vi[i1] - vi[i3];
    sk_rot<INV>(dr, di);                       // omega (a1 - a3), omega = -+i
    vr[i0] = ar + cr; vi[i0] = ai + ci;
    vr[i2] = ar - cr; vi[i2] = ai - ci;
    vr[i1] = br + dr; vi[i1] = bi + di;
    vr[i3] = br - dr; vi[i3] = bi - di;
}

// 8-point DFT of v[0..7] in place, natural order: X_k = E_k + w^k O_k,
// X_{k+4} = E_k - w^k O_k with E, O the 4-point DFTs of even and odd entries.
template <bool INV>
__host__ __device__ __forceinline__ void sk_dft8(double *vr, double *vi)
{
    sk_dft4<INV>(vr, vi, 0, 2);                // E_k at 0, 2, 4, 6
    sk_dft4<INV>(vr, vi, 1, 2);                // O_k at 1, 3, 5, 7
    const double h = SK_SQRT_HALF;
    double o1r = vr[3], o1i = vi[3], o2r = vr[5], o2i = vi[5], o3r = vr[7], o3i = vi[7];
    // w = exp(-+ i pi / 4)
    double w1r = (o1r + (INV ? -o1i : o1i)) * h, w1i = (o1i + (INV ? o1r : -o1r)) * h;
    double w2r = o2r, w2i = o2i;
    sk_rot<INV>(w2r, w2i);
    double w3r = (-o3r + (INV ? -o3i : o3i)) * h, w3i = (-o3i + (INV ? o3r : -o3r)) * h;
    double e0r = vr[0], e0i = vi[0], e1r = vr[2], e1i = vi[2], e2r = vr[4], e2i = vi[4], e3r = vr[6], e3i = vi[6];
    double o0r = vr[1], o0i = vi[1];
    vr[0] = e0r + o0r; vi[0] = e0i + o0i;
    vr[4] = e0r - o0r; vi[4] = e0i - o0i;
    vr[1] = e1r + w1r; vi[1] = e1i + w1i;
    vr[5] = e1r - w1r; vi[5] = e1i - w1i;
    vr[2] = e2r + w2r; vi[2] = e2i + w2i;
    vr[6] = e2r - w2r; vi[6] = e2i - w2i;
    vr[3] = e3r + w3r; vi[3] = e3i + w3i;
    vr[7] = e3r - w3r; vi[7] = e3i - w3i;
}

// Read butterfly bb of a radix-R pass (R = 2^LR), twiddle, DFT: v[0..R).
// All sizes are powers of two and passed as base-2 logarithms, so index
// arithmetic is shifts and masks.
template <int LR, bool INV>
__host__ __device__ __forceinline__ void sk_load(const double *sr, const double *si, int bb, int logM, int logNs,
                                                 const double *tw, double *vr, double *vi)
{
    constexpr int R = 1 << LR;
    const int stride = 1 << (logM - LR), k = bb & ((1 << logNs) - 1), step = k << (logM - logNs - LR);
#pragma unroll
    for (int r = 0; r < R; r++) {
        vr[r] = sr[bb + r * stride];
        vi[r] = si[bb + r * stride];
    }
    if (logNs > 0) {
#pragma unroll
        for (int r = 1; r < R; r++) {
            const int t = r * step;
            const double c = tw[2 * t], s = INV ? -tw[2 * t + 1] : tw[2 * t + 1];
            const double xr = vr[r], xi = vi[r];
            vr[r] = xr * c - xi * s;
            vi[r] = xr * s + xi * c;
        }
    }
    if (R == 8) sk_dft8<INV>(vr, vi);
    else if (R == 4) sk_dft4<INV>(vr, vi, 0, 1);
    else sk_dft2<INV>(vr, vi, 0, 1);
}

template <int LR>
__host__ __device__ __forceinline__ void sk_store(double *sr, double *si, int bb, int logNs, const double *vr,
                                                  const double *vi)
{
    constexpr int R = 1 << LR;
    const int Ns = 1 << logNs;
    const int base = ((bb >> logNs) << (logNs + LR)) + (bb & (Ns - 1));
#pragma unroll
    for (int r = 0; r < R; r++) {
        sr[base + r * Ns] = vr[r];
        si[base + r * Ns] = vi[r];
    }
}

// One radix-R pass over a transform whose T threads have lane ids 0..T-1;
// every thread of the block calls it (barriers inside).  bb >= M / R idle.
template <int LR, bool INV>
__device__ __forceinline__ void sk_pass(double *sr, double *si, int logM, int logNs, const double *tw, int lane,
                                        int T)
{
    constexpr int U = 8 >> LR;
    const int nbf = 1 << (logM - LR);
    double vr[8], vi[8];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int bb = lane + u * T;
        if (bb < nbf) sk_load<LR, INV>(sr, si, bb, logM, logNs, tw, vr + (u << LR), vi + (u << LR));
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int bb = lane + u * T;
        if (bb < nbf) sk_store<LR>(sr, si, bb, logNs, vr + (u << LR), vi + (u << LR));
    }
    __syncthreads();
}

// Whole transform of length M = 2^logM; the caller has synchronised after
// filling sr / si and reads the result after this returns.
template <bool INV>
__device__ __forceinline__ void sk_fft(double *sr, double *si, int logM, const double *tw, int lane, int T)
{
    int logNs = 0, rem = logM;
    for (; rem >= 3; rem -= 3, logNs += 3) sk_pass<3, INV>(sr, si, logM, logNs, tw, lane, T);
    if (rem == 2) sk_pass<2, INV>(sr, si, logM, logNs, tw, lane, T);
    else if (rem == 1) sk_pass<1, INV>(sr, si, logM, logNs, tw, lane, T);
}

// ---------------------------------------------------------------------------
// Mixed radix: lengths M = 2^a 3^b with 8 | M.  The same Stockham passes with
// general (not power-of-two) sub-transform sizes Ns: radix 8 while three
// factors of two remain, then a radix-4 or radix-2 pass, then one radix-3
// pass per factor of three.  T = M / 8 threads per transform; a radix-R pass
// has M / R butterflies, so a thread takes up to 8 / R of them (3 for R = 3).
// Threads whose transform index is out of range (act = false) only take part
// in the barriers.
// ---------------------------------------------------------------------------

// 3-point DFT of v[0..2] in place: X_1,2 = a - (b + c) / 2 -+ i (sqrt 3 / 2)(b - c)
// (forward: exp(-2 pi i / 3); the inverse swaps X_1 and X_2)
template <bool INV>
__host__ __device__ __forceinline__ void sk_dft3(double *vr, double *vi)
{
    const double h = 0.86602540378443864676;          // sqrt(3) / 2
    const double sr_ = vr[1] + vr[2], si_ = vi[1] + vi[2];
    const double tr = vr[0] - 0.5 * sr_, ti = vi[0] - 0.5 * si_;
    const double dr = h * (vr[1] - vr[2]), di = h * (vi[1] - vi[2]);
    vr[0] += sr_; vi[0] += si_;
    if (INV) { vr[1] = tr - di; vi[1] = ti + dr; vr[2] = tr + di; vi[2] = ti - dr; }
    else     { vr[1] = tr + di; vi[1] = ti - dr; vr[2] = tr - di; vi[2] = ti + dr; }
}

template <int R, bool INV>
__host__ __device__ __forceinline__ void sg_load(const double *sr, const double *si, int bb, int M, int Ns,
                                                 const double *tw, double *vr, double *vi)
{
    const int stride = M / R, k = bb % Ns, step = (M / (Ns * R)) * k;
#pragma unroll
    for (int r = 0; r < R; r++) {
        vr[r] = sr[bb + r * stride];
        vi[r] = si[bb + r * stride];
    }
    if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; r++) {
            const int t = r * step;
            const double c = tw[2 * t], s = INV ? -tw[2 * t + 1] : tw[2 * t + 1];
            const double xr = vr[r], xi = vi[r];
            vr[r] = xr * c - xi * s;
            vi[r] = xr * s + xi * c;
        }
    }
    if (R == 8) sk_dft8<INV>(vr, vi);
    else if (R == 4) sk_dft4<INV>(vr, vi, 0, 1);
    else if (R == 3) sk_dft3<INV>(vr, vi);
    else sk_dft2<INV>(vr, vi, 0, 1);
}

template <int R>
__host__ __device__ __forceinline__ void sg_store(double *sr, double *si, int bb, int Ns, const double *vr,
                                                  const double *vi)
{
    const int k = bb % Ns, base = (bb / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; r++) {
        sr[base + r * Ns] = vr[r];
        si[base + r * Ns] = vi[r];
    }
}

template <int R, bool INV>
__device__ __forceinline__ void sg_pass(double *sr, double *si, int M, int Ns, const double *tw, int lane, int T,
                                        bool act)
{
    constexpr int U = (R == 3) ? 3 : 8 / R;
    const int nbf = M / R;
    double vr[U * R], vi[U * R];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int bb = lane + u * T;
        if (act && bb < nbf) sg_load<R, INV>(sr, si, bb, M, Ns, tw, vr + u * R, vi + u * R);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int bb = lane + u * T;
        if (act && bb < nbf) sg_store<R>(sr, si, bb, Ns, vr + u * R, vi + u * R);
    }
    __syncthreads();
}

// Whole transform of length M = 2^a 3^b (8 | M)
template <bool INV>
__device__ __forceinline__ void sg_fft(double *sr, double *si, int M, const double *tw, int lane, int T, bool act)
{
    int Ns = 1, two = M;
    while (two % 3 == 0) two /= 3;                   // the power-of-two part
    for (; two % 8 == 0; two /= 8, Ns *= 8) sg_pass<8, INV>(sr, si, M, Ns, tw, lane, T, act);
    if (two == 4) { sg_pass<4, INV>(sr, si, M, Ns, tw, lane, T, act); Ns *= 4; }
    else if (two == 2) { sg_pass<2, INV>(sr, si, M, Ns, tw, lane, T, act); Ns *= 2; }
    for (; Ns < M; Ns *= 3) sg_pass<3, INV>(sr, si, M, Ns, tw, lane, T, act);
}
#endif
