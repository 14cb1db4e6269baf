// pv_block.h -- block-level helpers of the fused solver kernels (pv_pcg.hip,
// pv_cg.hip): 256-thread blocks that each own a band of rows.
//
// Scalars of a Krylov iteration (alpha, beta, stop tests) are never written
// by one block and read by another inside one kernel.  Each block stores one
// partial sum per reduction; the next kernel's blocks all re-add those few
// hundred partials in the same order, so every block holds bit-identical
// scalars and no one-block scalar kernel sits between the sweeps.
#ifndef PV_BLOCK_H
#define PV_BLOCK_H
#include <hip/hip_runtime.h>

#define NTH 256

// This thread's shares of up to three partial sums: sa = sum of pa[b] over
// b = t, t + NTH, ... < na in increasing b (likewise sb, sc).  The partials
// are loaded in batches of PVB per array, all loads of a batch in flight
// before its first add (one memory latency per batch, not one per partial);
// an array with n = 0 is not read.
#define PVB 4
__device__ __forceinline__ void thread_sums(const double *pa, int na, const double *pb, int nb, const double *pc,
                                            int nc, double &sa, double &sb, double &sc)
{
    const int t = threadIdx.x, n = max(na, max(nb, nc));
    sa = sb = sc = 0.0;
    for (int b0 = t; b0 < n; b0 += PVB * NTH) {
        double va[PVB], vb[PVB], vc[PVB];
#pragma unroll
        for (int k = 0; k < PVB; k++) {
            const int b = b0 + k * NTH;
            va[k] = vb[k] = vc[k] = 0.0;
            if (na > 0) va[k] = pa[min(b, na - 1)];
            if (nb > 0) vb[k] = pb[min(b, nb - 1)];
            if (nc > 0) vc[k] = pc[min(b, nc - 1)];
        }
#pragma unroll
        for (int k = 0; k < PVB; k++) {
            const int b = b0 + k * NTH;
            if (b < na) sa += va[k];
            if (b < nb) sb += vb[k];
            if (b < nc) sc += vc[k];
        }
    }
}

// Sum of n partials, same order in every block; all threads get the result.
__device__ __forceinline__ double block_sum(const double *part, int n, double *sh)
{
    const int t = threadIdx.x;
    double s, u0, u1;
    thread_sums(part, n, nullptr, 0, nullptr, 0, s, u0, u1);
    sh[t] = s;
    __syncthreads();
    for (int w = NTH / 2; w > 0; w >>= 1) {
        if (t < w) sh[t] += sh[t + w];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

// Two sums of partials in one pass (one tree, half the barriers of two
// block_sum calls), each added in exactly block_sum's order; sh holds
// 2 * NTH doubles.
__device__ __forceinline__ void block_sum2(const double *pa, int na, const double *pb, int nb, double *sh,
                                           double &ra, double &rb)
{
    const int t = threadIdx.x;
    double sa, sb, u;
    thread_sums(pa, na, pb, nb, nullptr, 0, sa, sb, u);
    sh[t] = sa;
    sh[NTH + t] = sb;
    __syncthreads();
    for (int w = NTH / 2; w > 0; w >>= 1) {
        if (t < w) {
            sh[t] += sh[t + w];
            sh[NTH + t] += sh[NTH + t + w];
        }
        __syncthreads();
    }
    ra = sh[0];
    rb = sh[NTH];
    __syncthreads();
}

// Three sums of partials in one pass (the loads of all three in flight
// together, one tree), each added in exactly block_sum's order; sh holds
// 3 * NTH doubles.
__device__ __forceinline__ void block_sum3(const double *pa, int na, const double *pb, int nb, const double *pc,
                                           int nc, double *sh, double &ra, double &rb, double &rc)
{
    const int t = threadIdx.x;
    double sa, sb, s3;
    thread_sums(pa, na, pb, nb, pc, nc, sa, sb, s3);
    sh[t] = sa;
    sh[NTH + t] = sb;
    sh[2 * NTH + t] = s3;
    __syncthreads();
    for (int w = NTH / 2; w > 0; w >>= 1) {
        if (t < w) {
            sh[t] += sh[t + w];
            sh[NTH + t] += sh[NTH + t + w];
            sh[2 * NTH + t] += sh[2 * NTH + t + w];
        }
        __syncthreads();
    }
    ra = sh[0];
    rb = sh[NTH];
    rc = sh[2 * NTH];
    __syncthreads();
}

// Block total of v (wave shuffles, then the 4 wave sums in order) -> part[slot].
__device__ __forceinline__ void block_store_partial(double v, double *part, int slot, double *sh)
{
    const int t = threadIdx.x;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((t & 63) == 0) sh[t >> 6] = v;
    __syncthreads();
    if (t == 0) {
        double s = sh[0];
        for (int w = 1; w < NTH / 64; w++) s += sh[w];
        part[slot] = s;
    }
    __syncthreads();
}

// Workgroups are dealt to the 8 XCDs round robin (block b runs on XCD b % 8).
// Give each XCD a contiguous run of bands (or tiles, in row-major order) so
// the data a band's stencil shares with its neighbours stays in that XCD's
// L2: XCD x owns q + (x < rem) consecutive indices from x q + min(x, rem),
// q = nblocks / 8, rem = nblocks % 8 (a bijection on [0, nblocks)).
__device__ __forceinline__ int xcd_band(int b, int nblocks)
{
    const int x = b & 7, q = nblocks >> 3, rem = nblocks & 7;
    return x * q + (x < rem ? x : rem) + (b >> 3);
}
#endif
