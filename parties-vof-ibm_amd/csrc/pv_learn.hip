// pv_learn.hip -- glue between the pressure solve and the learned
// preconditioner network (csrc/pv_cnn.hip): feature maps, the corrected
// preconditioner output, training targets and crop / batch assembly.
//
// For a residual r with FFT-preconditioned z = M_fft r the network input at
// cell (i, j) is (r / s_r, z / s_z, log10(rho) / 1.5 + 1), s = RMS over the
// cells, as an NHWC image of NYM rows and NXM columns (x periodic).  The
// learned preconditioner returns z + s_z * c with c the network output; the
// training target is (e - mean(e) - z) / s_z where e = A^-1 r is known from
// a converged solve.
#include <hip/hip_runtime.h>
#include <math.h>
#include "pv_kernels.h"
#include "pv_block.h"

#define AT(a, i, j) (a)[(ptrdiff_t)(j) * P + (i)]

// per-row partials of sum r^2, sum z^2, sum e (e = a - b) ; one block per row
__global__ void __launch_bounds__(NTH) k_learn_rowsums(pvk_grid g, const double *r, const double *z,
                                                      const double *a, const double *b, double *part)
{
    __shared__ double sh[NTH];
    const int P = g.P, j = blockIdx.x;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int i = threadIdx.x; i < g.NXM; i += NTH) {
        if (r) { const double v = AT(r, i, j); s0 += v * v; }
        if (z) { const double v = AT(z, i, j); s1 += v * v; }
        if (a) s2 += AT(a, i, j) - (b ? AT(b, i, j) : 0.0);
    }
    block_store_partial(s0, part, j, sh);
    block_store_partial(s1, part + g.NYM, j, sh);
    block_store_partial(s2, part + 2 * g.NYM, j, sh);
}

// sums of the row partials: sc->red[5] = s_r, red[6] = s_z (RMS), red[7] = mean(e)
__global__ void __launch_bounds__(NTH) k_learn_stats(pvk_grid g, pvk_scal *sc, const double *part)
{
    __shared__ double sh[NTH];
    const double n = (double)g.NXM * (double)g.NYM;
    const double rr = block_sum(part, g.NYM, sh);
    const double zz = block_sum(part + g.NYM, g.NYM, sh);
    const double es = block_sum(part + 2 * g.NYM, g.NYM, sh);
    if (threadIdx.x == 0) {
        sc->red[5] = sqrt(rr / n);
        sc->red[6] = sqrt(zz / n);
        sc->red[7] = es / n;
    }
}

__global__ void __launch_bounds__(NTH) k_learn_feat(pvk_grid g, const pvk_scal *sc, const double *r, const double *z,
                                                   const double *rho, float *in3)
{
    const int P = g.P, j = blockIdx.x;
    const double ir = (sc->red[5] > 0.0) ? 1.0 / sc->red[5] : 0.0;
    const double iz = (sc->red[6] > 0.0) ? 1.0 / sc->red[6] : 0.0;
    for (int i = threadIdx.x; i < g.NXM; i += NTH) {
        float *o = in3 + ((size_t)j * g.NXM + i) * 3;
        o[0] = (float)(AT(r, i, j) * ir);
        o[1] = (float)(AT(z, i, j) * iz);
        o[2] = (float)(log10(AT(rho, i, j)) / 1.5 + 1.0);
    }
}

// out = z + s_z * c   (out may alias z)
__global__ void __launch_bounds__(NTH) k_learn_combine(pvk_grid g, const pvk_scal *sc, const double *z,
                                                      const float *c, double *out)
{
    const int P = g.P, j = blockIdx.x;
    const double s = sc->red[6];
    for (int i = threadIdx.x; i < g.NXM; i += NTH)
        AT(out, i, j) = AT(z, i, j) + s * (double)c[(size_t)j * g.NXM + i];
}

// learned preconditioner output z <- z + s_z c with the flexible-CG partials
// <r, z> -> rzp and <Ad, z> -> zadp (one block and one partial per row)
__global__ void __launch_bounds__(NTH) k_learn_combine_dots(pvk_grid g, const pvk_scal *sc, double *z, const float *c,
                                                           const double *r, const double *Ad, double *rzp,
                                                           double *zadp)
{
    __shared__ double sh[NTH];
    const int P = g.P, j = blockIdx.x;
    const double s = sc->red[6];
    double a0 = 0.0, a1 = 0.0;
    for (int i = threadIdx.x; i < g.NXM; i += NTH) {
        const double v = AT(z, i, j) + s * (double)c[(size_t)j * g.NXM + i];
        AT(z, i, j) = v;
        a0 += AT(r, i, j) * v;
        a1 += AT(Ad, i, j) * v;
    }
    block_store_partial(a0, rzp, j, sh);
    block_store_partial(a1, zadp, j, sh);
}

// target t = (e - mean(e) - z) / s_z with e = phistar - phik
__global__ void __launch_bounds__(NTH) k_learn_target(pvk_grid g, const pvk_scal *sc, const double *phistar,
                                                     const double *phik, const double *z, float *tgt)
{
    const int P = g.P, j = blockIdx.x;
    const double iz = (sc->red[6] > 0.0) ? 1.0 / sc->red[6] : 0.0, em = sc->red[7];
    for (int i = threadIdx.x; i < g.NXM; i += NTH)
        tgt[(size_t)j * g.NXM + i] =
            (float)((AT(phistar, i, j) - (phik ? AT(phik, i, j) : 0.0) - em - AT(z, i, j)) * iz);
}

// crop n of side cs at origin org[2n], org[2n+1] (x, y) -> dataset slot base + n
__global__ void k_learn_crop(const float *in3, const float *tgt, int NXM, int cs, const int *org, int base,
                             float *ds_in, float *ds_tgt)
{
    const int n = blockIdx.y, p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= cs * cs) return;
    const int y = p / cs, x = p - y * cs;
    const size_t src = (size_t)(org[2 * n + 1] + y) * NXM + org[2 * n] + x;
    const size_t dst = (size_t)(base + n) * cs * cs + p;
    ds_in[3 * dst + 0] = in3[3 * src + 0];
    ds_in[3 * dst + 1] = in3[3 * src + 1];
    ds_in[3 * dst + 2] = in3[3 * src + 2];
    ds_tgt[dst] = tgt[src];
}

// batch b <- dataset sample idx[b]
__global__ void k_learn_gather(const float *ds_in, const float *ds_tgt, int cs, const int *idx, float *bin,
                               float *btgt)
{
    const int b = blockIdx.y, p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= cs * cs) return;
    const size_t src = (size_t)idx[b] * cs * cs + p, dst = (size_t)b * cs * cs + p;
    bin[3 * dst + 0] = ds_in[3 * src + 0];
    bin[3 * dst + 1] = ds_in[3 * src + 1];
    bin[3 * dst + 2] = ds_in[3 * src + 2];
    btgt[dst] = ds_tgt[src];
}

extern "C" void pvk_learn_stats(const pvk_grid *g, pvk_scal *sc, const double *r, const double *z, const double *a,
                                const double *b, double *part, hipStream_t s)
{
    hipLaunchKernelGGL(k_learn_rowsums, dim3(g->NYM), dim3(NTH), 0, s, *g, r, z, a, b, part);
    hipLaunchKernelGGL(k_learn_stats, dim3(1), dim3(NTH), 0, s, *g, sc, (const double *)part);
}
extern "C" void pvk_learn_feat(const pvk_grid *g, const pvk_scal *sc, const double *r, const double *z,
                               const double *rho, float *in3, hipStream_t s)
{
    hipLaunchKernelGGL(k_learn_feat, dim3(g->NYM), dim3(NTH), 0, s, *g, sc, r, z, rho, in3);
}
extern "C" void pvk_learn_combine(const pvk_grid *g, const pvk_scal *sc, const double *z, const float *c,
                                  double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_learn_combine, dim3(g->NYM), dim3(NTH), 0, s, *g, sc, z, c, out);
}
extern "C" void pvk_learn_combine_dots(const pvk_grid *g, const pvk_scal *sc, double *z, const float *c,
                                       const double *r, const double *Ad, double *rzp, double *zadp, hipStream_t s)
{
    hipLaunchKernelGGL(k_learn_combine_dots, dim3(g->NYM), dim3(NTH), 0, s, *g, sc, z, c, r, Ad, rzp, zadp);
}
extern "C" void pvk_learn_target(const pvk_grid *g, const pvk_scal *sc, const double *phistar, const double *phik,
                                 const double *z, float *tgt, hipStream_t s)
{
    hipLaunchKernelGGL(k_learn_target, dim3(g->NYM), dim3(NTH), 0, s, *g, sc, phistar, phik, z, tgt);
}
extern "C" void pvk_learn_crop(const float *in3, const float *tgt, int NXM, int cs, const int *org, int n,
                               int base, float *ds_in, float *ds_tgt, hipStream_t s)
{
    hipLaunchKernelGGL(k_learn_crop, dim3((cs * cs + 255) / 256, n), dim3(256), 0, s, in3, tgt, NXM, cs, org, base,
                       ds_in, ds_tgt);
}
extern "C" void pvk_learn_gather(const float *ds_in, const float *ds_tgt, int cs, const int *idx, int B, float *bin,
                                 float *btgt, hipStream_t s)
{
    hipLaunchKernelGGL(k_learn_gather, dim3((cs * cs + 255) / 256, B), dim3(256), 0, s, ds_in, ds_tgt, cs, idx, bin,
                       btgt);
}
