// pv_learn.hip -- glue between the pressure solve and the learned
// preconditioner network (csrc/pv_cnn.hip): feature maps, the corrected
// preconditioner output, training targets and crop / batch assembly.
//
// For a residual r with FFT-preconditioned z = M_fft r the network input at
// cell (i, j) is (r / s_r, z / s_z, log10(rho) / 1.5 + 1), s = RMS over the
// cells, as an NHWC image of NYM rows and NXM columns (x periodic).  The
// network predicts a source correction c: the learned approximation of
// A^-1 r is M_fft(r + s_r c).  With e = A^-1 r known from a converged solve
// the training target is (L e - r) / s_r, L the constant-coefficient
// Laplacian that M_fft inverts; it is nonzero only where the density differs
// from the liquid's (L - A = div((1 - beta) grad)), a local quantity, and
// M_fft damps grid-scale noise of the network output as 1/k^2.
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

// out = r + s_r c: the right-hand side whose constant-density solve M_fft
// gives the learned approximation of A^-1 r
__global__ void __launch_bounds__(NTH) k_learn_addr(pvk_grid g, const pvk_scal *sc, const double *r, const float *c,
                                                   double *out)
{
    const int P = g.P, j = blockIdx.x;
    const double s = sc->red[5];
    for (int i = threadIdx.x; i < g.NXM; i += NTH)
        AT(out, i, j) = AT(r, i, j) + s * (double)c[(size_t)j * g.NXM + i];
}

// flexible-CG partials <r, z> -> rzp and <Ad, z> -> zadp, one per row
__global__ void __launch_bounds__(NTH) k_learn_dots(pvk_grid g, const double *z, const double *r, const double *Ad,
                                                   double *rzp, double *zadp)
{
    __shared__ double sh[NTH];
    const int P = g.P, j = blockIdx.x;
    double a0 = 0.0, a1 = 0.0;
    for (int i = threadIdx.x; i < g.NXM; i += NTH) {
        const double v = AT(z, i, j);
        a0 += AT(r, i, j) * v;
        a1 += AT(Ad, i, j) * v;
    }
    block_store_partial(a0, rzp, j, sh);
    block_store_partial(a1, zadp, j, sh);
}

__global__ void __launch_bounds__(NTH) k_learn_copy(pvk_grid g, const double *src, double *dst)
{
    const int P = g.P, j = blockIdx.x;
    for (int i = threadIdx.x; i < g.NXM; i += NTH) AT(dst, i, j) = AT(src, i, j);
}

__global__ void __launch_bounds__(NTH) k_learn_axpy(pvk_grid g, const double *src, double *dst)
{
    const int P = g.P, j = blockIdx.x;
    for (int i = threadIdx.x; i < g.NXM; i += NTH) AT(dst, i, j) += AT(src, i, j);
}

// target t = (L e - r) / s_r, e = phistar - phik, L the constant-coefficient
// 5-point Laplacian that M_fft inverts (x periodic, zero-gradient y walls):
// the source correction that makes M_fft(r + s_r t) = e.
__global__ void __launch_bounds__(NTH) k_learn_target(pvk_grid g, const double *phistar, const double *phik,
                                                     const double *r, double sr, float *tgt)
{
    const int P = g.P, NXM = g.NXM, NYM = g.NYM, j = blockIdx.x;
    const double idx2 = g.idx_c[1] * g.idx_c[1], idy2 = g.idy_c[1] * g.idy_c[1], is = (sr > 0.0) ? 1.0 / sr : 0.0;
    const int jn = (j + 1 == NYM) ? j : j + 1, js = (j == 0) ? 0 : j - 1;
    for (int i = threadIdx.x; i < NXM; i += NTH) {
        const int ie = (i + 1 == NXM) ? 0 : i + 1, iw = (i == 0) ? NXM - 1 : i - 1;
#define EV(ii, jj) (AT(phistar, ii, jj) - AT(phik, ii, jj))
        const double e0 = EV(i, j);
        const double Le = (EV(ie, j) - 2.0 * e0 + EV(iw, j)) * idx2 + (EV(i, jn) - 2.0 * e0 + EV(i, js)) * idy2;
#undef EV
        tgt[(size_t)j * NXM + i] = (float)((Le - AT(r, i, j)) * is);
    }
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
extern "C" void pvk_learn_addr(const pvk_grid *g, const pvk_scal *sc, const double *r, const float *c, double *out,
                               hipStream_t s)
{
    hipLaunchKernelGGL(k_learn_addr, dim3(g->NYM), dim3(NTH), 0, s, *g, sc, r, c, out);
}
extern "C" void pvk_learn_dots(const pvk_grid *g, const double *z, const double *r, const double *Ad, double *rzp,
                               double *zadp, hipStream_t s)
{
    hipLaunchKernelGGL(k_learn_dots, dim3(g->NYM), dim3(NTH), 0, s, *g, z, r, Ad, rzp, zadp);
}
extern "C" void pvk_learn_copy(const pvk_grid *g, const double *src, double *dst, hipStream_t s)
{
    hipLaunchKernelGGL(k_learn_copy, dim3(g->NYM), dim3(NTH), 0, s, *g, src, dst);
}
extern "C" void pvk_learn_axpy(const pvk_grid *g, const double *src, double *dst, hipStream_t s)
{
    hipLaunchKernelGGL(k_learn_axpy, dim3(g->NYM), dim3(NTH), 0, s, *g, src, dst);
}
extern "C" void pvk_learn_target(const pvk_grid *g, const pvk_scal *sc, const double *phistar, const double *phik,
                                 const double *r, double sr, float *tgt, hipStream_t s)
{
    (void)sc;
    hipLaunchKernelGGL(k_learn_target, dim3(g->NYM), dim3(NTH), 0, s, *g, phistar, phik, r, sr, tgt);
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
