// pv_noalloc.hip -- the velocity-divergence check (Pressure.c:475-597) for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops): no handle, the caller's device arrays in, one
// device double out, and the block partials in a fixed file-scope device
// array instead of library-allocated memory.
//
// `scratch` below is a device symbol: in host code its name is not a device
// address.  The address is looked up once per device with
// hipGetSymbolAddress (with that device current) and only that pointer is
// passed to the kernels.  Every call on one GPU shares that GPU's array, so
// calls that overlap on one GPU race; each call overwrites every partial it
// reads, so nothing has to be cleared between calls.
#include <hip/hip_runtime.h>
#include <math.h>
#include <mutex>
#include "../../include/parties_vof.h"

#define NA_THREADS 256
#define NA_BLOCKS 1024                 // partials: 1024 doubles = 8 KiB per GPU
#define NA_MAX_DEV 64

static __device__ __align__(256) double scratch[NA_BLOCKS];

static std::mutex g_mu;
static double *g_scratch[NA_MAX_DEV];  // device address of `scratch` on each GPU

extern "C" int pv_set_error(int code, const char *fmt, ...);

// max over the block's cells of |(u_{i+1} - u_i) idx + (v_{j+1} - v_j) idy|,
// grid-stride over the NXM x NYM cells; one partial per block
__global__ void __launch_bounds__(NA_THREADS) k_div_partials(const double *__restrict__ u,
                                                            const double *__restrict__ v, int NXM, int NYM,
                                                            int pitch, double idx, double idy, double *part)
{
    __shared__ double sh[NA_THREADS / 64];
    const long long ncell = (long long)NXM * NYM;
    double m = 0.0;
    for (long long c = (long long)blockIdx.x * NA_THREADS + threadIdx.x; c < ncell;
         c += (long long)gridDim.x * NA_THREADS) {
        const int j = (int)(c / NXM), i = (int)(c - (long long)j * NXM);
        const ptrdiff_t o = (ptrdiff_t)j * pitch + i;
        const double dudx = (u[o + 1] - u[o]) * idx;
        const double dvdy = (v[o + pitch] - v[o]) * idy;
        m = fmax(m, fabs(dudx + dvdy));
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = sh[0];
        for (int w = 1; w < NA_THREADS / 64; w++) s = fmax(s, sh[w]);
        part[blockIdx.x] = s;
    }
}

__global__ void __launch_bounds__(NA_THREADS) k_div_final(const double *__restrict__ part, int nb, double *out)
{
    __shared__ double sh[NA_THREADS];
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += NA_THREADS) s = fmax(s, part[b]);
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = NA_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// the device address of `scratch` on the current device, looked up once
static int scratch_ptr(double **out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NA_MAX_DEV)
        return pv_set_error(PV_ENODEV, "no usable current HIP device (%d)", dev);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return pv_set_error(PV_EHIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev,
                                hipGetErrorString(e));
        g_scratch[dev] = (double *)p;
    }
    *out = g_scratch[dev];
    return PV_OK;
}

extern "C" int pv_Pressure_compute_velocity_divergence_noalloc(const double *u, const double *v, int NXM, int NYM,
                                                                int pitch, double idx, double idy,
                                                                double *div_max, void *stream)
{
    double *part = nullptr;
    int rc, nb;
    long long ncell;
    if (!u || !v || !div_max) return pv_set_error(PV_EINVAL, "divergence (no allocation): null pointer");
    if (NXM < 1 || NYM < 1 || NXM > PV_NOALLOC_MAX_N || NYM > PV_NOALLOC_MAX_N)
        return pv_set_error(PV_EINVAL, "divergence (no allocation): %d x %d outside 1 .. %d per direction", NXM,
                            NYM, PV_NOALLOC_MAX_N);
    if (pitch < NXM + 1) return pv_set_error(PV_EINVAL, "divergence (no allocation): pitch %d < NXM + 1", pitch);
    if ((rc = scratch_ptr(&part)) != PV_OK) return rc;
    ncell = (long long)NXM * NYM;
    nb = (int)((ncell + NA_THREADS - 1) / NA_THREADS);
    if (nb > NA_BLOCKS) nb = NA_BLOCKS;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_div_partials, dim3(nb), dim3(NA_THREADS), 0, s, u, v, NXM, NYM, pitch, idx, idy, part);
    hipLaunchKernelGGL(k_div_final, dim3(1), dim3(NA_THREADS), 0, s, (const double *)part, nb, div_max);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return pv_set_error(PV_EHIP, "divergence (no allocation) launch failed: %s", hipGetErrorString(e));
    return PV_OK;
}
