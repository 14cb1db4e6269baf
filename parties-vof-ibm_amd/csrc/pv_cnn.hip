// pv_cnn.hip -- the learned pressure preconditioner: a 6-layer 3x3
// convolutional network (3 -> 32 -> 32 -> 32 -> 32 -> 32 -> 1 channels, ReLU
// between layers) with BF16 operands and FP32 accumulation on the CDNA4
// matrix cores, trained on the GPU by the kernels below.
//
// Tensors are NHWC: batch b, row y, column x, channel c fastest.  The image
// side lengths are multiples of 16.  Padding: zero outside rows [0, H);
// columns wrap when px != 0 (the x-periodic domain) and are zero-padded
// otherwise (training crops).
//
// Parameter vector (FP32 master copy, offsets in floats):
//   layer 1   W1[co 32][tap 9][ci 3]            0 ..   863, b1[32] 864 .. 895
//   layer k   Wk[tap 9][co 32][ci 32], bk[32]   896 + (k-2) 9248, k = 2..5
//   layer 6   W6[tap 9][ci 32] + b6             37888 .. 38176
// tap = 3 ky + kx reads input pixel (y + ky - 1, x + kx - 1).
//
// Layers 2..5 run on v_mfma_f32_16x16x16_bf16: a block computes a 16 x 16
// pixel tile for all 32 output channels; wave w owns tile rows 4w .. 4w+3.
// GEMM view per row and tap: D[pixel 16][co 16] += A[pixel][ci 16] B[ci][co].
// Operand lanes (16x16x16): lane l holds A[l % 16][4 (l / 16) + q],
// B[4 (l / 16) + q][l % 16], D[4 (l / 16) + q][l % 16], q = 0..3.
#include <hip/hip_runtime.h>
#include <math.h>
#include "pv_kernels.h"

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef unsigned short bf16;

#define CNN_C 32
#define TILE 16
#define HALO 18
#define PSTR 40   // LDS shorts per pixel (32 channels + 8 pad: spreads lanes over banks)

__device__ __forceinline__ bf16 f2bf(float f)
{
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);     // round to nearest even (finite inputs)
    return (bf16)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16 h) { return __uint_as_float(((unsigned)h) << 16); }
__device__ __forceinline__ bool bfpos(bf16 h) { return h != 0 && !(h & 0x8000u); }

__device__ __forceinline__ int wrapx(int x, int W, int px, bool &ok)
{
    if (x >= 0 && x < W) return x;
    if (px) return (x < 0) ? x + W : x - W;
    ok = false;
    return 0;
}

// ---------------------------------------------------------------------------
// layer 1: 3 -> 32, FP32 FMA, ReLU, BF16 out
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_cnn_l1(const float *__restrict__ in3, const float *__restrict__ prm,
                                               bf16 *__restrict__ out, int H, int W, int px)
{
    __shared__ float w[896];
    for (int k = threadIdx.x; k < 896; k += 256) w[k] = prm[k];
    __syncthreads();
    const int b = blockIdx.y, p = blockIdx.x * 256 + threadIdx.x;
    if (p >= H * W) return;
    const int y = p / W, x = p - y * W;
    float v[27];
#pragma unroll
    for (int t = 0; t < 9; t++) {
        const int yy = y + t / 3 - 1;
        bool ok = (yy >= 0 && yy < H);
        const int xx = wrapx(x + t % 3 - 1, W, px, ok);
#pragma unroll
        for (int c = 0; c < 3; c++) v[3 * t + c] = ok ? in3[(((size_t)b * H + yy) * W + xx) * 3 + c] : 0.0f;
    }
    bf16 o[CNN_C];
#pragma unroll
    for (int co = 0; co < CNN_C; co++) {
        float a = w[864 + co];
#pragma unroll
        for (int k = 0; k < 27; k++) a += w[co * 27 + k] * v[k];
        o[co] = f2bf(fmaxf(a, 0.0f));
    }
    uint4 *dst = (uint4 *)(out + (((size_t)b * H + y) * W + x) * CNN_C);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        uint4 u;
        u.x = o[8 * q + 0] | ((unsigned)o[8 * q + 1] << 16);
        u.y = o[8 * q + 2] | ((unsigned)o[8 * q + 3] << 16);
        u.z = o[8 * q + 4] | ((unsigned)o[8 * q + 5] << 16);
        u.w = o[8 * q + 6] | ((unsigned)o[8 * q + 7] << 16);
        dst[q] = u;
    }
}

// ---------------------------------------------------------------------------
// layers 2..5 (and the input gradient of the backward pass): 32 -> 32 on MFMA.
// wts: BF16 [tap 9][co 32][ci 32].  bias may be null.  With relu the output
// is max(., 0); with mask the output is zeroed where mask <= 0 (ReLU'(a)).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load_halo_tile(const bf16 *__restrict__ in, bf16 *lds, int b, int ty0, int tx0, int H,
                                               int W, int px)
{
    for (int idx = threadIdx.x; idx < HALO * HALO * 4; idx += 256) {
        const int pix = idx >> 2, ch = idx & 3;
        const int ty = pix / HALO, tx = pix - ty * HALO;
        const int gy = ty0 + ty - 1;
        bool ok = (gy >= 0 && gy < H);
        const int gx = wrapx(tx0 + tx - 1, W, px, ok);
        uint4 u = make_uint4(0, 0, 0, 0);
        if (ok) u = *(const uint4 *)(in + (((size_t)b * H + gy) * W + gx) * CNN_C + ch * 8);
        *(uint4 *)(lds + pix * PSTR + ch * 8) = u;
    }
}

template <bool RELU, bool MASK>
__global__ void __launch_bounds__(256) k_cnn_mid(const bf16 *__restrict__ in, const bf16 *__restrict__ wts,
                                                const float *__restrict__ bias, bf16 *__restrict__ out,
                                                const bf16 *__restrict__ mask, int H, int W, int px)
{
    __shared__ __attribute__((aligned(16))) bf16 sin[HALO * HALO * PSTR];
    __shared__ __attribute__((aligned(16))) bf16 sw[9 * CNN_C * PSTR];
    const int ntx = W / TILE, b = blockIdx.y;
    const int ty0 = (blockIdx.x / ntx) * TILE, tx0 = (blockIdx.x % ntx) * TILE;
    load_halo_tile(in, sin, b, ty0, tx0, H, W, px);
    for (int idx = threadIdx.x; idx < 9 * CNN_C * 4; idx += 256) {
        const int row = idx >> 2, ch = idx & 3;
        *(uint4 *)(sw + row * PSTR + ch * 8) = *(const uint4 *)(wts + row * CNN_C + ch * 8);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, col = lane & 15, grp = lane >> 4;
    v4f acc[4][2];
#pragma unroll
    for (int r = 0; r < 4; r++) { acc[r][0] = (v4f){0, 0, 0, 0}; acc[r][1] = (v4f){0, 0, 0, 0}; }
#pragma unroll
    for (int tap = 0; tap < 9; tap++) {
        const int ky = tap / 3, kx = tap % 3;
#pragma unroll
        for (int kc = 0; kc < 2; kc++) {
            const int ci = kc * 16 + 4 * grp;
            const v4s b0 = *(const v4s *)(sw + (tap * CNN_C + col) * PSTR + ci);
            const v4s b1 = *(const v4s *)(sw + (tap * CNN_C + 16 + col) * PSTR + ci);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = wv * 4 + r;
                const v4s a = *(const v4s *)(sin + ((row + ky) * HALO + col + kx) * PSTR + ci);
                acc[r][0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b0, acc[r][0], 0, 0, 0);
                acc[r][1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b1, acc[r][1], 0, 0, 0);
            }
        }
    }
    __syncthreads();
    // stage the 16 x 16 x 32 output tile in LDS (reusing sin), then store coalesced
    bf16 *so = sin;
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int pcol = 4 * grp + q, co = h * 16 + col;
                float v = acc[r][h][q] + (bias ? bias[co] : 0.0f);
                if (RELU) v = fmaxf(v, 0.0f);
                so[((wv * 4 + r) * TILE + pcol) * PSTR + co] = f2bf(v);
            }
    __syncthreads();
    for (int idx = threadIdx.x; idx < TILE * TILE * 4; idx += 256) {
        const int pix = idx >> 2, ch = idx & 3;
        const int gy = ty0 + pix / TILE, gx = tx0 + pix % TILE;
        const size_t g = (((size_t)b * H + gy) * W + gx) * CNN_C + ch * 8;
        uint4 u = *(const uint4 *)(so + pix * PSTR + ch * 8);
        if (MASK) {
            const uint4 m = *(const uint4 *)(mask + g);
            unsigned *pu = &u.x;
            const unsigned *pm = &m.x;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                unsigned lo = bfpos((bf16)(pm[e] & 0xffffu)) ? (pu[e] & 0xffffu) : 0u;
                unsigned hi = bfpos((bf16)(pm[e] >> 16)) ? (pu[e] & 0xffff0000u) : 0u;
                pu[e] = lo | hi;
            }
        }
        *(uint4 *)(out + g) = u;
    }
}

// ---------------------------------------------------------------------------
// layer 6: 32 -> 1, FP32 FMA over BF16 activations; out = bias + sum
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_cnn_l6(const bf16 *__restrict__ a5, const float *__restrict__ prm,
                                               float *__restrict__ out, int H, int W, int px)
{
    __shared__ float w[289];
    for (int k = threadIdx.x; k < 289; k += 256) w[k] = prm[37888 + k];
    __syncthreads();
    const int b = blockIdx.y, p = blockIdx.x * 256 + threadIdx.x;
    if (p >= H * W) return;
    const int y = p / W, x = p - y * W;
    float a = w[288];
    for (int t = 0; t < 9; t++) {
        const int yy = y + t / 3 - 1;
        bool ok = (yy >= 0 && yy < H);
        const int xx = wrapx(x + t % 3 - 1, W, px, ok);
        if (!ok) continue;
        const uint4 *src = (const uint4 *)(a5 + (((size_t)b * H + yy) * W + xx) * CNN_C);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint4 u = src[q];
            const unsigned e[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                a += w[t * 32 + 8 * q + 2 * k] * bf2f((bf16)(e[k] & 0xffffu));
                a += w[t * 32 + 8 * q + 2 * k + 1] * bf2f((bf16)(e[k] >> 16));
            }
        }
    }
    out[(size_t)b * H * W + p] = a;
}

// ---------------------------------------------------------------------------
// backward
// ---------------------------------------------------------------------------
// Loss = sum over valid pixels of (out - t)^2 / nvalid; valid = at least
// `border` pixels from every image edge.  gout = dLoss/dout; per-block loss.
__global__ void __launch_bounds__(256) k_cnn_loss(const float *__restrict__ out, const float *__restrict__ tgt,
                                                 float *__restrict__ gout, float *__restrict__ part, int H, int W,
                                                 int border, float inv_n)
{
    __shared__ float sh[256];
    const int b = blockIdx.y, p = blockIdx.x * 256 + threadIdx.x;
    float l = 0.0f;
    if (p < H * W) {
        const int y = p / W, x = p - y * W;
        const size_t k = (size_t)b * H * W + p;
        const bool ok = y >= border && y < H - border && x >= border && x < W - border;
        const float d = ok ? out[k] - tgt[k] : 0.0f;
        gout[k] = 2.0f * d * inv_n;
        l = d * d * inv_n;
    }
    sh[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = sh[0];
}

// Layer 6 backward on a 16 x 16 tile: gz5 = (sum_tap W6[tap] gout(p - off)) * (a5 > 0)
// and per-block partials of dW6[tap][ci] = sum_p gout(p) a5(p + off)[ci], db6.
__global__ void __launch_bounds__(256) k_cnn_l6_back(const float *__restrict__ gout, const bf16 *__restrict__ a5,
                                                    const float *__restrict__ prm, bf16 *__restrict__ gz5,
                                                    float *__restrict__ part, int H, int W, int px)
{
    __shared__ float sg[HALO * HALO];                  // gout with halo
    __shared__ __attribute__((aligned(16))) bf16 sa[HALO * HALO * PSTR];
    __shared__ float w[289];
    const int ntx = W / TILE, b = blockIdx.y;
    const int ty0 = (blockIdx.x / ntx) * TILE, tx0 = (blockIdx.x % ntx) * TILE;
    for (int k = threadIdx.x; k < 289; k += 256) w[k] = prm[37888 + k];
    for (int idx = threadIdx.x; idx < HALO * HALO; idx += 256) {
        const int ty = idx / HALO, tx = idx - ty * HALO, gy = ty0 + ty - 1;
        bool ok = (gy >= 0 && gy < H);
        const int gx = wrapx(tx0 + tx - 1, W, px, ok);
        sg[idx] = ok ? gout[((size_t)b * H + gy) * W + gx] : 0.0f;
    }
    load_halo_tile(a5, sa, b, ty0, tx0, H, W, px);
    __syncthreads();
    {   // input gradient: thread = pixel
        const int ty = threadIdx.x / TILE, tx = threadIdx.x % TILE;
        const size_t g = (((size_t)b * H + ty0 + ty) * W + tx0 + tx) * CNN_C;
        for (int c = 0; c < CNN_C; c++) {
            float s = 0.0f;
#pragma unroll
            for (int t = 0; t < 9; t++) s += w[t * 32 + c] * sg[(ty + 2 - t / 3) * HALO + tx + 2 - t % 3];
            const bf16 a = sa[((ty + 1) * HALO + tx + 1) * PSTR + c];
            gz5[g + c] = bfpos(a) ? f2bf(s) : (bf16)0;
        }
    }
    // weight gradient partials: thread k < 288 -> (tap, ci); k == 288 -> bias
    float *pp = part + (size_t)(blockIdx.y * gridDim.x + blockIdx.x) * 289;
    for (int k = threadIdx.x; k < 289; k += 256) {
        float s = 0.0f;
        if (k < 288) {
            const int t = k / 32, c = k % 32, ky = t / 3, kx = t % 3;
            for (int p = 0; p < TILE * TILE; p++) {
                const int ty = p / TILE, tx = p % TILE;
                s += sg[(ty + 1) * HALO + tx + 1] * bf2f(sa[((ty + ky) * HALO + tx + kx) * PSTR + c]);
            }
        } else {
            for (int p = 0; p < TILE * TILE; p++) s += sg[(p / TILE + 1) * HALO + p % TILE + 1];
        }
        pp[k] = s;
    }
}

// Weight gradient of a 32 -> 32 layer on MFMA: per tile and tap,
// dW[tap][co][ci] += sum_p gz(p)[co] a(p + off(tap))[ci]  (M = co, N = ci,
// K = the tile's 256 pixels), plus db[co]; per-block partials of 9248.
__global__ void __launch_bounds__(256) k_cnn_dw(const bf16 *__restrict__ gz, const bf16 *__restrict__ a,
                                               float *__restrict__ part, int H, int W, int px)
{
    __shared__ __attribute__((aligned(16))) bf16 sa[HALO * HALO * PSTR];
    __shared__ __attribute__((aligned(16))) bf16 sg[TILE * TILE * PSTR];
    const int ntx = W / TILE, b = blockIdx.y;
    const int ty0 = (blockIdx.x / ntx) * TILE, tx0 = (blockIdx.x % ntx) * TILE;
    load_halo_tile(a, sa, b, ty0, tx0, H, W, px);
    for (int idx = threadIdx.x; idx < TILE * TILE * 4; idx += 256) {
        const int pix = idx >> 2, ch = idx & 3;
        const int gy = ty0 + pix / TILE, gx = tx0 + pix % TILE;
        *(uint4 *)(sg + pix * PSTR + ch * 8) = *(const uint4 *)(gz + (((size_t)b * H + gy) * W + gx) * CNN_C + ch * 8);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, col = lane & 15, grp = lane >> 4;
    const int mb = wv >> 1, nb = wv & 1;              // co block, ci block of this wave
    float *pp = part + (size_t)(blockIdx.y * gridDim.x + blockIdx.x) * 9248;
    for (int tap = 0; tap < 9; tap++) {
        const int ky = tap / 3, kx = tap % 3;
        v4f acc = (v4f){0, 0, 0, 0};
        for (int ks = 0; ks < TILE; ks++) {           // tile row ks = 16 pixels of K
            v4s av, bv;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int pc = 4 * grp + q;           // pixel column in the row
                av[q] = (short)sg[(ks * TILE + pc) * PSTR + mb * 16 + col];
                bv[q] = (short)sa[((ks + ky) * HALO + pc + kx) * PSTR + nb * 16 + col];
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int co = mb * 16 + 4 * grp + q, ci = nb * 16 + col;
            pp[(tap * CNN_C + co) * CNN_C + ci] = acc[q];
        }
    }
    if (threadIdx.x < CNN_C) {
        float s = 0.0f;
        for (int p = 0; p < TILE * TILE; p++) s += bf2f(sg[p * PSTR + threadIdx.x]);
        pp[9216 + threadIdx.x] = s;
    }
}

// Layer 1 weight gradient: per tile, dW1[co][tap][ci] = sum_p gz1(p)[co] in3(p + off)[ci], db1.
__global__ void __launch_bounds__(256) k_cnn_l1_back(const bf16 *__restrict__ gz, const float *__restrict__ in3,
                                                    float *__restrict__ part, int H, int W, int px)
{
    __shared__ float si[HALO * HALO * 3];
    __shared__ __attribute__((aligned(16))) bf16 sg[TILE * TILE * PSTR];
    const int ntx = W / TILE, b = blockIdx.y;
    const int ty0 = (blockIdx.x / ntx) * TILE, tx0 = (blockIdx.x % ntx) * TILE;
    for (int idx = threadIdx.x; idx < HALO * HALO * 3; idx += 256) {
        const int pix = idx / 3, c = idx % 3, ty = pix / HALO, tx = pix % HALO, gy = ty0 + ty - 1;
        bool ok = (gy >= 0 && gy < H);
        const int gx = wrapx(tx0 + tx - 1, W, px, ok);
        si[idx] = ok ? in3[(((size_t)b * H + gy) * W + gx) * 3 + c] : 0.0f;
    }
    for (int idx = threadIdx.x; idx < TILE * TILE * 4; idx += 256) {
        const int pix = idx >> 2, ch = idx & 3;
        const int gy = ty0 + pix / TILE, gx = tx0 + pix % TILE;
        *(uint4 *)(sg + pix * PSTR + ch * 8) = *(const uint4 *)(gz + (((size_t)b * H + gy) * W + gx) * CNN_C + ch * 8);
    }
    __syncthreads();
    float *pp = part + (size_t)(blockIdx.y * gridDim.x + blockIdx.x) * 896;
    for (int k = threadIdx.x; k < 896; k += 256) {
        float s = 0.0f;
        if (k < 864) {
            const int co = k / 27, r = k % 27, t = r / 3, c = r % 3, ky = t / 3, kx = t % 3;
            for (int p = 0; p < TILE * TILE; p++) {
                const int ty = p / TILE, tx = p % TILE;
                s += bf2f(sg[p * PSTR + co]) * si[((ty + ky) * HALO + tx + kx) * 3 + c];
            }
        } else {
            for (int p = 0; p < TILE * TILE; p++) s += bf2f(sg[p * PSTR + k - 864]);
        }
        pp[k] = s;
    }
}

// g[k] = sum over blocks q of part[q * n + k], k < n, in a fixed order over
// two launches: row group y (blockIdx.y) sums rows q = y, y + RSPLIT, ...
// into rpart[y * n + k]; then the RSPLIT group sums are added in order.
__global__ void k_cnn_reduce_rows(const float *__restrict__ part, int nblk, int n, float *__restrict__ rpart)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (k >= n) return;
    float s = 0.0f;
    for (int q = y; q < nblk; q += PVK_CNN_RSPLIT) s += part[(size_t)q * n + k];
    rpart[(size_t)y * n + k] = s;
}

__global__ void k_cnn_reduce_groups(const float *__restrict__ rpart, int n, float *__restrict__ g)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    float s = 0.0f;
#pragma unroll 8
    for (int y = 0; y < PVK_CNN_RSPLIT; y++) s += rpart[(size_t)y * n + k];
    g[k] = s;
}

static void cnn_reduce(const pvk_cnn *c, int nblk, int n, float *g, hipStream_t s)
{
    hipLaunchKernelGGL(k_cnn_reduce_rows, dim3((n + 255) / 256, PVK_CNN_RSPLIT), dim3(256), 0, s,
                       (const float *)c->part, nblk, n, c->rpart);
    hipLaunchKernelGGL(k_cnn_reduce_groups, dim3((n + 255) / 256), dim3(256), 0, s, (const float *)c->rpart, n, g);
}

// Adam (FP32 master weights)
__global__ void k_cnn_adam(float *w, const float *g, float *m, float *v, int n, float lr, float b1, float b2,
                           float eps, float c1, float c2)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.0f - b1) * gk, vk = b2 * v[k] + (1.0f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    w[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

// BF16 copies of the 32 -> 32 layers: forward [tap][co][ci] and, for the
// input gradient, the flipped transpose wt[tap'][ci][co] = W[8 - tap'][co][ci].
__global__ void k_cnn_pack(const float *__restrict__ prm, bf16 *__restrict__ wb, bf16 *__restrict__ wt)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 4 * 9216) return;
    const int l = k / 9216, r = k % 9216, tap = r / 1024, co = (r / 32) % 32, ci = r % 32;
    const float v = prm[896 + l * 9248 + r];
    wb[k] = f2bf(v);
    wt[l * 9216 + ((8 - tap) * CNN_C + ci) * CNN_C + co] = f2bf(v);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
extern "C" void pvk_cnn_pack(const pvk_cnn *c, hipStream_t s)
{
    hipLaunchKernelGGL(k_cnn_pack, dim3((4 * 9216 + 255) / 256), dim3(256), 0, s, c->w, c->wb, c->wt);
}

// forward of a batch: in3 [B][H][W][3] -> out [B][H][W]; activations kept in c->act
extern "C" void pvk_cnn_forward(const pvk_cnn *c, const float *in3, float *out, int B, int H, int W, int px,
                                hipStream_t s)
{
    const dim3 gp((H * W + 255) / 256, B), gt((H / TILE) * (W / TILE), B);
    hipLaunchKernelGGL(k_cnn_l1, gp, dim3(256), 0, s, in3, (const float *)c->w, c->act[0], H, W, px);
    for (int l = 0; l < 4; l++)
        hipLaunchKernelGGL((k_cnn_mid<true, false>), gt, dim3(256), 0, s, (const bf16 *)c->act[l],
                           (const bf16 *)(c->wb + l * 9216), (const float *)(c->w + 896 + l * 9248 + 9216),
                           c->act[l + 1], (const bf16 *)0, H, W, px);
    hipLaunchKernelGGL(k_cnn_l6, gp, dim3(256), 0, s, (const bf16 *)c->act[4], (const float *)c->w, out, H, W, px);
}

// loss and full gradient (into c->g) of the last forward; returns nothing on
// the host: the per-block loss partials are left in c->lpart
extern "C" void pvk_cnn_backward(const pvk_cnn *c, const float *in3, const float *out, const float *tgt, int B,
                                 int H, int W, int px, int border, hipStream_t s)
{
    const dim3 gp((H * W + 255) / 256, B), gt((H / TILE) * (W / TILE), B);
    const int nt = (H / TILE) * (W / TILE) * B;
    const int interior = (H - 2 * border) * (W - 2 * border) * B;
    hipLaunchKernelGGL(k_cnn_loss, gp, dim3(256), 0, s, out, tgt, c->gout, c->lpart, H, W, border,
                       1.0f / (float)(interior > 0 ? interior : 1));
    hipLaunchKernelGGL(k_cnn_l6_back, gt, dim3(256), 0, s, (const float *)c->gout, (const bf16 *)c->act[4],
                       (const float *)c->w, c->gz[0], c->part, H, W, px);
    cnn_reduce(c, nt, 289, c->g + 37888, s);
    int cur = 0;
    for (int l = 3; l >= 0; l--) {       // layer l + 2: input act[l], output-side gradient gz[cur]
        hipLaunchKernelGGL(k_cnn_dw, gt, dim3(256), 0, s, (const bf16 *)c->gz[cur], (const bf16 *)c->act[l], c->part,
                           H, W, px);
        cnn_reduce(c, nt, 9248, c->g + 896 + l * 9248, s);
        hipLaunchKernelGGL((k_cnn_mid<false, true>), gt, dim3(256), 0, s, (const bf16 *)c->gz[cur],
                           (const bf16 *)(c->wt + l * 9216), (const float *)0, c->gz[cur ^ 1],
                           (const bf16 *)c->act[l], H, W, px);
        cur ^= 1;
    }
    hipLaunchKernelGGL(k_cnn_l1_back, gt, dim3(256), 0, s, (const bf16 *)c->gz[cur], in3, c->part, H, W, px);
    cnn_reduce(c, nt, 896, c->g, s);
}

extern "C" void pvk_cnn_adam(const pvk_cnn *c, float lr, int t, hipStream_t s)
{
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
    hipLaunchKernelGGL(k_cnn_adam, dim3((PVK_CNN_NPARAM + 255) / 256), dim3(256), 0, s, c->w, (const float *)c->g,
                       c->m, c->v, PVK_CNN_NPARAM, lr, b1, b2, 1e-8f, c1, c2);
    pvk_cnn_pack(c, s);
}
