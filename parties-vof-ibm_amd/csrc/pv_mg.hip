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
//   smoother      red-black Gauss-Seidel, nu = 1 or 2 sweeps: (red, black)
//                 before the coarse correction, (black, red) after it;
//   restriction   sum of the four fine residuals;
//   prolongation  piecewise constant;
//   coarse faces  half the sum of the two fine faces across the coarse face;
//   the last level (<= PVK_MG_BOTTOM cells) is solved exactly with the
//   inverse of its operator grounded at cell 0, both sides projected to zero
//   mean (built per solve by a band Cholesky, one column per workgroup).
// Each level above PVK_MG_LDS_CELLS cells is two launches (pre-smoothing +
// restriction; correction + post-smoothing) on LDS tiles with a halo, so f,
// the faces and x cross HBM once per launch; the levels below run in one
// workgroup from LDS.  The V-cycle stores and computes in T = double, or in
// T = float inside the FP64 CG (the mixed form: level 0 then reads r and
// writes z in FP64, its own x and faces are FP32).
#include <hip/hip_runtime.h>
#include <math.h>
#include <mutex>
#include "pv_kernels.h"
#include "pv_block.h"

#define AT(a, i, j) (a)[(ptrdiff_t)(j) * P + (i)]
#define TXP 64      // cells per block row of the per-cell kernels
#define TYR 4       // rows per block; TXP * TYR = NTH

// a level with arrays of the V-cycle's type
template <typename T> struct Lev {
    int nx, ny, pitch;
    T *x, *f, *bx, *by, *xp;
};
template <typename T> __host__ __device__ static Lev<T> lev(const pvk_mglev &l)
{
    Lev<T> v;
    v.nx = l.nx;
    v.ny = l.ny;
    v.pitch = l.pitch;
    v.x = (T *)l.x;
    v.f = (T *)l.f;
    v.bx = (T *)l.bx;
    v.by = (T *)l.by;
    v.xp = (T *)l.xp;
    return v;
}

template <typename T> __device__ __forceinline__ T ld(const T *a, int pitch, int i, int j)
{
    return a[(ptrdiff_t)j * pitch + i];
}

// f of level L at (i, j); level 0 of the FP32 V-cycle reads the FP64 r0
template <typename T> __device__ __forceinline__ T fget(const Lev<T> &L, const double *r0, int p0, int i, int j)
{
    return r0 ? (T)r0[(ptrdiff_t)j * p0 + i] : ld(L.f, L.pitch, i, j);
}

// the CG step of iteration n folded into the level-0 pre-smoothing:
// r_{n+1} = r_n - alpha (A d_n - mean(A d_n)), phi += alpha d_n
struct MgStep {
    const double *Ad, *d;
    double *phi, *rnew;
    double alpha, mad;
    bool on;
};

// stop test of CG iteration n on the <r, r> partials (as k_transpose):
// n = -1 sets the reference norm, n = -2 keeps it (warm restart).  The
// partials of r_{n+1} come from the level-0 pre-smoothing (rrt, per tile)
// when it forms r_{n+1}, else from the start or update kernels (rr, per row).
__device__ void mg_stop_test(pvk_scal *sc, const pvk_pcg &w, int n, double tol, int maxit, double *sh)
{
    const bool tiles = (n >= 0 && w.nrrt > 0);
    const double rr = tiles ? block_sum(w.rrt, w.nrrt, sh) : block_sum(w.rr, w.nrr, sh);
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

// ---------------------------------------------------------------------------
// Fused smoothing of one level (levels with more than PVK_MG_LDS_CELLS cells).
// A block owns a core tile of TX x TY cells (64 x 16 on levels wider than
// 512, 32 x 16 up to 512, 32 x 8 up to 256: smaller tiles, more blocks, on
// the levels whose time is latency) and works on it extended by a halo of 2 nu
// cells (periodic in x, cut at the walls in y): half-sweep k updates the
// cells of its colour at distance >= k from the extended edge, so after
// 2 nu half-sweeps the core (and the ring the residual needs) holds exactly
// what sweeping the whole level would give.  f, the face coefficients and
// 1 / (their sum) stay in registers, x in LDS.
// ---------------------------------------------------------------------------
template <int NU, int SH> struct MgTile {
    static constexpr int TX = (SH == 0) ? 64 : 32;    // core tile
    static constexpr int TY = (SH == 2) ? 8 : 16;
    static constexpr int H = 2 * NU;                  // halo
    static constexpr int W = TX + 2 * H;              // extended width
    static constexpr int HT = TY + 2 * H;          // extended height
    static constexpr int NC = W * HT;
    static constexpr int K = (NC + NTH - 1) / NTH;    // cells per thread
    static constexpr int KC = (TX * TY + NTH - 1) / NTH; // core cells per thread
};

// is extended-tile cell (lx, ly) in the core, and in the level (x0 + lx the
// unwrapped column)?
template <int NU, int SH> __device__ __forceinline__ bool tile_core(int lx, int ly, int x0, int gj, int nx, int ny)
{
    using TT = MgTile<NU, SH>;
    return lx >= TT::H && lx < TT::H + TT::TX && ly >= TT::H && ly < TT::H + TT::TY && x0 + lx < nx && gj < ny;
}

// Column and row in the level of extended-tile cell c (clamped into the
// tile and, for rows, into the level: a cell outside it is masked later).
template <int NU, int SH>
__device__ __forceinline__ void tile_cell(int c, int x0, int y0, int nx, int ny, int &gi, int &gj)
{
    using TT = MgTile<NU, SH>;
    c = min(c, TT::NC - 1);
    const int ly = c / TT::W, lx = c - ly * TT::W;
    gj = min(max(y0 + ly, 0), ny - 1);
    gi = (x0 + lx) % nx;
    if (gi < 0) gi += nx;
}

// is extended-tile cell c a cell of the level?
template <int NU, int SH> __device__ __forceinline__ bool tile_in(int c, int y0, int ny)
{
    using TT = MgTile<NU, SH>;
    const int gj = y0 + c / TT::W;
    return c < TT::NC && gj >= 0 && gj < ny;
}

// The raw faces of this thread's cells of the extended tile.  Every load is
// unconditional, from indices clamped into the level (the face arrays hold
// every row, wall rows included), so the loads issue back to back with no
// wait between them; tile_mask zeroes what lies outside the level or on a
// wall once they have arrived.
template <int NU, int SH, typename T>
__device__ __forceinline__ void tile_faces(const Lev<T> &L, int x0, int y0, T *ce, T *cw, T *cn, T *cs)
{
    using TT = MgTile<NU, SH>;
    const int p = L.pitch;
#pragma unroll
    for (int u = 0; u < TT::K; u++) {
        int gi, gj;
        tile_cell<NU, SH>(threadIdx.x + u * NTH, x0, y0, L.nx, L.ny, gi, gj);
        const int giw = (gi == 0) ? L.nx - 1 : gi - 1;
        ce[u] = ld(L.bx, p, gi, gj);
        cw[u] = ld(L.bx, p, giw, gj);
        cn[u] = ld(L.by, p, gi, gj);
        cs[u] = ld(L.by, p, gi, max(gj - 1, 0));
    }
}

// The level's right-hand side of this thread's cells of the extended tile
// (raw, unconditional loads as tile_faces).
template <int NU, int SH, typename T>
__device__ __forceinline__ void tile_rhs(const Lev<T> &L, const double *r0, int p0, int x0, int y0, T *f)
{
    using TT = MgTile<NU, SH>;
    if (r0) {
#pragma unroll
        for (int u = 0; u < TT::K; u++) {
            int gi, gj;
            tile_cell<NU, SH>(threadIdx.x + u * NTH, x0, y0, L.nx, L.ny, gi, gj);
            f[u] = (T)r0[(ptrdiff_t)gj * p0 + gi];
        }
    } else {
#pragma unroll
        for (int u = 0; u < TT::K; u++) {
            int gi, gj;
            tile_cell<NU, SH>(threadIdx.x + u * NTH, x0, y0, L.nx, L.ny, gi, gj);
            f[u] = ld(L.f, L.pitch, gi, gj);
        }
    }
}

// Zero f and the faces outside the level and the wall faces, and form
// 1 / (sum of the faces).
template <int NU, int SH, typename T>
__device__ __forceinline__ void tile_mask(int y0, int ny, T *f, T *ce, T *cw, T *cn, T *cs, T *id)
{
    using TT = MgTile<NU, SH>;
#pragma unroll
    for (int u = 0; u < TT::K; u++) {
        const int c = threadIdx.x + u * NTH, gj = y0 + c / TT::W;
        const bool in = tile_in<NU, SH>(c, y0, ny);
        f[u] = in ? f[u] : T(0);
        ce[u] = in ? ce[u] : T(0);
        cw[u] = in ? cw[u] : T(0);
        cn[u] = (in && gj + 1 < ny) ? cn[u] : T(0);
        cs[u] = (in && gj > 0) ? cs[u] : T(0);
        id[u] = in ? T(1) / (ce[u] + cw[u] + cn[u] + cs[u]) : T(0);
    }
}

// The CG step's loads of the level-0 pre-smoothing: r_n and A d_n of this
// thread's cells of the extended tile (ra, aa; before the step's scalars
// are summed, in flight meanwhile), then d_n and phi of its cells of the core
// (core cell q = threadIdx.x + k NTH, row-major in the core: dd, pp; in
// flight while f is formed).  Every load is unconditional, from an index
// clamped into the level (cells outside it are masked when the values are
// used), so the loads issue back to back with no wait between them.
template <int NU, int SH, typename T>
__device__ __forceinline__ void tile_step_loads(const Lev<T> &L, const double *r0, int p0, const MgStep &st,
                                                int x0, int y0, double *ra, double *aa)
{
    using TT = MgTile<NU, SH>;
#pragma unroll
    for (int u = 0; u < TT::K; u++) {
        int gi, gj;
        tile_cell<NU, SH>(threadIdx.x + u * NTH, x0, y0, L.nx, L.ny, gi, gj);
        const size_t idx = (size_t)gj * p0 + gi;
        ra[u] = r0[idx];
        aa[u] = st.Ad[idx];
    }
}

template <int NU, int SH, typename T>
__device__ __forceinline__ void tile_step_core_loads(const Lev<T> &L, int p0, const MgStep &st, int bx, int by,
                                                     double *dd, double *pp)
{
    using TT = MgTile<NU, SH>;
#pragma unroll
    for (int k = 0; k < TT::KC; k++) {
        const int q = min(threadIdx.x + k * NTH, TT::TX * TT::TY - 1), ly = q / TT::TX, lx = q - ly * TT::TX;
        const int gi = min(bx * TT::TX + lx, L.nx - 1), gj = min(by * TT::TY + ly, L.ny - 1);
        const size_t idx = (size_t)gj * p0 + gi;
        dd[k] = st.d[idx];
        pp[k] = st.phi[idx];
    }
}

// f = r_{n+1} = r_n - alpha (A d_n - mean) of this thread's cells of the
// extended tile from the step's loads (tile_mask zeroes it outside the
// level); the core cells' r_{n+1} also go to rl (FP64, row-major in the
// core) for the core pass.
template <int NU, int SH, typename T>
__device__ __forceinline__ void tile_rhs_step(const Lev<T> &L, const MgStep &st, int x0, int y0, const double *ra,
                                              const double *aa, T *f, double *rl)
{
    using TT = MgTile<NU, SH>;
#pragma unroll
    for (int u = 0; u < TT::K; u++) {
        const int c = threadIdx.x + u * NTH, ly = c / TT::W, lx = c - ly * TT::W, gj = y0 + ly;
        const double rn = ra[u] - st.alpha * (aa[u] - st.mad);
        f[u] = (T)rn;
        if (tile_in<NU, SH>(c, y0, L.ny) && tile_core<NU, SH>(lx, ly, x0, gj, L.nx, L.ny))
            rl[(ly - TT::H) * TT::TX + lx - TT::H] = rn;
    }
}

// the core pass of the CG step: r_{n+1} (from rl) to rnew, phi += alpha d_n,
// and this thread's share of <r_{n+1}, r_{n+1}>
template <int NU, int SH, typename T>
__device__ __forceinline__ double tile_step_core(const Lev<T> &L, int p0, const MgStep &st, int bx, int by,
                                                 const double *dd, const double *pp, const double *rl)
{
    using TT = MgTile<NU, SH>;
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < TT::KC; k++) {
        const int q = threadIdx.x + k * NTH, ly = q / TT::TX, lx = q - ly * TT::TX;
        const int gi = bx * TT::TX + lx, gj = by * TT::TY + ly;
        if (q < TT::TX * TT::TY && gi < L.nx && gj < L.ny) {
            const size_t idx = (size_t)gj * p0 + gi;
            const double rn = rl[q];
            st.rnew[idx] = rn;
            st.phi[idx] = pp[k] + st.alpha * dd[k];
            acc += rn * rn;
        }
    }
    return acc;
}

// f, faces and 1 / (their sum) of this thread's cells of the extended tile:
// every global load of the smoothing is in this one batch.
template <int NU, int SH, typename T>
__device__ __forceinline__ void tile_coeffs(const Lev<T> &L, const double *r0, int p0, int x0, int y0, T *f,
                                            T *ce, T *cw, T *cn, T *cs, T *id)
{
    tile_faces<NU, SH>(L, x0, y0, ce, cw, cn, cs);
    tile_rhs<NU, SH>(L, r0, p0, x0, y0, f);
    tile_mask<NU, SH>(y0, L.ny, f, ce, cw, cn, cs, id);
}

// tile of this block, XCD-aware (consecutive tiles of a run share an L2)
__device__ __forceinline__ void tile_of(int &bx, int &by)
{
    const int t = xcd_band(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    by = t / gridDim.x;
    bx = t - by * gridDim.x;
}

// half-sweep k of colour col on the cells at distance >= k from the edge
template <int NU, int SH, typename T>
__device__ __forceinline__ void tile_sweep(T *xs, int x0, int y0, int ny, int k, int col, const T *f, const T *ce,
                                           const T *cw, const T *cn, const T *cs, const T *id)
{
    using TT = MgTile<NU, SH>;
    T v[TT::K];
    bool act[TT::K];
#pragma unroll
    for (int u = 0; u < TT::K; u++) {
        const int c = threadIdx.x + u * NTH, ly = c / TT::W, lx = c - ly * TT::W, gj = y0 + ly;
        act[u] = c < TT::NC && gj >= 0 && gj < ny && (((x0 + lx + gj) & 1) == col) && lx >= k && lx < TT::W - k &&
                 ly >= k && ly < TT::HT - k;
        v[u] = T(0);
        if (act[u]) {
            const T xe = (k > 0) ? xs[c + 1] : T(0), xw = (k > 0) ? xs[c - 1] : T(0);
            const T xn = (k > 0) ? xs[c + TT::W] : T(0), xq = (k > 0) ? xs[c - TT::W] : T(0);
            v[u] = (ce[u] * xe + cw[u] * xw + cn[u] * xn + cs[u] * xq - f[u]) * id[u];
        }
    }
    // a colour's cells read only the other colour: no barrier before writing
#pragma unroll
    for (int u = 0; u < TT::K; u++)
        if (act[u]) xs[threadIdx.x + u * NTH] = v[u];
    __syncthreads();
}

// pre-smoothing from x = 0 and restriction of the residual; the stop test of
// iteration n in block (0, 0) (check: the kernel after the one that formed
// the residual).  On level 0 with step (CG iteration n >= 0) the kernel
// first does the CG step: f = r_{n+1} is formed on the fly from r_n = r0
// (tile and halo), and the core writes r_{n+1} to rnew, phi += alpha d and
// the partial <r_{n+1}, r_{n+1}> of the tile.
template <int NU, int SH, typename T>
__global__ void __launch_bounds__(NTH, sizeof(T) == 4 ? 4 : 3) k_mg_pre(Lev<T> L, Lev<T> C, const double *r0, double *rnew, int p0,
                                               pvk_scal *sc, pvk_pcg w, int n, double tol, int maxit, int check,
                                               int step, int ncell)
{
    using TT = MgTile<NU, SH>;
    __shared__ T xs[TT::NC];
    __shared__ double rl[TT::TX * TT::TY];      // r_{n+1} of the core; then rs
    __shared__ double sh[3 * NTH];
    T *rs = (T *)rl;                              // core residual, after the sweeps
    if (sc->p_done) return;
    if (check && n != -3 && blockIdx.x == 0 && blockIdx.y == 0) mg_stop_test(sc, w, n, tol, maxit, sh);
    int bx, by;
    tile_of(bx, by);
    const int x0 = bx * TT::TX - TT::H, y0 = by * TT::TY - TT::H;
    T f[TT::K], ce[TT::K], cw[TT::K], cn[TT::K], cs[TT::K], id[TT::K];
    MgStep st;
    st.on = step && n >= 0;
    st.Ad = w.Ad;
    st.d = w.d[n & 1];
    st.phi = w.phi;
    st.rnew = rnew;
    st.alpha = st.mad = 0.0;
    double acc = 0.0;
    if (st.on) {
        // the tile's loads go out first and are in flight while the step's
        // scalars are summed (the three sums' loads together); nothing
        // waits on memory before the sums
        double ra[TT::K], aa[TT::K], dd[TT::KC], pp[TT::KC];
        tile_faces<NU, SH>(L, x0, y0, ce, cw, cn, cs);
        tile_step_loads<NU, SH>(L, r0, p0, st, x0, y0, ra, aa);
        double rz, dad, sad;
        block_sum3(w.rz[n & 1], w.nrz, w.dad[n & 1], w.ndad, w.sad, w.ndad, sh, rz, dad, sad);
        st.alpha = rz / dad;
        st.mad = sad / (double)ncell;
        tile_step_core_loads<NU, SH>(L, p0, st, bx, by, dd, pp);
        tile_rhs_step<NU, SH>(L, st, x0, y0, ra, aa, f, rl);
        tile_mask<NU, SH>(y0, L.ny, f, ce, cw, cn, cs, id);
        for (int c = threadIdx.x; c < TT::NC; c += NTH) xs[c] = T(0);
        __syncthreads();
        // (rs, on rl's memory, is written after the sweeps' barriers)
        acc = tile_step_core<NU, SH>(L, p0, st, bx, by, dd, pp, rl);
    } else {
        tile_coeffs<NU, SH>(L, r0, p0, x0, y0, f, ce, cw, cn, cs, id);
        for (int c = threadIdx.x; c < TT::NC; c += NTH) xs[c] = T(0);
        __syncthreads();
    }
    for (int k = 0; k < 2 * NU; k++) tile_sweep<NU, SH>(xs, x0, y0, L.ny, k, k & 1, f, ce, cw, cn, cs, id);
    // residual f - A x of each core cell from the registers (a wall face is
    // 0 there) to rs, and the core of x for the post-smoothing (to xp: the
    // post kernel reads it with halos while it writes x)
#pragma unroll
    for (int u = 0; u < TT::K; u++) {
        const int c = threadIdx.x + u * NTH, ly = c / TT::W, lx = c - ly * TT::W, gj = y0 + ly;
        if (c < TT::NC && gj >= 0 && tile_core<NU, SH>(lx, ly, x0, gj, L.nx, L.ny)) {
            const T x0c = xs[c];
            rs[(ly - TT::H) * TT::TX + lx - TT::H] =
                f[u] - (ce[u] * (xs[c + 1] - x0c) - cw[u] * (x0c - xs[c - 1]) + cn[u] * (xs[c + TT::W] - x0c) -
                        cs[u] * (x0c - xs[c - TT::W]));
            L.xp[(ptrdiff_t)gj * L.pitch + x0 + lx] = x0c;
        }
    }
    __syncthreads();
    // restriction: one coarse cell per thread (TT::TX/2 x TT::TY/2 <= NTH)
    if ((int)threadIdx.x < (TT::TX / 2) * (TT::TY / 2)) {
        const int I = threadIdx.x % (TT::TX / 2), J = threadIdx.x / (TT::TX / 2);
        const int gI = bx * (TT::TX / 2) + I, gJ = by * (TT::TY / 2) + J;
        if (gI < C.nx && gJ < C.ny) {
            T s = T(0);
#pragma unroll
            for (int dj = 0; dj < 2; dj++)
#pragma unroll
                for (int di = 0; di < 2; di++) s += rs[(2 * J + dj) * TT::TX + 2 * I + di];
            C.f[(ptrdiff_t)gJ * C.pitch + gI] = s;
        }
    }
    if (st.on) block_store_partial(acc, w.rrt, by * gridDim.x + bx, sh);
}

// coarse correction xp + xc(parent) and post-smoothing (black first); the
// result goes to x, or to the FP64 z0 on level 0 of the FP32 V-cycle.  The
// input (xp) and the output are separate arrays: a block reads its halo from
// the neighbouring tiles' cores, which their blocks overwrite, and when the
// level has more tiles than the GPU holds at once those blocks may already
// have finished.
// dot: partial <f, x> of the core (FP64, with r0 on that level) into w.rz[slot]
template <int NU, int SH, typename T>
__global__ void __launch_bounds__(NTH, sizeof(T) == 4 ? 4 : 3) k_mg_post(Lev<T> L, Lev<T> C, const double *r0, double *z0, int p0,
                                                const pvk_scal *sc, pvk_pcg w, int dot, int slot)
{
    using TT = MgTile<NU, SH>;
    __shared__ T xs[TT::NC];
    __shared__ double sh[NTH];
    if (sc->p_done) return;
    int bx, by;
    tile_of(bx, by);
    const int x0 = bx * TT::TX - TT::H, y0 = by * TT::TY - TT::H;
    T f[TT::K], ce[TT::K], cw[TT::K], cn[TT::K], cs[TT::K], id[TT::K], xa[TT::K], xb[TT::K];
    double fd[TT::KC];
    // every global load of the kernel first, unconditional (clamped indices)
    tile_faces<NU, SH>(L, x0, y0, ce, cw, cn, cs);
    tile_rhs<NU, SH>(L, r0, p0, x0, y0, f);
#pragma unroll
    for (int u = 0; u < TT::K; u++) {           // x = xp + the parent's correction
        int gi, gj;
        tile_cell<NU, SH>(threadIdx.x + u * NTH, x0, y0, L.nx, L.ny, gi, gj);
        xa[u] = ld(L.xp, L.pitch, gi, gj);
        xb[u] = ld(C.x, C.pitch, gi >> 1, gj >> 1);
    }
#pragma unroll
    for (int k = 0; k < TT::KC; k++) {          // the core's f for the dot
        fd[k] = 0.0;
        if (dot) {
            const int q = min(threadIdx.x + k * NTH, TT::TX * TT::TY - 1), ly = q / TT::TX, lx = q - ly * TT::TX;
            const int gi = min(bx * TT::TX + lx, L.nx - 1), gj = min(by * TT::TY + ly, L.ny - 1);
            fd[k] = r0 ? r0[(ptrdiff_t)gj * p0 + gi] : (double)ld(L.f, L.pitch, gi, gj);
        }
    }
    tile_mask<NU, SH>(y0, L.ny, f, ce, cw, cn, cs, id);
#pragma unroll
    for (int u = 0; u < TT::K; u++) {
        const int c = threadIdx.x + u * NTH;
        if (c < TT::NC) xs[c] = tile_in<NU, SH>(c, y0, L.ny) ? xa[u] + xb[u] : T(0);
    }
    __syncthreads();
    for (int k = 1; k <= 2 * NU; k++) tile_sweep<NU, SH>(xs, x0, y0, L.ny, k, k & 1, f, ce, cw, cn, cs, id);
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < TT::KC; k++) {
        const int c = threadIdx.x + k * NTH, ly = c / TT::TX, lx = c - ly * TT::TX;
        const int gi = bx * TT::TX + lx, gj = by * TT::TY + ly;
        if (c < TT::TX * TT::TY && gi < L.nx && gj < L.ny) {
            const T v = xs[(ly + TT::H) * TT::W + lx + TT::H];
            if (z0) z0[(ptrdiff_t)gj * p0 + gi] = (double)v;
            else L.x[(ptrdiff_t)gj * L.pitch + gi] = v;
            if (dot) acc += fd[k] * (double)v;
        }
    }
    if (dot) block_store_partial(acc, w.rz[slot], by * gridDim.x + bx, sh);
}

// coarse face coefficients from the fine ones (the fine level may be the
// FP64 level 0 of the FP32 V-cycle)
template <typename TF, typename T> __global__ void __launch_bounds__(NTH) k_mg_faces(Lev<TF> L, Lev<T> C)
{
    const int I = blockIdx.x * TXP + (threadIdx.x & (TXP - 1)), J = blockIdx.y * TYR + (threadIdx.x >> 6);
    if (I >= C.nx || J >= C.ny) return;
    const int p = L.pitch;
    C.bx[(ptrdiff_t)J * C.pitch + I] = (T)(0.5 * ((double)ld(L.bx, p, 2 * I + 1, 2 * J) +
                                                  (double)ld(L.bx, p, 2 * I + 1, 2 * J + 1)));
    C.by[(ptrdiff_t)J * C.pitch + I] = (T)(0.5 * ((double)ld(L.by, p, 2 * I, 2 * J + 1) +
                                                  (double)ld(L.by, p, 2 * I + 1, 2 * J + 1)));
}

// FP32 copies of the level-0 faces
__global__ void __launch_bounds__(NTH) k_mg_faces0(Lev<double> L, Lev<float> C)
{
    const int I = blockIdx.x * TXP + (threadIdx.x & (TXP - 1)), J = blockIdx.y * TYR + (threadIdx.x >> 6);
    if (I >= C.nx || J >= C.ny) return;
    C.bx[(ptrdiff_t)J * C.pitch + I] = (float)ld(L.bx, L.pitch, I, J);
    C.by[(ptrdiff_t)J * C.pitch + I] = (float)ld(L.by, L.pitch, I, J);
}

// H = P G P, G the inverse of the bottom operator M = -A_B + s e0 e0^T (s its
// (0,0) entry: grounded at cell 0, symmetric positive definite) and P the
// projection to zero mean, in FP64: the bottom solve is then x = -H f.
// Numbered x fastest, M is banded: the x neighbours (the periodic one
// included) lie within nx - 1 of the diagonal, the y neighbours at nx.  Once
// per solve: k_mg_bottom_factor (one wavefront) factors M = L L^T in LDS in
// band storage and writes the band out; k_mg_bottom_cols solves L L^T g = e_c
// for one column c per lane (the lanes of a block share the band in LDS and
// keep their vectors there; no lane waits on another), writes g as row c of
// G (G is symmetric) and its mean; k_mg_bottom_proj forms
// (P G P)_rc = G_rc - m_r - m_c + mean of the m, m the column means (the row
// means of a symmetric G), in the cycle's type.
#define BW 64
// NX: the largest band this instantiation takes.  NX = 16 (bottoms up to 16
// wide, all the common sizes) uses a padded layout: row stride 17 and 16 zero
// rows after the last, so every loop below runs its full, compile-time
// length with no bounds tests (entries past the band stay zero through the
// factorisation); wider bottoms use stride nx + 1 and tested bounds.
template <int NX> struct BandLayout {
    static constexpr bool PAD = (NX == 16);
    __device__ static int stride(int nx) { return PAD ? NX + 1 : nx + 1; }
    __device__ static int rows(int nb) { return PAD ? nb + NX : nb; }
};

template <typename T, int NX> __global__ void __launch_bounds__(BW) k_mg_bottom_factor(Lev<T> B, double *Lg, double *dinv)
{
    using BL = BandLayout<NX>;
    constexpr int TRI_PER_LANE = (NX * (NX + 1) / 2 + BW - 1) / BW;
    extern __shared__ double lb[];         // lb[r W + d] = M[r][r - d] (L after the factorisation)
    const int nx = B.nx, ny = B.ny, nb = nx * ny, p = B.pitch, t = threadIdx.x;
    const int W = BL::stride(nx), RP = BL::rows(nb), nbw = BL::PAD ? NX : nx;
    const int ntri = nbw * (nbw + 1) / 2;
    // this lane's pairs (i, j), 1 <= j <= i <= nbw, of the trailing update:
    // q = t + BW u in the order (1,1), (2,1), (2,2), (3,1), ...
    int ti[TRI_PER_LANE], tj[TRI_PER_LANE];
#pragma unroll
    for (int u = 0; u < TRI_PER_LANE; u++) {
        const int q = t + BW * u;
        int i = 1, base = 0;
        while (base + i <= q) { base += i; i++; }
        ti[u] = (q < ntri) ? i : nbw + 1;                // nbw + 1: never active
        tj[u] = q - base + 1;
    }
    for (int q = t; q < RP * W; q += BW) lb[q] = 0.0;
    __syncthreads();
    // row r's diagonal and its couplings to the columns before it: the west
    // neighbour, the periodic east neighbour (of i = nx - 1) and the south one
    for (int r = t; r < nb; r += BW) {
        const int j = r / nx, i = r - j * nx;
        const int ie = (i + 1 == nx) ? 0 : i + 1, iw = (i == 0) ? nx - 1 : i - 1;
        const double ce = ld(B.bx, p, i, j), cw = ld(B.bx, p, iw, j);
        const double cn = (j + 1 < ny) ? (double)ld(B.by, p, i, j) : 0.0;
        const double cs = (j > 0) ? (double)ld(B.by, p, i, j - 1) : 0.0;
        double dg = ce + cw + cn + cs;
        if (ie == i) dg -= ce;             // nx = 1: the periodic image is the cell itself
        if (iw == i) dg -= cw;
        if (r == 0) dg += dg;              // grounding: the (0, 0) entry doubled
        lb[r * W] = dg;
        if (iw < i) lb[r * W + (i - iw)] -= cw;
        if (ie < i) lb[r * W + (i - ie)] -= ce;
        if (j > 0) lb[r * W + nx] -= cs;
    }
    __syncthreads();
    // right-looking band Cholesky (padded: the steps touch the zero rows past
    // nb and the zero entries past the band, which stay zero).  The block is
    // one wavefront, whose LDS accesses complete in program order, so a step
    // needs no workgroup barrier: it reads the pivot, column k (unscaled) and
    // its pairs' entries, then writes L[k][k], the scaled column and the
    // updated entries -- no two lanes write, or write and read, one entry in
    // a step (the updated entries have j >= 1, the column entries d = i), and
    // the next step's reads come after these writes.  The wave barriers keep
    // the compiler from moving LDS accesses across steps.
    for (int k = 0; k < nb; k++) {
        const int m = BL::PAD ? NX : min(nx, nb - 1 - k);
        const double piv = lb[k * W];
        const double ck = (t >= 1 && t <= m) ? lb[(k + t) * W + t] : 0.0;
        double mi[TRI_PER_LANE], mj[TRI_PER_LANE], mij[TRI_PER_LANE];
#pragma unroll
        for (int u = 0; u < TRI_PER_LANE; u++) {
            const int i = ti[u], j = tj[u];
            mi[u] = mj[u] = mij[u] = 0.0;
            if (i <= m) {
                mi[u] = lb[(k + i) * W + i];
                mj[u] = lb[(k + j) * W + j];
                mij[u] = lb[(k + i) * W + (i - j)];
            }
        }
        // 1 / sqrt(piv): the hardware estimate (about 23 bits) and two Newton
        // steps (quadratic: to FP64 rounding), far shorter a chain than an
        // IEEE sqrt and divide; L[k][k] = piv / sqrt(piv)
        double inv = __builtin_amdgcn_rsq(piv);
        inv = inv * (1.5 - 0.5 * piv * inv * inv);
        inv = inv * (1.5 - 0.5 * piv * inv * inv);
        const double dkk = piv * inv;
        __builtin_amdgcn_wave_barrier();
        if (t == 0) {
            lb[k * W] = dkk;
            dinv[k] = inv;
        }
        if (t >= 1 && t <= m) lb[(k + t) * W + t] = ck * inv;
#pragma unroll
        for (int u = 0; u < TRI_PER_LANE; u++) {           // M[k+i][k+j] -= L[k+i][k] L[k+j][k]
            const int i = ti[u], j = tj[u];
            if (i <= m) lb[(k + i) * W + (i - j)] = mij[u] - (mi[u] * inv) * (mj[u] * inv);
        }
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    for (int q = t; q < RP * W; q += BW) Lg[q] = lb[q];
}

// sum over d = 2 .. NX of a[d] w[d - 1] as a pairwise tree: these terms do
// not depend on the newest value w[0], so only one product and two adds of a
// step wait on the step before
template <int NX> __device__ __forceinline__ double band_sum(const double *a, const double *w)
{
    double p[NX];
#pragma unroll
    for (int d = 0; d < NX; d++) p[d] = (d >= 1) ? a[d + 1] * w[d] : 0.0;   // p[d] = a[d+1] w[d], d >= 1
#pragma unroll
    for (int h = 1; h < NX; h <<= 1)
#pragma unroll
        for (int d = 0; d + h < NX; d += 2 * h) p[d] += p[d + h];
    return p[0];
}

// columns c = blockIdx * cpb + lane of G from the band L and 1 / diag(L),
// both staged in LDS.  Each lane keeps the last NX values of its recurrence
// in registers and reads the next band row while the current one is used,
// so a step waits on one product, not on LDS.
template <int NX>
__global__ void __launch_bounds__(BW) k_mg_bottom_cols(int nx, int nb, int cpb, const double *Lg, const double *dinv,
                                                       double *G, double *cm)
{
    using BL = BandLayout<NX>;
    extern __shared__ double sm[];
    const int W = BL::stride(nx), RP = BL::rows(nb), t = threadIdx.x;
    double *lb = sm, *di = sm + (size_t)RP * W, *y = di + nb;    // y[r cpb + lane]
    for (int q = t; q < RP * W; q += BW) lb[q] = Lg[q];
    for (int q = t; q < nb; q += BW) di[q] = dinv[q];
    __syncthreads();
    const int c = blockIdx.x * cpb + t;
    if (t >= cpb || c >= nb) return;
    // a band entry, 0 past the band or past the last row (no test when padded)
    auto band = [&](int row, int d) -> double {
        if (BL::PAD) return lb[row * W + d];
        return (d <= nx && row < nb) ? lb[row * W + d] : 0.0;
    };
    double w[NX], a[NX + 1], an[NX + 1];
    // L y = e_c: every lane runs all rows (y_r = 0 before row c comes out of
    // the same recurrence), so the loops stay uniform; w[d - 1] = y_{r - d},
    // a[d] = L[r][r - d], a[0] = 1 / L[r][r]
#pragma unroll
    for (int d = 0; d < NX; d++) w[d] = 0.0;
#pragma unroll
    for (int d = 0; d <= NX; d++) a[d] = (d == 0) ? di[0] : band(0, d);
    for (int r = 0; r < nb; r++) {
        const int rn = (r + 1 < nb) ? r + 1 : r;
#pragma unroll
        for (int d = 0; d <= NX; d++) an[d] = (d == 0) ? di[rn] : band(rn, d);
        const double v = ((r == c ? 1.0 : 0.0) - (band_sum<NX>(a, w) + a[1] * w[0])) * a[0];
#pragma unroll
        for (int d = NX - 1; d >= 1; d--) w[d] = w[d - 1];
        w[0] = v;
        y[r * cpb + t] = v;
#pragma unroll
        for (int d = 0; d <= NX; d++) a[d] = an[d];
    }
    // L^T g = y, in place; w[d - 1] = g_{r + d}, a[d] = L[r + d][r],
    // a[0] = 1 / L[r][r]
#pragma unroll
    for (int d = 0; d < NX; d++) w[d] = 0.0;
#pragma unroll
    for (int d = 0; d <= NX; d++) a[d] = (d == 0) ? di[nb - 1] : band(nb - 1 + d, d);
    for (int r = nb - 1; r >= 0; r--) {
        const int rp = (r > 0) ? r - 1 : 0;
#pragma unroll
        for (int d = 0; d <= NX; d++) an[d] = (d == 0) ? di[rp] : band(rp + d, d);
        const double yr = y[r * cpb + t];
        const double v = (yr - (band_sum<NX>(a, w) + a[1] * w[0])) * a[0];
#pragma unroll
        for (int d = NX - 1; d >= 1; d--) w[d] = w[d - 1];
        w[0] = v;
        y[r * cpb + t] = v;
#pragma unroll
        for (int d = 0; d <= NX; d++) a[d] = an[d];
    }
    double s = 0.0;
    for (int r = 0; r < nb; r++) {
        const double g = y[r * cpb + t];
        G[(size_t)c * nb + r] = g;
        s += g;
    }
    cm[c] = s / nb;
}

// The padded (16-wide) form of k_mg_bottom_cols: the row loop is unrolled by
// 16, so the window of the last 16 values is a ring indexed by row % 16 at
// compile time (no register moves per step) and the band rows alternate
// between two register sets; rows are taken up to a multiple of 16, past nb
// the band and 1 / diag are zero, so those rows yield zero.
__global__ void __launch_bounds__(BW) k_mg_bottom_cols16(int nb, int cpb, const double *Lg, const double *dinv,
                                                         double *G, double *cm)
{
    constexpr int K = 16, W = K + 1;
    extern __shared__ double sm[];
    const int t = threadIdx.x, nbp = (nb + K - 1) / K * K, RP = nbp + K;
    double *lb = sm, *di = sm + (size_t)RP * W, *y = di + RP;    // y[r cpb + lane]
    for (int q = t; q < RP * W; q += BW) lb[q] = (q < (nb + K) * W) ? Lg[q] : 0.0;
    for (int q = t; q < RP; q += BW) di[q] = (q < nb) ? dinv[q] : 0.0;
    __syncthreads();
    const int c = blockIdx.x * cpb + t;
    if (t >= cpb || c >= nb) return;
    double w[K], a0[W], a1[W];
    // L y = e_c, from the first row of this block's columns (y is zero
    // before its column): w[r % K] = y_r, a[d] = L[r][r - d], a[0] = 1 / L[r][r]
    const int r0 = (blockIdx.x * cpb) / K * K;
#pragma unroll
    for (int q = 0; q < K; q++) w[q] = 0.0;
#pragma unroll
    for (int d = 0; d < W; d++) a0[d] = (d == 0) ? di[r0] : lb[r0 * W + d];
    for (int rb = r0; rb < nbp; rb += K) {
#pragma unroll
        for (int s2 = 0; s2 < K; s2++) {
            const int r = rb + s2;
            double *a = (s2 & 1) ? a1 : a0, *an = (s2 & 1) ? a0 : a1;
#pragma unroll
            for (int d = 0; d < W; d++) an[d] = (d == 0) ? di[r + 1] : lb[(r + 1) * W + d];
            double p[K];
#pragma unroll
            for (int d = 0; d < K; d++) p[d] = (d >= 1) ? a[d + 1] * w[(s2 - d - 1 + 2 * K) % K] : 0.0;
#pragma unroll
            for (int h = 1; h < K; h <<= 1)
#pragma unroll
                for (int d = 0; d + h < K; d += 2 * h) p[d] += p[d + h];
            const double v = ((r == c ? 1.0 : 0.0) - (p[0] + a[1] * w[(s2 - 1 + K) % K])) * a[0];
            w[s2] = v;
            y[r * cpb + t] = v;
        }
    }
    // L^T g = y in place, from the last row: w[r % K] = g_r, a[d] = L[r + d][r]
#pragma unroll
    for (int q = 0; q < K; q++) w[q] = 0.0;
#pragma unroll
    for (int d = 0; d < W; d++) a0[d] = (d == 0) ? di[nbp - 1] : lb[(nbp - 1 + d) * W + d];
    for (int rb = nbp - K; rb >= 0; rb -= K) {
#pragma unroll
        for (int s2 = K - 1; s2 >= 0; s2--) {
            const int r = rb + s2;
            double *a = ((K - 1 - s2) & 1) ? a1 : a0, *an = ((K - 1 - s2) & 1) ? a0 : a1;
            const int rp = (r > 0) ? r - 1 : 0;
#pragma unroll
            for (int d = 0; d < W; d++) an[d] = (d == 0) ? di[rp] : lb[(rp + d) * W + d];
            const double yr = (r >= r0) ? y[r * cpb + t] : 0.0;
            double p[K];
#pragma unroll
            for (int d = 0; d < K; d++) p[d] = (d >= 1) ? a[d + 1] * w[(s2 + d + 1) % K] : 0.0;
#pragma unroll
            for (int h = 1; h < K; h <<= 1)
#pragma unroll
                for (int d = 0; d + h < K; d += 2 * h) p[d] += p[d + h];
            const double v = (yr - (p[0] + a[1] * w[(s2 + 1) % K])) * a[0];
            w[s2] = v;
            y[r * cpb + t] = v;
        }
    }
    double sum = 0.0;
    for (int r = 0; r < nb; r++) {
        const double g = y[r * cpb + t];
        G[(size_t)c * nb + r] = g;
        sum += g;
    }
    cm[c] = sum / nb;
}

template <typename T> __global__ void __launch_bounds__(NTH) k_mg_bottom_proj(int nb, const double *G, const double *cm, T *H)
{
    __shared__ double sh[NTH];
    const int r = blockIdx.x, t = threadIdx.x;
    double s = 0.0;
    for (int q = t; q < nb; q += NTH) s += cm[q];
    sh[t] = s;
    __syncthreads();
    for (int w = NTH / 2; w > 0; w >>= 1) {
        if (t < w) sh[t] += sh[t + w];
        __syncthreads();
    }
    const double tm = sh[0] / nb, mr = cm[r];
    for (int c = t; c < nb; c += NTH) H[(size_t)r * nb + c] = (T)(((G[(size_t)r * nb + c] - mr) - cm[c]) + tm);
}

// ---------------------------------------------------------------------------
// Levels c0 .. nlev-1 in one workgroup of 1024 threads.  x and f of every
// level live in LDS, and the faces of every level below c0; those of level
// c0 are copied too when they fit (topf), else read from its arrays.  The
// kernel runs on one CU, so its time is instruction issue: each thread reads
// the coefficients and neighbour indices of its cells once per level and
// pass into registers, and a Gauss-Seidel update is then four LDS reads, the
// weighted sum and one store.
// ---------------------------------------------------------------------------
#define CT 1024
#define MG_LOW PVK_MG_LOW   // 4096 -> 1024 -> 256 -> 64 cells
static_assert(PVK_MG_LDS_CELLS <= 4 * CT, "level c0 is loaded as at most 4 cells a thread");

template <typename T> struct LdsLev {
    int nx, ny, fp;         // cells; pitch of bx, by
    int lnx;                // log2(nx) if nx is a power of two, else -1
    T *x, *f;               // LDS, compact
    const T *bx, *by;       // LDS (pitch nx) or the level's arrays
};

// q / n and q % n, by shift and mask when n = 2^l (l >= 0)
__device__ __forceinline__ void divmod(int q, int n, int l, int &d, int &r)
{
    d = (l >= 0) ? (q >> l) : q / n;
    r = q - d * n;
}
__device__ __forceinline__ int ilog2_or_neg(int n) { return (n > 0 && (n & (n - 1)) == 0) ? __ffs(n) - 1 : -1; }

// a / b of the Gauss-Seidel update: FP64 divides; FP32 multiplies by the
// hardware reciprocal (1 ulp), which keeps the V-cycle a fixed symmetric
// operator (both sweep directions use the same value)
template <typename T> __device__ __forceinline__ T gs_div(T a, T b) { return a / b; }
template <> __device__ __forceinline__ float gs_div<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// A cell of an in-LDS level as one thread smooths it: the LDS indices of the
// cell and of its four neighbours, the faces, 1 / (their sum) and f, read
// once per level and pass and kept in registers for all the half-sweeps of
// that pass.  A neighbour beyond a y wall has face 0 and is pointed at the
// neighbour on the other side (a smoothing level has at least two rows):
// a cell of the other colour, so already written when it is read (a cell of
// the same colour may still hold stale LDS, and 0 * NaN is NaN).
template <typename T> struct GsCell {
    int s, e, w, n, so;
    T ce, cw, cn, cs, id, f;
};

template <typename T> __device__ __forceinline__ void gs_cell(const LdsLev<T> &L, int i, int j, GsCell<T> &c)
{
    const int nx = L.nx, fp = L.fp, ie = (i + 1 == nx) ? 0 : i + 1, iw = (i == 0) ? nx - 1 : i - 1;
    c.s = j * nx + i;
    c.e = j * nx + ie;
    c.w = j * nx + iw;
    c.n = (j + 1 < L.ny) ? c.s + nx : c.s - nx;
    c.so = (j > 0) ? c.s - nx : c.s + nx;
    c.ce = L.bx[j * fp + i];
    c.cw = L.bx[j * fp + iw];
    c.cn = (j + 1 < L.ny) ? L.by[j * fp + i] : T(0);
    c.cs = (j > 0) ? L.by[(j - 1) * fp + i] : T(0);
    c.id = gs_div(T(1), c.ce + c.cw + c.cn + c.cs);
    c.f = L.f[c.s];
}

// Gauss-Seidel update of the cell from its neighbours' current x
template <typename T> __device__ __forceinline__ void gs_update(T *x, const GsCell<T> &c)
{
    x[c.s] = (c.ce * x[c.e] + c.cw * x[c.w] + c.cn * x[c.n] + c.cs * x[c.so] - c.f) * c.id;
}

// residual f - A x at the cell
template <typename T> __device__ __forceinline__ T gs_res(const T *x, const GsCell<T> &c)
{
    const T x0 = x[c.s];
    return c.f - (c.ce * (x[c.e] - x0) - c.cw * (x0 - x[c.w]) + c.cn * (x[c.n] - x0) - c.cs * (x0 - x[c.so]));
}

// The cells of in-LDS level L pair along x: pair p = (ip, j) holds cells
// 2 ip and 2 ip + 1 of row j, one red ((i + j) even) and one black.  Thread
// t owns pairs t and t + CT (a level has at most 2 CT pairs).
template <typename T> struct GsPairs {
    GsCell<T> r[2], b[2];
    int p[2];        // pair index, or -1
};

template <typename T> __device__ __forceinline__ void gs_pairs(const LdsLev<T> &L, GsPairs<T> &P)
{
    const int hx = L.nx >> 1, npair = hx * L.ny, lhx = (L.lnx > 0) ? L.lnx - 1 : -1;
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int q = threadIdx.x + k * CT;
        P.p[k] = -1;
        if (q < npair) {
            int j, ip;
            divmod(q, hx, lhx, j, ip);
            P.p[k] = q;
            gs_cell(L, 2 * ip + (j & 1), j, P.r[k]);
            gs_cell(L, 2 * ip + 1 - (j & 1), j, P.b[k]);
        }
    }
}

// Pre-smoothing of in-LDS level L (red, from x = 0 when zero, black, then
// nu - 1 more red-black pairs) and restriction: C.f = the sum of the four
// fine residuals, through rs (one residual sum per fine pair).
template <typename T> __device__ void lds_pre(const LdsLev<T> &L, const LdsLev<T> &C, T *rs, int nu, bool zero)
{
    GsPairs<T> P;
    gs_pairs(L, P);
    // From x = 0 the first red update is -f / diag.  Otherwise (the second
    // correction of a W) it is left out: the post-smoothing just ended with
    // a red half-sweep and no black cell has changed since, so it would
    // reproduce the same red values.
    if (zero) {
#pragma unroll
        for (int k = 0; k < 2; k++)
            if (P.p[k] >= 0) L.x[P.r[k].s] = -P.r[k].f * P.r[k].id;
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < 2; k++)
        if (P.p[k] >= 0) gs_update(L.x, P.b[k]);
    __syncthreads();
    for (int sw = 1; sw < nu; sw++) {
#pragma unroll
        for (int k = 0; k < 2; k++)
            if (P.p[k] >= 0) gs_update(L.x, P.r[k]);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 2; k++)
            if (P.p[k] >= 0) gs_update(L.x, P.b[k]);
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < 2; k++)
        if (P.p[k] >= 0) rs[P.p[k]] = gs_res(L.x, P.r[k]) + gs_res(L.x, P.b[k]);
    __syncthreads();
    const int hx = L.nx >> 1;
    for (int q = threadIdx.x; q < C.nx * C.ny; q += CT) {
        int J, I;
        divmod(q, C.nx, C.lnx, J, I);
        C.f[q] = rs[2 * J * hx + I] + rs[(2 * J + 1) * hx + I];
    }
    __syncthreads();
}

// parent (coarse cell) of the cell at LDS index q of level L
template <typename T> __device__ __forceinline__ int gs_parent(const LdsLev<T> &L, const LdsLev<T> &C, int q)
{
    int j, i;
    divmod(q, L.nx, L.lnx, j, i);
    return (j >> 1) * C.nx + (i >> 1);
}

// Coarse correction x += xc(parent) and post-smoothing (black first) of
// in-LDS level L; C holds the coarse result.  The correction rides in the
// first black half-sweep: its red neighbours are read as x + xc(their
// parent), the value the separate correction would have stored; the red
// cells themselves need not be corrected, the red half-sweep that follows
// overwrites them from the updated black ones.
template <typename T> __device__ void lds_post(const LdsLev<T> &L, const LdsLev<T> &C, int nu)
{
    GsPairs<T> P;
    gs_pairs(L, P);
#pragma unroll
    for (int k = 0; k < 2; k++)
        if (P.p[k] >= 0) {
            const GsCell<T> &c = P.b[k];
            const T *x = L.x, *xc = C.x;
            const T xe = x[c.e] + xc[gs_parent(L, C, c.e)], xw = x[c.w] + xc[gs_parent(L, C, c.w)];
            const T xn = x[c.n] + xc[gs_parent(L, C, c.n)], xs = x[c.so] + xc[gs_parent(L, C, c.so)];
            L.x[c.s] = (c.ce * xe + c.cw * xw + c.cn * xn + c.cs * xs - c.f) * c.id;
        }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 2; k++)
        if (P.p[k] >= 0) gs_update(L.x, P.r[k]);
    __syncthreads();
    for (int sw = 1; sw < nu; sw++) {
#pragma unroll
        for (int k = 0; k < 2; k++)
            if (P.p[k] >= 0) gs_update(L.x, P.b[k]);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 2; k++)
            if (P.p[k] >= 0) gs_update(L.x, P.r[k]);
        __syncthreads();
    }
}

// in-LDS level l: x, f, bx, by (level c0 without topf: x, f)
template <typename T> __device__ __forceinline__ LdsLev<T> lds_level(const pvk_mg &m, T *lds, int l)
{
    T *q = lds;
    for (int k = m.c0; k < l; k++) q += ((k == m.c0 && !m.topf) ? 2 : 4) * m.L[k].nx * m.L[k].ny;
    const int nc = m.L[l].nx * m.L[l].ny;
    LdsLev<T> v;
    v.nx = m.L[l].nx;
    v.ny = m.L[l].ny;
    v.lnx = ilog2_or_neg(v.nx);
    v.x = q;
    v.f = q + nc;
    if (l == m.c0 && !m.topf) {
        v.fp = m.L[l].pitch;
        v.bx = (const T *)m.L[l].bx;
        v.by = (const T *)m.L[l].by;
    } else {
        v.fp = v.nx;
        v.bx = q + 2 * nc;
        v.by = q + 3 * nc;
    }
    return v;
}

// residual sums of the restriction (one per fine pair of the largest
// smoothing level, c0), after the levels
template <typename T> __device__ __forceinline__ T *lds_scratch(const pvk_mg &m, T *lds)
{
    T *q = lds;
    for (int k = m.c0; k < m.nlev; k++) q += ((k == m.c0 && !m.topf) ? 2 : 4) * m.L[k].nx * m.L[k].ny;
    return q;
}

// the bottom solve's partial sums (4 PVK_MG_BOTTOM doubles) after the
// restriction scratch, 8-byte aligned
template <typename T> __device__ __forceinline__ double *lds_bottom(const pvk_mg &m, T *lds)
{
    T *q = lds_scratch(m, lds);
    if (m.nlev > m.c0 + 1) q += (m.L[m.c0].nx / 2) * m.L[m.c0].ny;
    return (double *)(((uintptr_t)q + 7) & ~(uintptr_t)7);
}

// bottom level, x = -H f (H = P G P, symmetric, in the cycle's type: the
// FP32 cycle reads half the bytes), H read from memory (up to 512 KiB; the
// one workgroup streams it through one CU, so many loads are kept in
// flight).  Thread (q, r) = (threadIdx / 256, threadIdx % 256) sums a
// quarter of the columns for row r, in FP64, reading H[c][r] = H[r][c]:
// consecutive threads read consecutive entries of row c.  The four quarters
// are added in order through LDS.
template <typename T> __device__ __forceinline__ void bottom_solve(const LdsLev<T> &B, const T *H, double *part)
{
    static_assert(CT == 4 * PVK_MG_BOTTOM, "four quarters of the bottom's columns, one row a thread");
    const int n = B.nx * B.ny, r = threadIdx.x % PVK_MG_BOTTOM, q = threadIdx.x / PVK_MG_BOTTOM;
    const int qn = (n + 3) / 4, c0 = q * qn, c1 = min(n, c0 + qn);
    double v = 0.0;
    if (r < n) {
#pragma unroll 16
        for (int c = c0; c < c1; c++) v -= (double)H[(size_t)c * n + r] * (double)B.f[c];
    }
    part[threadIdx.x] = v;
    __syncthreads();
    if (q == 0 && r < n)
        B.x[r] = (T)(((part[r] + part[PVK_MG_BOTTOM + r]) + part[2 * PVK_MG_BOTTOM + r]) + part[3 * PVK_MG_BOTTOM + r]);
    __syncthreads();
}

template <typename T> __global__ void __launch_bounds__(CT) k_mg_coarse(pvk_mg m, pvk_scal *sc)
{
    extern __shared__ double lds_raw[];
    __shared__ LdsLev<T> lv[PVK_MG_LOW + 1];      // levels c0 .. nb
    T *lds = (T *)lds_raw;
    const int nb = m.nlev - 1;
    // The loads below are valid whatever the stop flag says, so they are
    // issued before it is read: one memory round trip instead of two.
    // f of level c0 (and its faces with topf), the faces of the levels below
    // it: every global load of this thread is issued before the first LDS
    // store, so the loads wait for memory once, not once per level and pass.
    // Level c0 has at most 4 CT cells, each level below at most CT, and there
    // are at most MG_LOW of those (pv_mg_create).
    T f0[4], bx0[4], by0[4], lbx[MG_LOW], lby[MG_LOW];
    const Lev<T> G0 = lev<T>(m.L[m.c0]);
    const int n0 = G0.nx * G0.ny, l0 = ilog2_or_neg(G0.nx);
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int q = threadIdx.x + u * CT;
        f0[u] = bx0[u] = by0[u] = T(0);
        if (q < n0) {
            int j, i;
            divmod(q, G0.nx, l0, j, i);
            f0[u] = (m.c0 == 0 && m.r0) ? (T)m.r0[(ptrdiff_t)j * m.p0 + i] : ld(G0.f, G0.pitch, i, j);
            if (m.topf) {
                bx0[u] = ld(G0.bx, G0.pitch, i, j);
                by0[u] = ld(G0.by, G0.pitch, i, j);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < MG_LOW; k++) {
        const Lev<T> G = lev<T>(m.L[min(m.c0 + 1 + k, nb)]);
        const bool in = (m.c0 + 1 + k <= nb) && (int)threadIdx.x < G.nx * G.ny;
        lbx[k] = in ? G.bx[threadIdx.x] : T(0);      // levels below c0: pitch = width
        lby[k] = in ? G.by[threadIdx.x] : T(0);
    }
    if (sc->p_done) return;
    // the level descriptors, formed once: the schedule below reads them from
    // LDS instead of walking the kernel argument with a run-time index
    if (threadIdx.x <= (unsigned)(nb - m.c0)) lv[threadIdx.x] = lds_level(m, lds, m.c0 + (int)threadIdx.x);
    double *part = lds_bottom(m, lds);             // the bottom solve's partial sums
    {
        const LdsLev<T> L0 = lds_level(m, lds, m.c0);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int q = threadIdx.x + u * CT;
            if (q < n0) {
                L0.f[q] = f0[u];
                if (m.topf) {
                    ((T *)L0.bx)[q] = bx0[u];
                    ((T *)L0.by)[q] = by0[u];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < MG_LOW; k++) {
            if (m.c0 + 1 + k > nb) break;
            const LdsLev<T> Ll = lds_level(m, lds, m.c0 + 1 + k);
            if ((int)threadIdx.x < Ll.nx * Ll.ny) {
                ((T *)Ll.bx)[threadIdx.x] = lbx[k];
                ((T *)Ll.by)[threadIdx.x] = lby[k];
            }
        }
    }
    __syncthreads();
    T *rs = lds_scratch(m, lds);
    // The in-LDS levels: a V-cycle, except that the levels in m.wmask take
    // two coarse corrections, smoothing between them (pre-smooth, correct,
    // post-smooth, pre-smooth, correct, post-smooth: the error propagation
    // S^T C S S^T C S stays symmetric in the energy inner product).  One
    // schedule for all threads: level l, going down or coming back up, and
    // the levels whose second correction is under way.
    int l = m.c0;
    bool down = true;
    unsigned second = 0u;
    for (;;) {
        if (l == nb) {
            bottom_solve(lv[nb - m.c0], (const T *)m.H, part);
            l--;
            down = false;
            if (l < m.c0) break;
            continue;
        }
        const LdsLev<T> Ll = lv[l - m.c0], Lc = lv[l + 1 - m.c0];
        if (down) {
            lds_pre(Ll, Lc, rs, m.nu, true);
            l++;
            continue;
        }
        lds_post(Ll, Lc, m.nu);
        if (((m.wmask >> l) & 1) && !((second >> l) & 1u)) {
            second |= 1u << l;
            lds_pre(Ll, Lc, rs, m.nu, false);
            l++;
            down = true;
            continue;
        }
        second &= ~(1u << l);
        if (l == m.c0) break;
        l--;
    }
    {
        const Lev<T> G = lev<T>(m.L[m.c0]);
        const LdsLev<T> top = lds_level(m, lds, m.c0);
        for (int q = threadIdx.x; q < G.nx * G.ny; q += CT) {
            int j, i;
            divmod(q, G.nx, top.lnx, j, i);
            if (m.c0 == 0 && m.z0) m.z0[(ptrdiff_t)j * m.p0 + i] = (double)top.x[q];
            else G.x[(ptrdiff_t)j * G.pitch + i] = top.x[q];
        }
    }
}

// stop test alone, and <r, z> alone (grids small enough for one workgroup)
__global__ void __launch_bounds__(NTH) k_mg_check(pvk_scal *sc, pvk_pcg w, int n, double tol, int maxit)
{
    __shared__ double sh[NTH];
    if (sc->p_done || n == -3) return;
    mg_stop_test(sc, w, n, tol, maxit, sh);
}
__global__ void __launch_bounds__(NTH) k_mg_dot(const pvk_scal *sc, pvk_pcg w, pvk_grid g, int slot)
{
    __shared__ double sh[NTH];
    if (sc->p_done) return;
    const int P = g.P;
    double acc = 0.0;
    for (int q = threadIdx.x; q < g.NXM * g.NYM; q += NTH) {
        const int j = q / g.NXM, i = q - j * g.NXM;
        acc += AT(w.r, i, j) * AT(w.z, i, j);
    }
    block_store_partial(acc, w.rz[slot], 0, sh);
}

// ---------------------------------------------------------------------------
// CG vector kernels, one block per row
// ---------------------------------------------------------------------------
// The pressure operator is singular (constants are its null space) and the
// V-cycle, unlike the FFT solve, does not drop the mean of its input: a mean
// left in r by rounding in A d is a part of the system no iteration can
// reduce, and once ||r|| comes down to it the CG stalls.  So the residual is
// kept at zero mean: every A d is used less its mean (sum(A d) is zero in
// exact arithmetic; the partials come from the kernel that forms A d).
//
// start: r = rhs - mean (phi = 0), or with warm != 0 r = rhs - mean - (Ad - mean(Ad))
// (Ad = A phi0, phi kept); partial <r, r>
__global__ void __launch_bounds__(NTH) k_mg_start(pvk_grid g, const pvk_scal *sc, pvk_pcg w, int warm)
{
    __shared__ double sh[NTH];
    const int P = g.P, j = xcd_band(blockIdx.x, gridDim.x);
    const double ncell = (double)g.NXM * (double)g.NYM;
    const double mean = block_sum(w.mean, w.nmean, sh) / ncell;
    const double mad = (warm && w.sad) ? block_sum(w.sad, w.ndad, sh) / ncell : 0.0;
    double acc = 0.0;
    for (int i = threadIdx.x; i < g.NXM; i += NTH) {
        double r = AT(w.rhs, i, j) - mean;
        if (warm) r -= AT(w.Ad, i, j) - mad;
        else AT(w.phi, i, j) = 0.0;
        AT(w.r, i, j) = r;
        acc += r * r;
    }
    block_store_partial(acc, w.rr, j, sh);
    (void)sc;
}

// alpha = <r, z> / <d, A d>; phi += alpha d; r -= alpha (A d - mean(A d));
// partial <r, r>
__global__ void __launch_bounds__(NTH) k_mg_update(pvk_grid g, const pvk_scal *sc, pvk_pcg w, int n)
{
    __shared__ double sh[3 * NTH];
    if (sc->p_done) return;
    const int P = g.P, j = xcd_band(blockIdx.x, gridDim.x);
    double rz, dad, sad = 0.0;
    if (w.sad) block_sum3(w.rz[n & 1], w.nrz, w.dad[n & 1], w.ndad, w.sad, w.ndad, sh, rz, dad, sad);
    else block_sum2(w.rz[n & 1], w.nrz, w.dad[n & 1], w.ndad, sh, rz, dad);
    const double alpha = rz / dad;
    const double mad = sad / ((double)g.NXM * (double)g.NYM);
    const double *d = w.d[n & 1];
    double acc = 0.0;
    for (int i = threadIdx.x; i < g.NXM; i += NTH) {
        AT(w.phi, i, j) += alpha * AT(d, i, j);
        const double r = AT(w.r, i, j) - alpha * (AT(w.Ad, i, j) - mad);
        AT(w.r, i, j) = r;
        acc += r * r;
    }
    block_store_partial(acc, w.rr, j, sh);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// tile shape by level width: 0 = 64 x 16, 1 = 32 x 16, 2 = 32 x 8
// (level 0: m->sh0 when it is set)
static int tile_shape(const pvk_mg *m, int l)
{
    const pvk_mglev &L = m->L[l];
    if (l == 0 && m->sh0 >= 0) return m->sh0;
    return (L.nx > 512) ? 0 : (L.nx > 256 ? 1 : 2);
}
static dim3 tile_grid(const pvk_mg *m, int l)
{
    const pvk_mglev &L = m->L[l];
    const int sh = tile_shape(m, l), tx = sh ? 32 : 64, ty = (sh == 2) ? 8 : 16;
    return dim3((L.nx + tx - 1) / tx, (L.ny + ty - 1) / ty);
}
static dim3 cell_grid(const pvk_mglev &L) { return dim3((L.nx + TXP - 1) / TXP, (L.ny + TYR - 1) / TYR); }

// LDS bytes of the one-workgroup kernel
extern "C" size_t pvk_mg_lds_bytes(const pvk_mg *m)
{
    size_t n = 0;
    for (int l = m->c0; l < m->nlev; l++) n += ((l == m->c0 && !m->topf) ? 2 : 4) * (size_t)m->L[l].nx * m->L[l].ny;
    if (m->nlev > m->c0 + 1) n += (size_t)(m->L[m->c0].nx / 2) * m->L[m->c0].ny;   // restriction scratch
    // + the bottom solve's partial sums (FP64, 8-byte aligned)
    return n * (m->fp32 ? sizeof(float) : sizeof(double)) + 8 + sizeof(double) * 4 * PVK_MG_BOTTOM;
}

extern "C" int pvk_mg_nrz(const pvk_mg *m)
{
    if (m->c0 == 0) return 1;
    const dim3 gr = tile_grid(m, 0);
    return (int)(gr.x * gr.y);
}

// the bottom matrix of a solve: G column by column, then P G P.  Above 64 KiB
// of LDS (bottoms wider than 16 columns) the factorisation kernel must be
// allowed it, raised once per device under a lock as for k_mg_coarse.
// above 64 KiB of dynamic LDS a kernel must be allowed it (gfx950: up to 160
// KiB per workgroup); raised once per device and kernel, under a lock
static void allow_lds(const void *k, int slot)
{
    static std::mutex mu;
    static unsigned char raised[64][4];
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(mu);
    if (dev >= 0 && dev < 64 && !raised[dev][slot]) {
        if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) == hipSuccess)
            raised[dev][slot] = 1;
        else
            (void)hipGetLastError();
    }
}

// the bottom matrix of a solve: factor, columns of G, then P G P
template <typename T> static void bottom_setup(const pvk_mg *m, hipStream_t s)
{
    const pvk_mglev &Lb = m->L[m->nlev - 1];
    const int nx = Lb.nx, nb = nx * Lb.ny, pad = (nx <= 16);
    const size_t W = pad ? 17 : (size_t)nx + 1, RP = pad ? (size_t)nb + 16 : (size_t)nb;
    const size_t band = sizeof(double) * RP * W, budget = 160 * 1024;
    int cpb = BW;                                   // columns per block: as many as LDS holds
    while (cpb > 1 && band + sizeof(double) * (size_t)nb * (cpb + 1) > budget) cpb >>= 1;
    const size_t need_cols = band + sizeof(double) * (size_t)nb * (cpb + 1);
    double *dinv = m->Lb + (size_t)PVK_MG_BOTTOM * (PVK_MG_BOTTOM_NX + 1);
    if (pad) {
        hipLaunchKernelGGL((k_mg_bottom_factor<T, 16>), dim3(1), dim3(BW), band, s, lev<T>(Lb), m->Lb, dinv);
        const size_t nbp = ((size_t)nb + 15) / 16 * 16;
        int cp = BW;
        while (cp > 1 && sizeof(double) * ((nbp + 16) * 18 + nbp * cp) > budget) cp >>= 1;
        const size_t need16 = sizeof(double) * ((nbp + 16) * 18 + nbp * cp);
        if (need16 > 65536) allow_lds((const void *)k_mg_bottom_cols16, 2);
        hipLaunchKernelGGL(k_mg_bottom_cols16, dim3((nb + cp - 1) / cp), dim3(BW), need16, s, nb, cp,
                           (const double *)m->Lb, (const double *)dinv, m->G, m->cm);
    } else {
        if (band > 65536) allow_lds((const void *)k_mg_bottom_factor<T, PVK_MG_BOTTOM_NX>, sizeof(T) == 4 ? 0 : 1);
        hipLaunchKernelGGL((k_mg_bottom_factor<T, PVK_MG_BOTTOM_NX>), dim3(1), dim3(BW), band, s, lev<T>(Lb), m->Lb,
                           dinv);
        if (need_cols > 65536) allow_lds((const void *)k_mg_bottom_cols<PVK_MG_BOTTOM_NX>, 3);
        hipLaunchKernelGGL(k_mg_bottom_cols<PVK_MG_BOTTOM_NX>, dim3((nb + cpb - 1) / cpb), dim3(BW), need_cols, s,
                           nx, nb, cpb, (const double *)m->Lb, (const double *)dinv, m->G, m->cm);
    }
    hipLaunchKernelGGL(k_mg_bottom_proj<T>, dim3(nb), dim3(NTH), 0, s, nb, (const double *)m->G,
                       (const double *)m->cm, (T *)m->H);
}

// per solve: coarse faces of every level (FP32: and the level-0 copies) and,
// when asked, the bottom inverse, from the FP64 level-0 faces bx0, by0 (pitch pf)
extern "C" void pvk_mg_setup(const pvk_mg *m, const double *bx0, const double *by0, int pf, int bottom,
                             hipStream_t s)
{
    Lev<double> F0;
    F0.nx = m->L[0].nx;
    F0.ny = m->L[0].ny;
    F0.pitch = pf;
    F0.x = F0.f = nullptr;
    F0.bx = (double *)bx0;
    F0.by = (double *)by0;
    if (m->fp32) {
        hipLaunchKernelGGL(k_mg_faces0, cell_grid(m->L[0]), dim3(NTH), 0, s, F0, lev<float>(m->L[0]));
        if (m->nlev > 1)
            hipLaunchKernelGGL((k_mg_faces<double, float>), cell_grid(m->L[1]), dim3(NTH), 0, s, F0,
                               lev<float>(m->L[1]));
        for (int l = 1; l + 1 < m->nlev; l++)
            hipLaunchKernelGGL((k_mg_faces<float, float>), cell_grid(m->L[l + 1]), dim3(NTH), 0, s,
                               lev<float>(m->L[l]), lev<float>(m->L[l + 1]));
        if (bottom) bottom_setup<float>(m, s);
    } else {
        for (int l = 0; l + 1 < m->nlev; l++)
            hipLaunchKernelGGL((k_mg_faces<double, double>), cell_grid(m->L[l + 1]), dim3(NTH), 0, s,
                               lev<double>(m->L[l]), lev<double>(m->L[l + 1]));
        if (bottom) bottom_setup<double>(m, s);
    }
}

// One V-cycle z = V(r) for CG iteration n is, in order: pvk_mg_pre for
// levels 0 .. c0-1, pvk_mg_coarse, pvk_mg_post for levels c0-1 .. 0.  The
// stop test of n runs in the kernel after the one that formed r_{n+1}
// (n = -3: none); <r, z> partials go to w->rz[(n < 0) ? 0 : (n + 1) & 1].
// With tiled levels (c0 > 0) the level-0 pre kernel of iteration n >= 0 does
// the CG step, r_n -> r_{n+1}, between w->r and w->r2 by parity; without
// them (c0 = 0) pvk_mg_update does it in place in w->r.
static int rz_slot(int n) { return (n < 0) ? 0 : ((n + 1) & 1); }
static double *r_of(const pvk_mg *m, const pvk_pcg *w, int k)     // r_k, k >= 0
{
    return (m->c0 > 0 && (k & 1)) ? w->r2 : w->r;
}
// the residual a V-cycle of iteration n reads: r_{n+1} (r_0 for n < 0)
static const double *r_in(const pvk_mg *m, const pvk_pcg *w, int n) { return r_of(m, w, n < 0 ? 0 : n + 1); }

#define MG_SHAPES(KERNEL, NU, ...)                                                                         \
    do {                                                                                                   \
        if (sh == 0) hipLaunchKernelGGL((KERNEL<NU, 0, T>), gr, dim3(NTH), 0, s, __VA_ARGS__);             \
        else if (sh == 1) hipLaunchKernelGGL((KERNEL<NU, 1, T>), gr, dim3(NTH), 0, s, __VA_ARGS__);        \
        else hipLaunchKernelGGL((KERNEL<NU, 2, T>), gr, dim3(NTH), 0, s, __VA_ARGS__);                     \
    } while (0)

template <typename T>
static void pre_t(const pvk_mg *m, int l, pvk_scal *sc, const pvk_pcg *w, int n, double tol, int maxit,
                  hipStream_t s)
{
    const dim3 gr = tile_grid(m, l);
    const Lev<T> L = lev<T>(m->L[l]), C = lev<T>(m->L[l + 1]);
    const int step = (l == 0 && n >= 0);
    const double *r0 = (l == 0) ? (step ? r_of(m, w, n) : r_in(m, w, n)) : nullptr;
    double *rnew = step ? r_of(m, w, n + 1) : nullptr;
    const int sh = tile_shape(m, l), chk = (l == 1), nc = m->L[0].nx * m->L[0].ny;
    if (m->nu == 1) MG_SHAPES(k_mg_pre, 1, L, C, r0, rnew, m->p0, sc, *w, n, tol, maxit, chk, step, nc);
    else MG_SHAPES(k_mg_pre, 2, L, C, r0, rnew, m->p0, sc, *w, n, tol, maxit, chk, step, nc);
    // levels 0 .. 1 only: the stop test of the residual pre level 0 formed
    if (l == 0 && m->c0 == 1 && n != -3)
        hipLaunchKernelGGL(k_mg_check, dim3(1), dim3(NTH), 0, s, sc, *w, n, tol, maxit);
}

template <typename T> static void post_t(const pvk_mg *m, int l, pvk_scal *sc, const pvk_pcg *w, int n, hipStream_t s)
{
    const dim3 gr = tile_grid(m, l);
    const Lev<T> L = lev<T>(m->L[l]), C = lev<T>(m->L[l + 1]);
    const double *r0 = (l == 0) ? r_in(m, w, n) : nullptr;
    double *z0 = (l == 0) ? m->z0 : nullptr;
    const pvk_scal *cs = sc;
    const int sh = tile_shape(m, l), slot = rz_slot(n), dot = (l == 0);
    if (m->nu == 1) MG_SHAPES(k_mg_post, 1, L, C, r0, z0, m->p0, cs, *w, dot, slot);
    else MG_SHAPES(k_mg_post, 2, L, C, r0, z0, m->p0, cs, *w, dot, slot);
}

extern "C" void pvk_mg_pre(const pvk_mg *m, int l, pvk_scal *sc, const pvk_pcg *w, int n, double tol, int maxit,
                           hipStream_t s)
{
    if (m->fp32) pre_t<float>(m, l, sc, w, n, tol, maxit, s);
    else pre_t<double>(m, l, sc, w, n, tol, maxit, s);
}

extern "C" void pvk_mg_coarse(const pvk_mg *m, const pvk_grid *g, pvk_scal *sc, const pvk_pcg *w, int n,
                              double tol, int maxit, hipStream_t s)
{
    // above 64 KiB of dynamic LDS the kernel must be allowed it (gfx950: up to
    // 160 KiB per workgroup).  Handles of any size may launch from several host
    // threads at once, so the limit is raised once per device and kernel, to
    // the whole 160 KiB, under a lock; a refusal leaves the launch to report
    // the error.
    static std::mutex mu;
    static unsigned char raised[64][2];
    const size_t need = pvk_mg_lds_bytes(m);
    if (need > 65536) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        std::lock_guard<std::mutex> lk(mu);
        if (dev >= 0 && dev < 64 && !raised[dev][m->fp32]) {
            const void *k = m->fp32 ? (const void *)k_mg_coarse<float> : (const void *)k_mg_coarse<double>;
            if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) == hipSuccess)
                raised[dev][m->fp32] = 1;
            else
                (void)hipGetLastError();
        }
    }
    if (m->c0 == 0) hipLaunchKernelGGL(k_mg_check, dim3(1), dim3(NTH), 0, s, sc, *w, n, tol, maxit);
    if (m->fp32) hipLaunchKernelGGL(k_mg_coarse<float>, dim3(1), dim3(CT), need, s, *m, sc);
    else hipLaunchKernelGGL(k_mg_coarse<double>, dim3(1), dim3(CT), need, s, *m, sc);
    if (m->c0 == 0)
        hipLaunchKernelGGL(k_mg_dot, dim3(1), dim3(NTH), 0, s, (const pvk_scal *)sc, *w, *g, rz_slot(n));
}

extern "C" void pvk_mg_post(const pvk_mg *m, int l, pvk_scal *sc, const pvk_pcg *w, int n, hipStream_t s)
{
    if (m->fp32) post_t<float>(m, l, sc, w, n, s);
    else post_t<double>(m, l, sc, w, n, s);
}

// phi = src unless the solve has already stopped (a zero right-hand side:
// phi stays 0)
__global__ void __launch_bounds__(NTH) k_mg_take(pvk_grid g, const pvk_scal *sc, const double *src, double *phi)
{
    if (sc->p_done) return;
    const int P = g.P, j = blockIdx.x;
    for (int i = threadIdx.x; i < g.NXM; i += NTH) AT(phi, i, j) = AT(src, i, j);
}

extern "C" void pvk_mg_check(pvk_scal *sc, const pvk_pcg *w, int n, double tol, int maxit, hipStream_t s)
{
    hipLaunchKernelGGL(k_mg_check, dim3(1), dim3(NTH), 0, s, sc, *w, n, tol, maxit);
}

extern "C" void pvk_mg_take(const pvk_grid *g, const pvk_scal *sc, const double *src, double *phi, hipStream_t s)
{
    hipLaunchKernelGGL(k_mg_take, dim3(g->NYM), dim3(NTH), 0, s, *g, sc, src, phi);
}

extern "C" void pvk_mg_start(const pvk_grid *g, pvk_scal *sc, const pvk_pcg *w, int warm, hipStream_t s)
{
    hipLaunchKernelGGL(k_mg_start, dim3(g->NYM), dim3(NTH), 0, s, *g, (const pvk_scal *)sc, *w, warm);
}

extern "C" void pvk_mg_update(const pvk_grid *g, pvk_scal *sc, const pvk_pcg *w, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_mg_update, dim3(g->NYM), dim3(NTH), 0, s, *g, (const pvk_scal *)sc, *w, n);
}
