// pv_kernels.hip -- gfx950 kernels of the PARTIES 2-D two-phase RK3 substep.
//
// Every kernel is HBM-bound FP64 stencil or vector work.  Launch shape: 64 x 4
// threads per block, one cell per thread, x (unit stride) across the 64-wide
// wavefront so each wave row is one coalesced 512-byte segment.  Built with
// -ffp-contract=off: every expression rounds as written, in the operand order
// of the reference (PARTIES/src), so stencil outputs match the CPU oracle bit
// for bit and only reduction order differs.
//
// Reductions are deterministic: each block writes one partial per sum, a
// one-block kernel adds the partials in block order and then runs the solver's
// scalar logic (alpha, beta, stop tests) on the device.  Solver loops never
// need a host round trip per iteration; kernels early-exit on a device flag.
#include <hip/hip_runtime.h>
#include <math.h>
#include "pv_kernels.h"
#include "pv_stencil.h"

#define AT(a, i, j) (a)[(ptrdiff_t)(j) * P + (i)]
#define BX 64
#define BY 4
#define NT (BX * BY)


static inline dim3 grid2d(const pvk_grid *g) { return dim3(g->nbx, g->nby); }

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
__device__ __forceinline__ int flat_tid() { return threadIdx.y * BX + threadIdx.x; }
__device__ __forceinline__ int flat_bid() { return blockIdx.y * gridDim.x + blockIdx.x; }

// Sum NS values over the block in a fixed order and store them as this
// block's partials part[k * nb + bid].
template <int NS>
__device__ __forceinline__ void block_partials(double (&v)[NS], double *part, int nb)
{
    __shared__ double sh[NS][NT / 64];
    const int tid = flat_tid(), lane = tid & 63, w = tid >> 6;
#pragma unroll
    for (int k = 0; k < NS; k++) {
        double x = v[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if (lane == 0) sh[k][w] = x;
    }
    __syncthreads();
    if (tid < NS) {
        double s = sh[tid][0];
#pragma unroll
        for (int q = 1; q < NT / 64; q++) s += sh[tid][q];
        part[(size_t)tid * nb + flat_bid()] = s;
    }
}

template <int NS>
__device__ __forceinline__ void block_max_partials(double (&v)[NS], double *part, int nb)
{
    __shared__ double sh[NS][NT / 64];
    const int tid = flat_tid(), lane = tid & 63, w = tid >> 6;
#pragma unroll
    for (int k = 0; k < NS; k++) {
        double x = v[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_down(x, off, 64));
        if (lane == 0) sh[k][w] = x;
    }
    __syncthreads();
    if (tid < NS) {
        double s = sh[tid][0];
#pragma unroll
        for (int q = 1; q < NT / 64; q++) s = fmax(s, sh[tid][q]);
        part[(size_t)tid * nb + flat_bid()] = s;
    }
}

// One-block final sum of partial set k (fixed order: strided serial, then tree).
__device__ double final_sum(const double *part, int nb, int k, bool is_max)
{
    __shared__ double sh[256];
    const int tid = threadIdx.x;
    double s = is_max ? 0.0 : 0.0;
    for (int b = tid; b < nb; b += 256) {
        double x = part[(size_t)k * nb + b];
        s = is_max ? fmax(s, x) : s + x;
    }
    sh[tid] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) sh[tid] = is_max ? fmax(sh[tid], sh[tid + w]) : sh[tid] + sh[tid + w];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------
// ghost fills (final state of the reference's BC + Sendrecv sequences)
// ---------------------------------------------------------------------------

// kind: 0 = CH zero-gradient (VOF_DIFFUSE.c:571-683, 3 layers),
//       1 = pressure Neumann (psolve_cg.c:358-529, 1 layer) + cell halo,
//       2 = cell halo only, 3 = u-type halo only,
//       4 = u velocity: -mirror row -1 / NYM + u-type halo (Velocity.c:2600-2682),
//       5 = v velocity NORMAL: rows 0, NYM = 0 + cell halo (Velocity.c:2566-2644),
//       6 = v velocity CG: cell halo only.
__global__ void k_ghost(pvk_grid g, double *f0, double *f1, double *f2, double *f3, int nf, int kind)
{
    const int P = g.P, NXM = g.NXM, NYM = g.NYM;
    const int nrow = NXM + 7;                       // i in [-3, NXM+3]
    const int nrowcells = 6 * nrow;                 // 3 ghost rows top and bottom
    const int ncol = (kind == 3 || kind == 4) ? 7 : 6;
    const int ncolcells = ncol * NYM;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrowcells + ncolcells) return;
    int i, j;
    if (t < nrowcells) {
        int r = t / nrow;
        i = t % nrow - 3;
        j = (r < 3) ? -1 - r : NYM + (r - 3);
    } else {
        int c = (t - nrowcells) / NYM;
        j = (t - nrowcells) % NYM;
        if (ncol == 7) i = (c < 4) ? -c : NXM + 1 + (c - 4);
        else i = (c < 3) ? -1 - c : NXM + (c - 3);
    }
    double *fs[4] = {f0, f1, f2, f3};
    for (int q = 0; q < nf; q++) {
        double *f = fs[q];
        double val;
        bool write = true;
        switch (kind) {
        case 0: {
            if (i == NXM + 3 || j == NYM + 3) { write = false; val = 0.0; break; }
            int ii = wrapc(i, NXM);
            int jj = j < 0 ? 0 : (j >= NYM ? NYM - 1 : j);
            val = AT(f, ii, jj);
        } break;
        case 1: {
            if (i == NXM + 3) { write = false; val = 0.0; break; }
            int ii = wrapc(i, NXM);
            if (j == -1) val = AT(f, ii, 0);
            else if (j == NYM) val = AT(f, ii, NYM - 1);
            else if (j < -1 || j > NYM) { if (i >= 0 && i < NXM) { write = false; val = 0.0; break; } val = AT(f, ii, j); }
            else val = AT(f, ii, j);
        } break;
        case 2: case 6: {
            if (i >= 0 && i < NXM) { write = false; val = 0.0; break; }
            if (i == NXM + 3) { write = false; val = 0.0; break; }
            val = AT(f, wrapc(i, NXM), j);
        } break;
        case 3: {
            if (i >= 1 && i <= NXM) { write = false; val = 0.0; break; }
            val = AT(f, wrapu(i, NXM), j);
        } break;
        case 4: {
            int ii = wrapu(i, NXM);
            if (j == -1) val = -AT(f, ii, 0);
            else if (j == NYM) val = -AT(f, ii, NYM - 1);
            else if (i >= 1 && i <= NXM) { write = false; val = 0.0; }
            else val = AT(f, ii, j);
        } break;
        case 5: default: {
            if (i == NXM + 3) { write = false; val = 0.0; break; }
            if (j == 0 || j == NYM) val = 0.0;
            else if (i >= 0 && i < NXM) { write = false; val = 0.0; }
            else val = AT(f, wrapc(i, NXM), j);
        } break;
        }
        if (write) AT(f, i, j) = val;
    }
}

// Rows 0 and NYM of a v field are interior-indexed rows that the NORMAL BC
// zeroes; the frame kernel above only visits j < 0, j >= NYM and ghost
// columns, so row 0 is cleared here.
__global__ void k_zero_row(pvk_grid g, double *f, int j)
{
    const int P = g.P;
    int i = blockIdx.x * blockDim.x + threadIdx.x - 3;
    if (i <= g.NXM + 3) AT(f, i, j) = 0.0;
}

static void launch_ghost(const pvk_grid *g, double *f0, double *f1, double *f2, double *f3, int nf,
                         int kind, hipStream_t s)
{
    int ncol = (kind == 3 || kind == 4) ? 7 : 6;
    int n = 6 * (g->NXM + 7) + ncol * g->NYM;
    hipLaunchKernelGGL(k_ghost, dim3((n + 255) / 256), dim3(256), 0, s, *g, f0, f1, f2, f3, nf, kind);
}

extern "C" void pvk_ghost_multi(const pvk_grid *g, double *f0, double *f1, double *f2, double *f3,
                                int nf, int kind, hipStream_t s)
{
    launch_ghost(g, f0, f1, f2, f3, nf, kind, s);
}
extern "C" void pvk_bc_ch(const pvk_grid *g, double *f, hipStream_t s) { launch_ghost(g, f, 0, 0, 0, 1, 0, s); }
extern "C" void pvk_halo_cell(const pvk_grid *g, double *f, hipStream_t s) { launch_ghost(g, f, 0, 0, 0, 1, 2, s); }
extern "C" void pvk_halo_u(const pvk_grid *g, double *f, hipStream_t s) { launch_ghost(g, f, 0, 0, 0, 1, 3, s); }
extern "C" void pvk_p_bc(const pvk_grid *g, double *f, hipStream_t s) { launch_ghost(g, f, 0, 0, 0, 1, 1, s); }
extern "C" void pvk_vel_bc(const pvk_grid *g, double *f, int comp, int type, hipStream_t s)
{
    if (comp == 'u') { launch_ghost(g, f, 0, 0, 0, 1, 4, s); return; }
    if (type == 0) {
        hipLaunchKernelGGL(k_zero_row, dim3((g->NXM + 7 + 255) / 256), dim3(256), 0, s, *g, f, 0);
        launch_ghost(g, f, 0, 0, 0, 1, 5, s);
    } else {
        launch_ghost(g, f, 0, 0, 0, 1, 6, s);
    }
}

// Communication.c:146-295 on one rank that is its own x neighbour: thread
// (slot, j) copies one ghost of row j.  Slots 0 .. nlo-1 fill the columns
// below the domain, the rest the columns above it.
__global__ void k_xwrap(pvk_grid g, double *f, int ulike, int pnodes)
{
    const int P = g.P, NXM = g.NXM;
    const int nlo = ulike ? pnodes + 1 : pnodes, nslot = nlo + pnodes;
    const int nrow = g.NYM + 1 + 2 * pnodes;            // j in [-pnodes, NYM + pnodes]
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nslot * nrow) return;
    const int slot = t / nrow, j = t % nrow - pnodes;
    int dst, src;
    if (slot < nlo) {                                  // data[Is-q-1] = data[Ie-q-1]
        dst = ulike ? -slot : -1 - slot;
        src = ulike ? NXM - slot : NXM - 1 - slot;
    } else {                                           // data[Ie+q] = data[Is+q]
        const int q = slot - nlo;
        dst = ulike ? NXM + 1 + q : NXM + q;
        src = ulike ? 1 + q : q;
    }
    AT(f, dst, j) = AT(f, src, j);
}
extern "C" void pvk_xwrap(const pvk_grid *g, double *f, int ulike, int pnodes, hipStream_t s)
{
    const int n = ((ulike ? pnodes + 1 : pnodes) + pnodes) * (g->NYM + 1 + 2 * pnodes);
    hipLaunchKernelGGL(k_xwrap, dim3((n + 255) / 256), dim3(256), 0, s, *g, f, ulike, pnodes);
}

// Temporal start of the momentum solves: mode 0 keeps x in s and adds
// scale * dlt (the same stage's increment of the previous step, scaled by the
// ratio of the time steps) on the unknowns of each component (u: i = 1..NXM,
// j < NYM; v: i < NXM, 1 <= j < NYM), x untouched when scale is 0; mode 1
// stores the new increment x - s.
__global__ void k_mom_warm(pvk_grid g, int mode, double scale, double *u, double *v, double *su, double *sv,
                           double *du, double *dv)
{
    const int P = g.P;
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (j >= g.NYM || i > g.NXM) return;
    if (i >= 1) {
        if (mode == 0) {
            const double x = AT(u, i, j);
            AT(su, i, j) = x;
            if (scale != 0.0) AT(u, i, j) = x + scale * AT(du, i, j);
        } else AT(du, i, j) = AT(u, i, j) - AT(su, i, j);
    }
    if (i < g.NXM && j >= 1) {
        if (mode == 0) {
            const double x = AT(v, i, j);
            AT(sv, i, j) = x;
            if (scale != 0.0) AT(v, i, j) = x + scale * AT(dv, i, j);
        } else AT(dv, i, j) = AT(v, i, j) - AT(sv, i, j);
    }
}
extern "C" void pvk_mom_warm(const pvk_grid *g, int mode, double scale, double *u, double *v, double *su,
                             double *sv, double *du, double *dv, hipStream_t s)
{
    dim3 gr((g->NXM + 1 + BX - 1) / BX, (g->NYM + BY - 1) / BY);
    hipLaunchKernelGGL(k_mom_warm, gr, dim3(BX, BY), 0, s, *g, mode, scale, u, v, su, sv, du, dv);
}

__global__ void k_fill(double *f, size_t n, double val)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) f[t] = val;
}
extern "C" void pvk_fill(double *f, size_t n, double val, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, f, n, val);
}

// ---------------------------------------------------------------------------
// Cahn-Hilliard stencils
// ---------------------------------------------------------------------------
#define CELL_IJ                                                   \
    const int i = blockIdx.x * BX + threadIdx.x;                  \
    const int j = blockIdx.y * BY + threadIdx.y;                  \
    const bool in = (i < g.NXM) && (j < g.NYM);                   \
    const int P = g.P;                                            \
    (void)P;

__global__ void k_laplacian(pvk_grid g, const double *in_f, double *out)
{
    CELL_IJ
    if (in) AT(out, i, j) = lap_at(g, in_f, i, j);
}
extern "C" void pvk_laplacian(const pvk_grid *g, const double *in, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_laplacian, grid2d(g), dim3(BX, BY), 0, s, *g, in, out);
}

__global__ void k_bulk_S(pvk_grid g, const double *C, double *S)
{
    CELL_IJ
    if (in) AT(S, i, j) = bulkS(AT(C, i, j));
}
extern "C" void pvk_bulk_S(const pvk_grid *g, const pvk_phys *ph, const double *C, double *S, hipStream_t s)
{
    (void)ph;
    hipLaunchKernelGGL(k_bulk_S, grid2d(g), dim3(BX, BY), 0, s, *g, C, S);
}

__global__ void k_weno_advect(pvk_grid g, pvk_phys ph, const double *C, const double *u, const double *v, double *rhs)
{
    CELL_IJ
    if (in) AT(rhs, i, j) = weno_rhs_at(g, ph, C, u, v, i, j);
}
extern "C" void pvk_weno_advect(const pvk_grid *g, const pvk_phys *ph, const double *C, const double *u,
                                const double *v, double *rhs, hipStream_t s)
{
    hipLaunchKernelGGL(k_weno_advect, grid2d(g), dim3(BX, BY), 0, s, *g, *ph, C, u, v, rhs);
}

__global__ void k_explicit_diffusion(pvk_grid g, pvk_phys ph, const double *C, double *rhs)
{
    CELL_IJ
    if (in) AT(rhs, i, j) = diffusion_at(g, ph, C, i, j);
}
extern "C" void pvk_explicit_diffusion(const pvk_grid *g, const pvk_phys *ph, const double *C, double *rhs,
                                       hipStream_t s)
{
    hipLaunchKernelGGL(k_explicit_diffusion, grid2d(g), dim3(BX, BY), 0, s, *g, *ph, C, rhs);
}

// b -= (Cn^2/Pe) L^2 C on the interior (VOF_DIFFUSE.c:1047-1057)
__global__ void k_ch_sub_bih(pvk_grid g, pvk_phys ph, const double *C, double *b)
{
    CELL_IJ
    if (in) AT(b, i, j) -= ph.bih * bilap_at(g, C, i, j);
}
extern "C" void pvk_ch_sub_bih(const pvk_grid *g, const pvk_phys *ph, const double *C, double *b, hipStream_t s)
{
    hipLaunchKernelGGL(k_ch_sub_bih, grid2d(g), dim3(BX, BY), 0, s, *g, *ph, C, b);
}

// diffuse_bound_liquid_fraction + update_phase_cache + density/viscosity
// (VOF_DIFFUSE.c:1093-1127, :108-126, :1453-1488); partial of the mass delta.
__global__ void k_clamp_props(pvk_grid g, pvk_phys ph, double *C, double *G, double *rho, double *mu,
                              int do_clamp, double *part, int nb)
{
    CELL_IJ
    double vals[2] = {0.0, 0.0};
    if (in) {
        const double cs = 0.0;
        double old_cl = AT(C, i, j), new_cl = old_cl;
        if (do_clamp) {
            double max_cl = 1.0 - cs;
            if (max_cl < 0.0) max_cl = 0.0;
            if (max_cl > 1.0) max_cl = 1.0;
            if (new_cl < 0.0) new_cl = 0.0;
            if (new_cl > max_cl) new_cl = max_cl;
            double d = new_cl - old_cl;
            double m = (g.dx_u[i] * g.dy_v[j] * g.slab) / 2.0;
            double s = d * m;
            if (i == 0) s += d * ((g.dx_u[g.NXM] * g.dy_v[j] * g.slab) / 2.0);
            if (j == g.NYM - 1) s += d * ((g.dx_u[i] * g.dy_v[g.NYM] * g.slab) / 2.0);
            if (i == 0 && j == g.NYM - 1) s += d * ((g.dx_u[g.NXM] * g.dy_v[g.NYM] * g.slab) / 2.0);
            vals[0] = s;
            vals[1] = d * m;    // the interior cell alone (no periodic image, no ghost row)
            AT(C, i, j) = new_cl;
        }
        double gg = 1.0 - new_cl - cs;
        gg = (gg < 0.0) ? 0.0 : ((gg > 1.0) ? 1.0 : gg);
        AT(G, i, j) = gg;
        AT(rho, i, j) = new_cl + 1.0 * cs + ph.rho_g * gg;
        AT(mu, i, j) = new_cl + 1.0 * cs + ph.mu_g * gg;
    }
    if (do_clamp) block_partials<2>(vals, part, nb);
}
__global__ void k_clamp_finalize(pvk_scal *sc, const double *part, int nb)
{
    double s = 2.0 * final_sum(part, nb, 0, false);
    double si = 2.0 * final_sum(part, nb, 1, false);
    if (threadIdx.x == 0) {
        sc->clamp_delta = s;
        sc->clamp_delta_int = si;
    }
}
extern "C" void pvk_clamp_props(const pvk_grid *g, const pvk_phys *ph, double *C, double *G, double *rho,
                                double *mu, const pvk_red *rd, pvk_scal *sc, hipStream_t s)
{
    hipLaunchKernelGGL(k_clamp_props, grid2d(g), dim3(BX, BY), 0, s, *g, *ph, C, G, rho, mu, 1, rd->part,
                       rd->nblocks);
    hipLaunchKernelGGL(k_clamp_finalize, dim3(1), dim3(256), 0, s, sc, rd->part, rd->nblocks);
    launch_ghost(g, C, G, rho, mu, 4, 0, s);
}
extern "C" void pvk_phase_props(const pvk_grid *g, const pvk_phys *ph, const double *C, double *G, double *rho,
                                double *mu, hipStream_t s)
{
    hipLaunchKernelGGL(k_clamp_props, grid2d(g), dim3(BX, BY), 0, s, *g, *ph, (double *)C, G, rho, mu, 0,
                       (double *)0, 0);
    launch_ghost(g, G, rho, mu, 0, 3, 0, s);
}

// psi_LG (VOF_DIFFUSE.c:907-929) and the potential-form capillary force on
// u and v faces (VOF_DIFFUSE.c:1363-1445).  Thread i handles u-face
// iu = (i == 0 ? NXM : i) (unknowns 1..NXM) and v-face (i, j), j >= 1.
template <class R = double>
__device__ __forceinline__ R psi_at(const pvk_grid &g, const pvk_phys &ph, const double *C, int i, int j)
{
    const int P = g.P;
    R c = R(AT(C, i, j));
    return c * c * c - R(1.5) * c * c + R(0.5) * c - R(ph.Cn) * R(ph.Cn) * lap_at<R>(g, C, i, j);
}

// The band test reads the FP64 fields; psi and the force are formed in R.
template <class R>
__global__ void k_psi_fsigma(pvk_grid g, pvk_phys ph, const double *C, const double *G, double *psi_out,
                             double *fx, double *fy)
{
    CELL_IJ
    if (!in) return;
    const R scale = R(2.0 * 6.0 * sqrt(2.0) / (ph.Cn * ph.We + 1e-30));
    const double cs = 0.0;
    R psi_c = psi_at<R>(g, ph, C, i, j);
    AT(psi_out, i, j) = (double)psi_c;
    {   // u face iu, between cells iu-1 and iu
        int iu = (i == 0) ? g.NXM : i;
        R fval = R(0.0);
        if (j < g.NYM) {
            double cl = 0.5 * (AT(C, iu, j) + AT(C, iu - 1, j));
            double cg = 0.5 * (AT(G, iu, j) + AT(G, iu - 1, j));
            if (!(cl < 0.005 || cl > 0.995 || cg < 0.005 || cg > 0.995 || cs > 0.05)) {
                R psi_e = (iu == g.NXM) ? psi_at<R>(g, ph, C, 0, j) : psi_c;
                R psi_w = psi_at<R>(g, ph, C, iu - 1, j);
                R psi_f = R(0.5) * (psi_e + psi_w);
                R gradC = (R(AT(C, iu, j)) - R(AT(C, iu - 1, j))) * R(g.idx_c[iu - 1]);
                fval = scale * psi_f * gradC;
            }
        }
        AT(fx, iu, j) = (double)fval;
    }
    if (j >= 1) {   // v face (i, j), between cells j-1 and j
        R fval = R(0.0);
        double cl = 0.5 * (AT(C, i, j) + AT(C, i, j - 1));
        double cg = 0.5 * (AT(G, i, j) + AT(G, i, j - 1));
        if (!(cl < 0.005 || cl > 0.995 || cg < 0.005 || cg > 0.995 || cs > 0.05)) {
            R psi_f = R(0.5) * (psi_c + psi_at<R>(g, ph, C, i, j - 1));
            R gradC = (R(AT(C, i, j)) - R(AT(C, i, j - 1))) * R(g.idy_c[j - 1]);
            fval = scale * psi_f * gradC;
        }
        AT(fy, i, j) = (double)fval;
    }
}
extern "C" void pvk_psi_fsigma(const pvk_grid *g, const pvk_phys *ph, const double *C, const double *G,
                               double *lap, double *psi, double *fx, double *fy, hipStream_t s)
{
    (void)lap;
    if (ph->precision)
        hipLaunchKernelGGL(k_psi_fsigma<float>, grid2d(g), dim3(BX, BY), 0, s, *g, *ph, C, G, psi, fx, fy);
    else
        hipLaunchKernelGGL(k_psi_fsigma<double>, grid2d(g), dim3(BX, BY), 0, s, *g, *ph, C, G, psi, fx, fy);
}

// ---------------------------------------------------------------------------
// momentum
// ---------------------------------------------------------------------------

// Velocity_cell_center (Velocity.c:278-357) on the G range.
__global__ void k_cell_center(pvk_grid g, const double *u, const double *v, double *ub, double *vb)
{
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int P = g.P;
    if (i <= g.NXM && j <= g.NYM) {
        AT(ub, i, j) = 0.5 * (AT(u, i, j) + AT(u, i + 1, j));
        AT(vb, i, j) = 0.5 * (AT(v, i, j) + AT(v, i, j + 1));
    }
}
extern "C" void pvk_cell_center(const pvk_grid *g, const double *u, const double *v, double *ub, double *vb,
                                hipStream_t s)
{
    dim3 gr((g->NXM + 1 + BX - 1) / BX, (g->NYM + 1 + BY - 1) / BY);
    hipLaunchKernelGGL(k_cell_center, gr, dim3(BX, BY), 0, s, *g, u, v, ub, vb);
}

// Explicit, implicit-diagonal and explicit-viscous terms for u at face
// (iu, j) (Velocity.c:366-762, VOF + VOF_DIFFUSE branches, planar).
template <class R = double>
__device__ __forceinline__ void u_terms(const pvk_grid &g, const pvk_phys &ph, const double *u, const double *v,
                                        const double *mu, int i, int j, R &ex, R &im, R &ve)
{
    const int P = g.P;
    const R iRe = R(ph.iRe);
    R muE = R(AT(mu, i, j)), muW = R(AT(mu, i - 1, j));
    R mu_yp = R(2.0) * (R(0.5) * (R(AT(mu, i - 1, j + 1)) + R(AT(mu, i, j + 1)))) *
                   (R(0.5) * (R(AT(mu, i - 1, j)) + R(AT(mu, i, j)))) /
                   ((R(0.5) * (R(AT(mu, i - 1, j + 1)) + R(AT(mu, i, j + 1)))) +
                    (R(0.5) * (R(AT(mu, i - 1, j)) + R(AT(mu, i, j)))) + R(1e-12));
    R mu_ym = R(2.0) * (R(0.5) * (R(AT(mu, i - 1, j)) + R(AT(mu, i, j)))) *
                   (R(0.5) * (R(AT(mu, i - 1, j - 1)) + R(AT(mu, i, j - 1)))) /
                   ((R(0.5) * (R(AT(mu, i - 1, j)) + R(AT(mu, i, j)))) +
                    (R(0.5) * (R(AT(mu, i - 1, j - 1)) + R(AT(mu, i, j - 1)))) + R(1e-12));
    R uc = R(AT(u, i, j));
    R dudxE = (R(AT(u, i + 1, j)) - uc) * R(g.idx_u[i]);
    R dudxW = (uc - R(AT(u, i - 1, j))) * R(g.idx_u[i - 1]);
    R dudyN = (R(AT(u, i, j + 1)) - uc) * R(g.idy_c[j]);
    R dudyS = (uc - R(AT(u, i, j - 1))) * R(g.idy_c[j - 1]);
    R d2udy2 = iRe * (mu_yp * dudyN - mu_ym * dudyS) * R(g.idy_v[j]);
    R dvdxN = (R(AT(v, i, j + 1)) - R(AT(v, i - 1, j + 1))) * R(g.idx_c[i - 1]);
    R dvdxS = (R(AT(v, i, j)) - R(AT(v, i - 1, j))) * R(g.idx_c[i - 1]);
    R d2udx2 = iRe * ((muE * dudxE - muW * dudxW) * R(g.idx_c[i - 1]));
    R ddydvdx = iRe * (mu_yp * dvdxN - mu_ym * dvdxS) * R(g.idy_v[j]);
    R vCN = R(0.5) * (R(AT(v, i, j + 1)) + R(AT(v, i - 1, j + 1)));
    R vCS = R(0.5) * (R(AT(v, i, j)) + R(AT(v, i - 1, j)));
    R ubE = R(0.5) * (uc + R(AT(u, i + 1, j)));            // u_bc[i]
    R ubW = R(0.5) * (R(AT(u, i - 1, j)) + uc);            // u_bc[i-1]
    R uuE = ubE * ubE, uuW = ubW * ubW;
    R uvN = R(0.5) * (uc + R(AT(u, i, j + 1))) * vCN;
    R uvS = R(0.5) * (uc + R(AT(u, i, j - 1))) * vCS;
    R duudx = (uuE - uuW) * R(g.idx_c[i - 1]);
    R duvdy = (uvN - uvS) * R(g.idy_v[j]);
    ex = -(duudx + duvdy + R(0.0));
    ve = d2udx2 + ddydvdx + R(0.0);
    im = d2udx2 + d2udy2 + R(0.0) + R(0.0);
}

// Same for v at face (i, j) (Velocity.c:771-1144).
template <class R = double>
__device__ __forceinline__ void v_terms(const pvk_grid &g, const pvk_phys &ph, const double *u, const double *v,
                                        const double *mu, int i, int j, R &ex, R &im, R &ve)
{
    const int P = g.P;
    const R iRe = R(ph.iRe);
    R mu_xp = R(2.0) * (R(0.5) * (R(AT(mu, i + 1, j - 1)) + R(AT(mu, i + 1, j)))) *
                   (R(0.5) * (R(AT(mu, i, j - 1)) + R(AT(mu, i, j)))) /
                   ((R(0.5) * (R(AT(mu, i + 1, j - 1)) + R(AT(mu, i + 1, j)))) +
                    (R(0.5) * (R(AT(mu, i, j - 1)) + R(AT(mu, i, j)))) + R(1e-12));
    R mu_xm = R(2.0) * (R(0.5) * (R(AT(mu, i, j - 1)) + R(AT(mu, i, j)))) *
                   (R(0.5) * (R(AT(mu, i - 1, j - 1)) + R(AT(mu, i - 1, j)))) /
                   ((R(0.5) * (R(AT(mu, i, j - 1)) + R(AT(mu, i, j)))) +
                    (R(0.5) * (R(AT(mu, i - 1, j - 1)) + R(AT(mu, i - 1, j)))) + R(1e-12));
    R mu_yp = R(AT(mu, i, j)), mu_ym = R(AT(mu, i, j - 1));
    R vc = R(AT(v, i, j));
    R dvdxE = (R(AT(v, i + 1, j)) - vc) * R(g.idx_c[i]);
    R dvdxW = (vc - R(AT(v, i - 1, j))) * R(g.idx_c[i - 1]);
    R dvdyN = (R(AT(v, i, j + 1)) - vc) * R(g.idy_v[j]);
    R dvdyS = (vc - R(AT(v, i, j - 1))) * R(g.idy_v[j - 1]);
    R d2vdx2 = iRe * ((mu_xp * dvdxE - mu_xm * dvdxW) * R(g.idx_u[i]));
    R d2vdy2 = iRe * (mu_yp * dvdyN - mu_ym * dvdyS) * R(g.idy_c[j - 1]);
    R dudyE = (R(AT(u, i + 1, j)) - R(AT(u, i + 1, j - 1))) * R(g.idy_c[j - 1]);
    R dudyW = (R(AT(u, i, j)) - R(AT(u, i, j - 1))) * R(g.idy_c[j - 1]);
    R ddxdudy = iRe * ((mu_xp * dudyE - mu_xm * dudyW) * R(g.idx_u[i]));
    R uCE = R(0.5) * (R(AT(u, i + 1, j)) + R(AT(u, i + 1, j - 1)));
    R uCW = R(0.5) * (R(AT(u, i, j)) + R(AT(u, i, j - 1)));
    R vuE = R(0.5) * (vc + R(AT(v, i + 1, j))) * uCE;
    R vuW = R(0.5) * (vc + R(AT(v, i - 1, j))) * uCW;
    R vbN = R(0.5) * (vc + R(AT(v, i, j + 1)));            // v_bc[j]
    R vbS = R(0.5) * (R(AT(v, i, j - 1)) + vc);            // v_bc[j-1]
    R vvN = vbN * vbN, vvS = vbS * vbS;
    R dvudx = (vuE - vuW) * R(g.idx_u[i]);
    R dvvdy = (vvN - vvS) * R(g.idy_c[j - 1]);
    ex = -(dvudx + dvvdy + R(0.0));
    ve = ddxdudy + d2vdy2 + R(0.0);
    im = d2vdx2 + d2vdy2 + R(0.0);
}

// Fused momentum right-hand side per face (Velocity.c:1494-1947) plus the
// lagged capillary force (VOF_SurfaceTension.c:1983-2041) and gravity
// (:2075-2154).  mode bit 1: store ex/im/ve; bit 2: assemble rhs and shift
// the explicit history; bit 4: add f_sigma_old and gravity.
template <class R>
__global__ void k_mom_terms(pvk_grid g, pvk_phys ph, int stage, const pvk_scal *sc, int comp,
                            const double *u, const double *v, const double *mu, const double *rho,
                            const double *p, double *eold, double *vold, const double *fold, double *rhs,
                            double *exo, double *imo, double *veo, int mode)
{
    CELL_IJ
    if (!in) return;
    int ii = i;
    if (comp == 'u') ii = (i == 0) ? g.NXM : i;
    else if (j == 0) return;
    R exr, imr, ver;
    if (comp == 'u') u_terms<R>(g, ph, u, v, mu, ii, j, exr, imr, ver);
    else v_terms<R>(g, ph, u, v, mu, ii, j, exr, imr, ver);
    const double ex = (double)exr, im = (double)imr, ve = (double)ver;
    if (mode & 1) { AT(exo, ii, j) = ex; AT(imo, ii, j) = im; AT(veo, ii, j) = ve; }
    if (mode & 2) {
        const double a_dt = 1.0 / (sc->dt * c_BET[stage]);
        const double G = c_GAMB[stage], Z = c_ZETB[stage];
        double dp, rho_face, vel;
        if (comp == 'u') {
            dp = (AT(p, ii, j) - AT(p, ii - 1, j)) * g.idx_c[ii - 1];
            rho_face = 0.5 * (AT(rho, ii, j) + AT(rho, ii - 1, j));
            vel = AT(u, ii, j);
        } else {
            dp = (AT(p, ii, j) - AT(p, ii, j - 1)) * g.idy_c[j - 1];
            rho_face = 0.5 * (AT(rho, ii, j) + AT(rho, ii, j - 1));
            vel = AT(v, ii, j);
        }
        double conv = G * ex + Z * AT(eold, ii, j);
        conv *= rho_face;
        double r = rho_face * a_dt * vel + conv + im + (G * ve + Z * AT(vold, ii, j)) - 2.0 * dp;
        if (comp == 'u') r += 2.0 * (-ph.dp_dx);
        if (mode & 4) {
            r += AT(fold, ii, j);
            r += 2.0 * rho_face * ph.Ri * (comp == 'u' ? ph.gx : ph.gy);
        }
        AT(rhs, ii, j) = r;
        AT(eold, ii, j) = ex;
        AT(vold, ii, j) = ve;
    }
}
extern "C" void pvk_mom_terms(const pvk_grid *g, const pvk_phys *ph, int stage, const pvk_scal *sc, int comp,
                              const double *u, const double *v, const double *mu, const double *rho,
                              const double *p, double *eold, double *vold, const double *fold, double *rhs,
                              double *ex_out, double *im_out, double *ve_out, int write_terms, int add_force,
                              hipStream_t s)
{
    int mode = write_terms | (add_force ? 4 : 0);
    if (ph->precision)
        hipLaunchKernelGGL(k_mom_terms<float>, grid2d(g), dim3(BX, BY), 0, s, *g, *ph, stage, sc, comp, u, v, mu, rho,
                           p, eold, vold, fold, rhs, ex_out, im_out, ve_out, mode);
    else
        hipLaunchKernelGGL(k_mom_terms<double>, grid2d(g), dim3(BX, BY), 0, s, *g, *ph, stage, sc, comp, u, v, mu, rho,
                           p, eold, vold, fold, rhs, ex_out, im_out, ve_out, mode);
}

// Force terms alone (VOF_apply_f_sigma_old / Velocity_add_gravity_2_RHS).
__global__ void k_mom_force(pvk_grid g, pvk_phys ph, int comp, const double *rho, const double *fold,
                            double *rhs, int which)
{
    CELL_IJ
    if (!in) return;
    int ii = i;
    if (comp == 'u') ii = (i == 0) ? g.NXM : i;
    else if (j == 0) return;
    double rho_face = (comp == 'u') ? 0.5 * (AT(rho, ii, j) + AT(rho, ii - 1, j))
                                    : 0.5 * (AT(rho, ii, j) + AT(rho, ii, j - 1));
    if (which & 1) AT(rhs, ii, j) += AT(fold, ii, j);
    if (which & 2) AT(rhs, ii, j) += 2.0 * rho_face * ph.Ri * (comp == 'u' ? ph.gx : ph.gy);
}
extern "C" void pvk_mom_force(const pvk_grid *g, const pvk_phys *ph, int comp, const double *rho,
                              const double *fold, double *rhs, int which, hipStream_t s)
{
    hipLaunchKernelGGL(k_mom_force, grid2d(g), dim3(BX, BY), 0, s, *g, *ph, comp, rho, fold, rhs, which);
}

#define MOM_IJ                                        \
    CELL_IJ                                           \
    int ii = i;                                       \
    bool act = in;                                    \
    if (comp == 'u') ii = (i == 0) ? g.NXM : i;       \
    else if (j == 0) act = false;

__global__ void k_mom_matvec(pvk_grid g, pvk_phys ph, int stage, const pvk_scal *sc, int comp, const double *x,
                             const double *rho, const double *mu, double *Ax)
{
    MOM_IJ
    if (act) AT(Ax, ii, j) = mom_op(g, ph, c_BET[stage], sc->dt, rho, mu, x, comp, ii, j);
}
extern "C" void pvk_mom_matvec(const pvk_grid *g, const pvk_phys *ph, int stage, const pvk_scal *sc, int comp,
                               const double *x, const double *rho, const double *mu, double *Ax, hipStream_t s)
{
    hipLaunchKernelGGL(k_mom_matvec, grid2d(g), dim3(BX, BY), 0, s, *g, *ph, stage, sc, comp, x, rho, mu, Ax);
}

// ---------------------------------------------------------------------------
// pressure
// ---------------------------------------------------------------------------
// Pressure_set_RHS (Pressure.c:100-196, GRID_UNIFORM) and the harmonic-mean
// Jacobi preconditioner (psolve_cg.c:63-101), cells i < NXM, j < NYM.
__global__ void k_p_rhs(pvk_grid g, int stage, const pvk_scal *sc, const double *u, const double *v,
                        const double *rho, double *rhs, double *M)
{
    CELL_IJ
    if (!in) return;
    const double a_dt = 1.0 / (sc->dt * 2.0 * c_BET[stage]);
    const double ihdt = g.idx_u[1] * a_dt;
    double dudx = (AT(u, i + 1, j) - AT(u, i, j)) * ihdt;
    double dvdy = (AT(v, i, j + 1) - AT(v, i, j)) * ihdt;
    AT(rhs, i, j) = dudx + dvdy;
    if (M) {
        const double idx2 = g.idx_c[1] * g.idx_c[1], idy2 = g.idy_c[1] * g.idy_c[1];
        double r0 = AT(rho, i, j);
        double bxp = 2.0 / (r0 + AT(rho, i + 1, j)), bxm = 2.0 / (r0 + AT(rho, i - 1, j));
        double byp = 2.0 / (r0 + AT(rho, i, j + 1)), bym = 2.0 / (r0 + AT(rho, i, j - 1));
        double aii = bxm * idx2 + bxp * idx2 + bym * idy2 + byp * idy2 + 0.0 + 0.0;
        AT(M, i, j) = 1.0 / (1.0 * aii + 1e-12);
    }
}
extern "C" void pvk_p_rhs2(const pvk_grid *g, int stage, const pvk_scal *sc, const double *u, const double *v,
                           const double *rho, double *rhs, double *M, hipStream_t s)
{
    hipLaunchKernelGGL(k_p_rhs, grid2d(g), dim3(BX, BY), 0, s, *g, stage, sc, u, v, rho, rhs, M);
}

// Pressure_operator_variableCoeff (psolve_cg.c:302-344, TwodOps.c:116-142)
__device__ __forceinline__ double p_op(const pvk_grid &g, const double *rho, const double *phi, int i, int j)
{
    const int P = g.P;
    const double idx2 = g.idx_c[1] * g.idx_c[1], idy2 = g.idy_c[1] * g.idy_c[1];
    double r0 = AT(rho, i, j), c0 = AT(phi, i, j);
    double cxp = (2.0 / (r0 + AT(rho, i + 1, j))) * idx2;
    double cxm = (2.0 / (r0 + AT(rho, i - 1, j))) * idx2;
    double cyp = (2.0 / (r0 + AT(rho, i, j + 1))) * idy2;
    double cym = (2.0 / (r0 + AT(rho, i, j - 1))) * idy2;
    double op = cxp * (AT(phi, i + 1, j) - c0) - cxm * (c0 - AT(phi, i - 1, j)) +
                cyp * (AT(phi, i, j + 1) - c0) - cym * (c0 - AT(phi, i, j - 1)) + 0.0 * 0.0 - 0.0 * 0.0;
    return 1.0 * op;
}

__global__ void k_p_operator(pvk_grid g, const double *rho, const double *x, double *Ax)
{
    CELL_IJ
    if (in) AT(Ax, i, j) = p_op(g, rho, x, i, j);
}
extern "C" void pvk_p_operator(const pvk_grid *g, const double *rho, const double *x, double *Ax, hipStream_t s)
{
    hipLaunchKernelGGL(k_p_operator, grid2d(g), dim3(BX, BY), 0, s, *g, rho, x, Ax);
}

// PCG start with x0 = 0 (psolve_cg.c:166-194): r = rhs - A 0 = rhs, z = M r,
// d = z, partial <r, z>.  use_M = 0 leaves z = r (the FFT path overwrites d).
__global__ void k_p_cg_init(pvk_grid g, const double *rhs, const double *M, double *phi, double *r, double *d,
                            int use_M, double shift_mean_slot, const pvk_scal *sc, double *part, int nb)
{
    CELL_IJ
    double vals[2] = {0.0, 0.0};
    if (in) {
        double rr = AT(rhs, i, j);
        if (shift_mean_slot != 0.0) rr = rr - sc->p_mean;
        double z = use_M ? rr * AT(M, i, j) : rr;
        AT(phi, i, j) = 0.0;
        AT(r, i, j) = rr;
        AT(d, i, j) = z;
        vals[0] = rr * z;
        vals[1] = rr * rr;
    }
    block_partials<2>(vals, part, nb);
}
extern "C" void pvk_p_cg_init2(const pvk_grid *g, const double *rhs, const double *M, double *phi, double *r,
                               double *d, int use_M, int shift_mean, const pvk_scal *sc, const pvk_red *rd,
                               hipStream_t s)
{
    hipLaunchKernelGGL(k_p_cg_init, grid2d(g), dim3(BX, BY), 0, s, *g, rhs, M, phi, r, d, use_M,
                       shift_mean ? 1.0 : 0.0, sc, rd->part, rd->nblocks);
}

// Ad = A d; partial <d, Ad> (psolve_cg.c:203-208)
__global__ void k_p_cg_apply(pvk_grid g, const pvk_scal *sc, const double *rho, const double *d, double *Ad,
                             double *part, int nb)
{
    if (sc->p_done) return;
    CELL_IJ
    double vals[1] = {0.0};
    if (in) {
        double ad = p_op(g, rho, d, i, j);
        AT(Ad, i, j) = ad;
        vals[0] = AT(d, i, j) * ad;
    }
    block_partials<1>(vals, part, nb);
}
extern "C" void pvk_p_cg_apply(const pvk_grid *g, pvk_scal *sc, const double *rho, const double *d, double *Ad,
                               const pvk_red *rd, hipStream_t s)
{
    hipLaunchKernelGGL(k_p_cg_apply, grid2d(g), dim3(BX, BY), 0, s, *g, sc, rho, d, Ad, rd->part, rd->nblocks);
}

// phi += alpha d; r -= alpha Ad; z = M r; partials <r, z>, <r, r> (:211-230)
__global__ void k_p_cg_update(pvk_grid g, const pvk_scal *sc, double *phi, const double *d, double *r,
                              const double *Ad, const double *M, int use_M, double *part, int nb)
{
    if (sc->p_done) return;
    CELL_IJ
    double vals[2] = {0.0, 0.0};
    if (in) {
        const double alpha = sc->p_alpha;
        AT(phi, i, j) += alpha * AT(d, i, j);
        double rr = AT(r, i, j) - alpha * AT(Ad, i, j);
        AT(r, i, j) = rr;
        double z = use_M ? rr * AT(M, i, j) : rr;
        vals[0] = rr * z;
        vals[1] = rr * rr;
    }
    block_partials<2>(vals, part, nb);
}
extern "C" void pvk_p_cg_update(const pvk_grid *g, pvk_scal *sc, double *phi, const double *d, double *r,
                                const double *Ad, const double *M, int use_M, const pvk_red *rd, hipStream_t s)
{
    hipLaunchKernelGGL(k_p_cg_update, grid2d(g), dim3(BX, BY), 0, s, *g, sc, phi, d, r, Ad, M, use_M, rd->part,
                       rd->nblocks);
}

// d = z + beta d, z = M r (Jacobi) or z given (FFT preconditioner) (:238-246)
__global__ void k_p_cg_dupdate(pvk_grid g, const pvk_scal *sc, const double *zr, const double *M, int use_M,
                               double *d)
{
    if (sc->p_done) return;
    CELL_IJ
    if (in) {
        double z = use_M ? AT(zr, i, j) * AT(M, i, j) : AT(zr, i, j);
        AT(d, i, j) = z + sc->p_beta * AT(d, i, j);
    }
}
extern "C" void pvk_p_cg_dupdate(const pvk_grid *g, pvk_scal *sc, const double *z, const double *M, int use_M,
                                 double *d, hipStream_t s)
{
    hipLaunchKernelGGL(k_p_cg_dupdate, grid2d(g), dim3(BX, BY), 0, s, *g, sc, z, M, use_M, d);
}

// Scalar logic.  mode 0 = Jacobi-PCG (sqrt(r.z) absolute stop, psolve_cg.c),
// mode 1 = FFT-PCG (||r|| / ||b|| stop).  which: 0 start, 1 alpha, 2 stop/beta,
// 3 gamma after a new preconditioned residual (FFT path).
__global__ void k_p_scalar(pvk_scal *sc, int which, int mode, int maxit, double tol, const double *part, int nb)
{
    if (which == 0) {
        double rz = final_sum(part, nb, 0, false);
        double rr = final_sum(part, nb, 1, false);
        if (threadIdx.x == 0) {
            sc->p_gamma = rz;
            sc->p_iter = 0;
            sc->p_done = 0;
            if (mode == 0) {
                sc->p_res = sqrt(rz);
                if (sqrt(rz) < tol) sc->p_done = 4;   // early return: no gauge (:182-185)
            } else {
                double bn = sqrt(rr);
                sc->p_bnorm = (bn == 0.0) ? 1.0 : bn;
                sc->p_res = sqrt(rr) / sc->p_bnorm;
                if (sc->p_res <= tol) sc->p_done = 1;
            }
            if (maxit <= 0 && !sc->p_done) sc->p_done = 3;
        }
        return;
    }
    if (sc->p_done) return;
    if (which == 1) {
        double dAd = final_sum(part, nb, 0, false);
        if (threadIdx.x == 0) {
            sc->p_alpha = (mode == 0) ? sc->p_gamma / (dAd + 1e-30) : sc->p_gamma / dAd;
            sc->p_iter += 1;
        }
        return;
    }
    if (which == 2) {
        double rz = final_sum(part, nb, 0, false);
        double rr = final_sum(part, nb, 1, false);
        if (threadIdx.x == 0) {
            if (mode == 0) {
                sc->p_res = sqrt(rz);
                if (sqrt(rz) < tol) { sc->p_done = 1; return; }
                sc->p_beta = rz / (sc->p_gamma + 1e-30);
                sc->p_gamma = rz;
            } else {
                sc->p_res = sqrt(rr) / sc->p_bnorm;
                if (sc->p_res <= tol) { sc->p_done = 1; return; }
            }
            if (sc->p_iter >= maxit) sc->p_done = 3;
        }
        return;
    }
    // which == 3: gamma_new = <r, z> with z from the FFT preconditioner;
    // which == 4: first gamma = <r, z>, beta = 0 (so d = z)
    double rz = final_sum(part, nb, 0, false);
    if (threadIdx.x == 0) {
        sc->p_beta = (which == 4) ? 0.0 : rz / sc->p_gamma;
        sc->p_gamma = rz;
    }
}
extern "C" void pvk_p_scalar2(const pvk_grid *g, pvk_scal *sc, int which, int mode, int maxit, double tol,
                              const pvk_red *rd, int nb, hipStream_t s)
{
    (void)g;
    hipLaunchKernelGGL(k_p_scalar, dim3(1), dim3(256), 0, s, sc, which, mode, maxit, tol, rd->part,
                       nb > 0 ? nb : rd->nblocks);
}

// partial <a, b> over cells (i < NXM, j < NYM)
__global__ void k_dot(pvk_grid g, const double *a, const double *b, double *part, int nb)
{
    CELL_IJ
    double vals[1] = {0.0};
    if (in) vals[0] = AT(a, i, j) * AT(b, i, j);
    block_partials<1>(vals, part, nb);
}
extern "C" void pvk_dot(const pvk_grid *g, const double *a, const double *b, const pvk_red *rd, hipStream_t s)
{
    hipLaunchKernelGGL(k_dot, grid2d(g), dim3(BX, BY), 0, s, *g, a, b, rd->part, rd->nblocks);
}

// partial sums of x for the mean (psolve_cg.c:255-279)
__global__ void k_sum(pvk_grid g, const double *x, double *part, int nb)
{
    CELL_IJ
    double vals[1] = {0.0};
    if (in) vals[0] = AT(x, i, j);
    block_partials<1>(vals, part, nb);
}
extern "C" void pvk_p_mean(const pvk_grid *g, const double *x, const pvk_red *rd, hipStream_t s)
{
    hipLaunchKernelGGL(k_sum, grid2d(g), dim3(BX, BY), 0, s, *g, x, rd->part, rd->nblocks);
}
__global__ void k_mean_finalize(pvk_scal *sc, double cnt, const double *part, int nb)
{
    double s = final_sum(part, nb, 0, false);
    if (threadIdx.x == 0) sc->p_mean = s / cnt;
}
extern "C" void pvk_p_mean_finalize(const pvk_grid *g, pvk_scal *sc, const pvk_red *rd, hipStream_t s)
{
    double cnt = (double)((long)g->NXM * (long)g->NYM);
    hipLaunchKernelGGL(k_mean_finalize, dim3(1), dim3(256), 0, s, sc, cnt, rd->part, rd->nblocks);
}
__global__ void k_sub_mean(pvk_grid g, const pvk_scal *sc, double *x)
{
    CELL_IJ
    if (in) AT(x, i, j) -= sc->p_mean;
}
extern "C" void pvk_p_sub_mean(const pvk_grid *g, const pvk_scal *sc, double *x, int ghosts, hipStream_t s)
{
    hipLaunchKernelGGL(k_sub_mean, grid2d(g), dim3(BX, BY), 0, s, *g, sc, x);
    if (ghosts) launch_ghost(g, x, 0, 0, 0, 1, 2, s);
}

// Pressure_project_velocity (Pressure.c:218-464): p += phi on cells, then
// u -= dphi/dx * 2 beta dt / rho_face, v likewise; phi ghosts must be valid.
__global__ void k_project(pvk_grid g, int stage, const pvk_scal *sc, const double *rho, const double *phi,
                          double *p, double *u, double *v)
{
    CELL_IJ
    if (!in) return;
    const double mdt = 2.0 * c_BET[stage] * sc->dt;
    AT(p, i, j) += AT(phi, i, j);
    {
        int iu = (i == 0) ? g.NXM : i;
        double inv_rho_face = 2.0 / (AT(rho, iu, j) + AT(rho, iu - 1, j));
        double idxdt = g.idx_c[iu - 1] * mdt;
        AT(u, iu, j) -= (AT(phi, iu, j) - AT(phi, iu - 1, j)) * (idxdt * inv_rho_face);
    }
    if (j >= 1) {
        double inv_rho_face = 2.0 / (AT(rho, i, j) + AT(rho, i, j - 1));
        double idydt = g.idy_c[j - 1] * mdt;
        AT(v, i, j) -= (AT(phi, i, j) - AT(phi, i, j - 1)) * (idydt * inv_rho_face);
    }
}
extern "C" void pvk_project(const pvk_grid *g, int stage, const pvk_scal *sc, const double *rho, const double *phi,
                            double *p, double *u, double *v, hipStream_t s)
{
    hipLaunchKernelGGL(k_project, grid2d(g), dim3(BX, BY), 0, s, *g, stage, sc, rho, phi, p, u, v);
}

// max |div u| (Pressure.c:475-597)
__global__ void k_divergence(pvk_grid g, const double *u, const double *v, double *part, int nb)
{
    CELL_IJ
    double vals[1] = {0.0};
    if (in) {
        double dudx = (AT(u, i + 1, j) - AT(u, i, j)) * g.idx_u[i];
        double dvdy = (AT(v, i, j + 1) - AT(v, i, j)) * g.idy_v[j];
        vals[0] = fabs(dudx + dvdy);
    }
    block_max_partials<1>(vals, part, nb);
}
extern "C" void pvk_divergence(const pvk_grid *g, const double *u, const double *v, const pvk_red *rd, hipStream_t s)
{
    hipLaunchKernelGGL(k_divergence, grid2d(g), dim3(BX, BY), 0, s, *g, u, v, rd->part, rd->nblocks);
}
__global__ void k_div_finalize(pvk_scal *sc, const double *part, int nb)
{
    double m = final_sum(part, nb, 0, true);
    if (threadIdx.x == 0) sc->div_max = m;
}
extern "C" void pvk_div_finalize(const pvk_grid *g, pvk_scal *sc, const pvk_red *rd, hipStream_t s)
{
    (void)g;
    hipLaunchKernelGGL(k_div_finalize, dim3(1), dim3(256), 0, s, sc, rd->part, rd->nblocks);
}

// ---------------------------------------------------------------------------
// time step (Dtime.c:24-249, FULLY_IMPLICIT, SURFACE_TENSION)
// ---------------------------------------------------------------------------
__global__ void k_cfl(pvk_grid g, const double *u, const double *v, double *part, int nb)
{
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int P = g.P;
    double vals[1] = {0.0};
    if (i <= g.NXM && j <= g.NYM) {
        double ub = 0.5 * (AT(u, i, j) + AT(u, i + 1, j));
        double vb = 0.5 * (AT(v, i, j) + AT(v, i, j + 1));
        double u_cfl = fabs(ub), v_cell = fabs(vb), v_part = fabs(vb - 0.0);
        double v_cfl = (v_cell < v_part) ? v_part : v_cell;
        vals[0] = u_cfl * g.idx_u[i] + v_cfl * g.idy_v[j] + 0.0;
    }
    block_max_partials<1>(vals, part, nb);
}
extern "C" void pvk_cfl(const pvk_grid *g, const double *u, const double *v, const pvk_red *rd, hipStream_t s)
{
    dim3 gr((g->NXM + 1 + BX - 1) / BX, (g->NYM + 1 + BY - 1) / BY);
    hipLaunchKernelGGL(k_cfl, gr, dim3(BX, BY), 0, s, *g, u, v, rd->part, (int)(gr.x * gr.y));
}
__global__ void k_cfl_finalize(pvk_scal *sc, pvk_phys ph, double dx_min, const double *part, int nb)
{
    double max_idt = final_sum(part, nb, 0, true);
    if (threadIdx.x == 0) {
        sc->max_idt = max_idt;
        double dt_cfl = ph.cfl / max_idt;
        double dt_sigma = ph.cfl * sqrt(((ph.rho1 + ph.rho2) * pow(dx_min, 3) * ph.We) / (4.0 * 3.14159265358979323846));
        dt_cfl = (dt_cfl > dt_sigma) ? dt_sigma : dt_cfl;
        double dt;
        if (ph.constant_dt == 1) dt = ph.default_dt;
        else {
            double dt_old = sc->dt;          // caller's dt becomes dt_old
            dt = dt_cfl;
            if (dt > 1.05 * dt_old) dt = 1.05 * dt_old;
            if (dt < 0.95 * dt_old) dt = 0.95 * dt_old;
            else dt = (dt > ph.max_dt) ? ph.max_dt : dt;
        }
        sc->dt_new = dt;
    }
}
extern "C" void pvk_cfl_finalize2(const pvk_grid *g, const pvk_phys *ph, pvk_scal *sc, double dx_min,
                                  const pvk_red *rd, hipStream_t s)
{
    int nb = ((g->NXM + 1 + BX - 1) / BX) * ((g->NYM + 1 + BY - 1) / BY);
    hipLaunchKernelGGL(k_cfl_finalize, dim3(1), dim3(256), 0, s, sc, *ph, dx_min, rd->part, nb);
}
// dt_old = dt; dt = dt_new (Temporal_int.c:427-429)
__global__ void k_take_dt(pvk_scal *sc)
{
    sc->dt_old = sc->dt;
    sc->dt = sc->dt_new;
}
extern "C" void pvk_take_dt(pvk_scal *sc, hipStream_t s) { hipLaunchKernelGGL(k_take_dt, dim3(1), dim3(1), 0, s, sc); }
// time += dt * BETA2[rk] (Temporal_int.c:445)
__global__ void k_advance_time(pvk_scal *sc, int stage) { sc->time += sc->dt * c_BET2[stage]; }
extern "C" void pvk_advance_time(pvk_scal *sc, int stage, hipStream_t s)
{
    hipLaunchKernelGGL(k_advance_time, dim3(1), dim3(1), 0, s, sc, stage);
}

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__global__ void k_psi_only(pvk_grid g, pvk_phys ph, const double *C, double *psi)
{
    CELL_IJ
    if (in) AT(psi, i, j) = psi_at(g, ph, C, i, j);
}
extern "C" void pvk_psi_only(const pvk_grid *g, const pvk_phys *ph, const double *C, double *psi, hipStream_t s)
{
    hipLaunchKernelGGL(k_psi_only, grid2d(g), dim3(BX, BY), 0, s, *g, *ph, C, psi);
}

// C_L = clamp01(F) (VOF_DIFFUSE.c:531-537); ghosts follow from a CH ghost fill.
__global__ void k_clamp_init(pvk_grid g, const double *F, double *C)
{
    CELL_IJ
    if (in) {
        double f = AT(F, i, j);
        AT(C, i, j) = (f < 0.0) ? 0.0 : ((f > 1.0) ? 1.0 : f);
    }
}
extern "C" void pvk_clamp_init(const pvk_grid *g, const double *F, double *C, hipStream_t s)
{
    hipLaunchKernelGGL(k_clamp_init, grid2d(g), dim3(BX, BY), 0, s, *g, F, C);
}

// VOF_DIFFUSE.c:531-537: the L-range loop, ghosts included
__global__ void k_clamp_slab(pvk_grid g, const double *F, double *C)
{
    const int P = g.P;
    int i = blockIdx.x * BX + threadIdx.x - 3, j = blockIdx.y * BY + threadIdx.y - 3;
    if (i <= g.NXM + 3 && j <= g.NYM + 3) {
        double f = AT(F, i, j);
        AT(C, i, j) = (f < 0.0) ? 0.0 : ((f > 1.0) ? 1.0 : f);
    }
}
extern "C" void pvk_clamp_slab(const pvk_grid *g, const double *F, double *C, hipStream_t s)
{
    dim3 gr((g->NXM + 7 + BX - 1) / BX, (g->NYM + 7 + BY - 1) / BY);
    hipLaunchKernelGGL(k_clamp_slab, gr, dim3(BX, BY), 0, s, *g, F, C);
}

// whole-slab copy including ghosts (Array_copy_withghost)
__global__ void k_copy_slab(pvk_grid g, const double *src, double *dst)
{
    const int P = g.P;
    int i = blockIdx.x * BX + threadIdx.x - 3, j = blockIdx.y * BY + threadIdx.y - 3;
    if (i <= g.NXM + 3 && j <= g.NYM + 3) AT(dst, i, j) = AT(src, i, j);
}
extern "C" void pvk_copy(const pvk_grid *g, const double *src, double *dst, hipStream_t s)
{
    dim3 gr((g->NXM + 7 + BX - 1) / BX, (g->NYM + 7 + BY - 1) / BY);
    hipLaunchKernelGGL(k_copy_slab, gr, dim3(BX, BY), 0, s, *g, src, dst);
}

// partial <a, b>, skipped once the pressure solve has stopped
__global__ void k_dot_sc(pvk_grid g, const pvk_scal *sc, const double *a, const double *b, double *part, int nb)
{
    if (sc->p_done) return;
    CELL_IJ
    double vals[1] = {0.0};
    if (in) vals[0] = AT(a, i, j) * AT(b, i, j);
    block_partials<1>(vals, part, nb);
}
extern "C" void pvk_dot_sc(const pvk_grid *g, const pvk_scal *sc, const double *a, const double *b,
                           const pvk_red *rd, hipStream_t s)
{
    hipLaunchKernelGGL(k_dot_sc, grid2d(g), dim3(BX, BY), 0, s, *g, sc, a, b, rd->part, rd->nblocks);
}

