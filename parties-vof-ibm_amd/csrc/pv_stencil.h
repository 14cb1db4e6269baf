// pv_stencil.h -- device-side stencil formulas of the PARTIES two-phase
// substep shared by the kernel files (pv_kernels.hip, pv_cg.hip).  Every
// expression keeps the operand order of the reference routine it cites so
// that, compiled with -ffp-contract=off, results round exactly like the CPU
// oracle.  Field access: AT(a, i, j) with the pitch P of the caller's scope.
#ifndef PV_STENCIL_H
#define PV_STENCIL_H
#include <hip/hip_runtime.h>
#include <math.h>
#include "pv_kernels.h"

#ifndef AT
#define AT(a, i, j) (a)[(ptrdiff_t)(j) * P + (i)]
#endif

// low-storage RK3 tables (definitions.h:198-204)
static __constant__ double c_BET[3]  = {4.0 / 15.0, 1.0 / 15.0, 1.0 / 6.0};
static __constant__ double c_BET2[3] = {8.0 / 15.0, 2.0 / 15.0, 1.0 / 3.0};
static __constant__ double c_GAMB[3] = {2.0, 25.0 / 4.0, 9.0 / 2.0};
static __constant__ double c_ZETB[3] = {0.0, -17.0 / 4.0, -2.5};

// Cell-type image of column i (Communication.c:146-295, period NXM).
__device__ __forceinline__ int wrapc(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }
// u-type image (Communication.c:162-169): unknowns 1..n, column 0 = column n.
__device__ __forceinline__ int wrapu(int i, int n) { return i <= 0 ? i + n : (i > n ? i - n : i); }

// TwodOps_scalar_laplacian, planar (TwodOps.c:159-199)
__device__ __forceinline__ double lap_at(const pvk_grid &g, const double *f, int i, int j)
{
    const int P = g.P;
    double c = AT(f, i, j);
    double dfdxE = (AT(f, i + 1, j) - c) * g.idx_c[i];
    double dfdxW = (c - AT(f, i - 1, j)) * ((i != 0) ? g.idx_c[i - 1] : g.idx_c[i]);
    double dfdyN = (AT(f, i, j + 1) - c) * g.idy_c[j];
    double dfdyS = (c - AT(f, i, j - 1)) * ((j != 0) ? g.idy_c[j - 1] : g.idy_c[j]);
    return (dfdxE - dfdxW) * g.idx_u[i] + (dfdyN - dfdyS) * g.idy_v[j];
}

// Same formula applied to five already-evaluated centre/neighbour values.
__device__ __forceinline__ double lap_vals(const pvk_grid &g, int i, int j, double c, double E,
                                           double W, double N, double S)
{
    double dfdxE = (E - c) * g.idx_c[i];
    double dfdxW = (c - W) * ((i != 0) ? g.idx_c[i - 1] : g.idx_c[i]);
    double dfdyN = (N - c) * g.idy_c[j];
    double dfdyS = (c - S) * ((j != 0) ? g.idy_c[j - 1] : g.idy_c[j]);
    return (dfdxE - dfdxW) * g.idx_u[i] + (dfdyN - dfdyS) * g.idy_v[j];
}

// L(L f) at interior (i,j) where the inner Laplacian obeys the CH boundary
// rules of its own ghost fill: column NXM is column 0, row -1 is row 0 and
// row NYM is row NYM-1 (compute_laplacian then set_boundary_values).
__device__ __forceinline__ double bilap_at(const pvk_grid &g, const double *f, int i, int j)
{
    const int NXM = g.NXM, NYM = g.NYM;
    double L0 = lap_at(g, f, i, j);
    double LE = lap_at(g, f, (i + 1 == NXM) ? 0 : i + 1, j);
    double LW = lap_at(g, f, (i == 0) ? NXM - 1 : i - 1, j);
    double LN = (j + 1 == NYM) ? L0 : lap_at(g, f, i, j + 1);
    double LS = (j == 0) ? L0 : lap_at(g, f, i, j - 1);
    return lap_vals(g, i, j, L0, LE, LW, LN, LS);
}

// S(c) = 0.5 c - 1.5 c^2 + c^3 - c cs (1 - c - cs), cs = 0 (VOF_DIFFUSE.c:865-885)
__device__ __forceinline__ double bulkS(double c)
{
    const double cs = 0.0;
    return 0.5 * c - 1.5 * c * c + c * c * c - c * cs * (1.0 - c - cs);
}

// diffuse_weno5_left (VOF_DIFFUSE.c:431-453); pow(x, 2.0) evaluated as x*x
__device__ __forceinline__ double weno5(double v0, double v1, double v2, double v3, double v4)
{
    const double eps = 1e-6;
    double q0 = (2.0 * v0 - 7.0 * v1 + 11.0 * v2) / 6.0;
    double q1 = (-v1 + 5.0 * v2 + 2.0 * v3) / 6.0;
    double q2 = (2.0 * v2 + 5.0 * v3 - v4) / 6.0;
    double t0 = v0 - 2.0 * v1 + v2, t1 = v0 - 4.0 * v1 + 3.0 * v2;
    double t2 = v1 - 2.0 * v2 + v3, t3 = v1 - v3;
    double t4 = v2 - 2.0 * v3 + v4, t5 = 3.0 * v2 - 4.0 * v3 + v4;
    double b0 = 13.0 / 12.0 * (t0 * t0) + 0.25 * (t1 * t1);
    double b1 = 13.0 / 12.0 * (t2 * t2) + 0.25 * (t3 * t3);
    double b2 = 13.0 / 12.0 * (t4 * t4) + 0.25 * (t5 * t5);
    double a0 = 0.1 / ((eps + b0) * (eps + b0));
    double a1 = 0.6 / ((eps + b1) * (eps + b1));
    double a2 = 0.3 / ((eps + b2) * (eps + b2));
    double asum = a0 + a1 + a2;
    return (a0 * q0 + a1 * q1 + a2 * q2) / asum;
}

// diffuse_face_state_x / _y (VOF_DIFFUSE.c:455-483); face i sits between i-1 and i
__device__ __forceinline__ double face_x(const pvk_grid &g, const double *C, int i, int j, double vel, int wo)
{
    const int P = g.P;
    if (wo == 5) {
        if (vel >= 0.0) return weno5(AT(C, i - 3, j), AT(C, i - 2, j), AT(C, i - 1, j), AT(C, i, j), AT(C, i + 1, j));
        return weno5(AT(C, i + 2, j), AT(C, i + 1, j), AT(C, i, j), AT(C, i - 1, j), AT(C, i - 2, j));
    }
    return (vel >= 0.0) ? AT(C, i - 1, j) : AT(C, i, j);
}
__device__ __forceinline__ double face_y(const pvk_grid &g, const double *C, int i, int j, double vel, int wo)
{
    const int P = g.P;
    if (wo == 5) {
        if (vel >= 0.0) return weno5(AT(C, i, j - 3), AT(C, i, j - 2), AT(C, i, j - 1), AT(C, i, j), AT(C, i, j + 1));
        return weno5(AT(C, i, j + 2), AT(C, i, j + 1), AT(C, i, j), AT(C, i, j - 1), AT(C, i, j - 2));
    }
    return (vel >= 0.0) ? AT(C, i, j - 1) : AT(C, i, j);
}

// -(div(u C)) at interior cell (i,j): VOF_DIFFUSE.c:931-996 + TwodOps.c:201-229
__device__ __forceinline__ double weno_rhs_at(const pvk_grid &g, const pvk_phys &ph, const double *C,
                                              const double *u, const double *v, int i, int j)
{
    const int P = g.P;
    double uw = AT(u, i, j), ue = AT(u, i + 1, j);
    double vs = AT(v, i, j), vn = AT(v, i, j + 1);
    double fxW = uw * face_x(g, C, i, j, uw, ph.weno_order);
    double fxE = ue * face_x(g, C, i + 1, j, ue, ph.weno_order);
    double fyS = vs * face_y(g, C, i, j, vs, ph.weno_order);
    double fyN = vn * face_y(g, C, i, j + 1, vn, ph.weno_order);
    double div_x = (fxE - fxW) / g.dx_c[i];
    double div_y = (fyN - fyS) / g.dy_c[j];
    double div = div_x + div_y;
    return -div;
}

// L(S(C)) / Pe at interior (i,j) (VOF_DIFFUSE.c:998-1016); C ghosts valid.
__device__ __forceinline__ double diffusion_at(const pvk_grid &g, const pvk_phys &ph, const double *C, int i, int j)
{
    const int P = g.P;
    double S0 = bulkS(AT(C, i, j));
    double lapS = lap_vals(g, i, j, S0, bulkS(AT(C, i + 1, j)), bulkS(AT(C, i - 1, j)),
                           bulkS(AT(C, i, j + 1)), bulkS(AT(C, i, j - 1)));
    return lapS / ph.Pe_CH;
}

// measure-weighted CH product contribution of interior cell (i,j) including
// its duplicates in the G range (column NXM, row NYM); the second z slab is
// accounted for by doubling the final sum.
__device__ __forceinline__ double ch_w(const pvk_grid &g, int i, int j, double a, double b)
{
    double m = g.dx_u[i] * g.dy_v[j] * g.slab;
    double s = (m * a) * b;
    bool ix = (i == 0), jy = (j == g.NYM - 1);
    if (ix) s += ((g.dx_u[g.NXM] * g.dy_v[j] * g.slab) * a) * b;
    if (jy) s += ((g.dx_u[i] * g.dy_v[g.NYM] * g.slab) * a) * b;
    if (ix && jy) s += ((g.dx_u[g.NXM] * g.dy_v[g.NYM] * g.slab) * a) * b;
    return s;
}
__device__ __forceinline__ double harm(double a, double b) { return 2.0 * (a) * (b) / ((a) + (b) + 1e-12); }
__device__ __forceinline__ double avg2(double a, double b) { return 0.5 * ((a) + (b)); }

// Jacobi diagonal (lsolver/msolve_cg.c:103-249) and operator (:255-493).
__device__ __forceinline__ double mom_diag(const pvk_grid &g, const pvk_phys &ph, double alpha_k, double dt,
                                           const double *rho, const double *mu, int comp, int i, int j)
{
    const int P = g.P;
    const double iRe = ph.iRe;
    double rho_face, d_time, dx, dy;
    if (comp == 'u') {
        rho_face = 0.5 * (AT(rho, i, j) + AT(rho, i - 1, j));
        d_time = rho_face / (alpha_k * dt);
        double mu_xp = AT(mu, i, j), mu_xm = AT(mu, i - 1, j);
        double mu_yp = harm(avg2(AT(mu, i - 1, j + 1), AT(mu, i, j + 1)), avg2(AT(mu, i - 1, j), AT(mu, i, j)));
        double mu_ym = harm(avg2(AT(mu, i - 1, j), AT(mu, i, j)), avg2(AT(mu, i - 1, j - 1), AT(mu, i, j - 1)));
        dx = iRe * ((mu_xp * g.idx_u[i] + mu_xm * g.idx_u[i - 1]) * g.idx_c[i - 1]);
        dy = iRe * (mu_yp * g.idy_c[j] + mu_ym * g.idy_c[j - 1]) * g.idy_v[j];
    } else {
        rho_face = 0.5 * (AT(rho, i, j) + AT(rho, i, j - 1));
        d_time = rho_face / (alpha_k * dt);
        double mu_yp = AT(mu, i, j), mu_ym = AT(mu, i, j - 1);
        double mu_xp = harm(avg2(AT(mu, i + 1, j - 1), AT(mu, i + 1, j)), avg2(AT(mu, i, j - 1), AT(mu, i, j)));
        double mu_xm = harm(avg2(AT(mu, i, j - 1), AT(mu, i, j)), avg2(AT(mu, i - 1, j - 1), AT(mu, i - 1, j)));
        dx = iRe * ((mu_xp * g.idx_c[i] + mu_xm * g.idx_c[i - 1]) * g.idx_u[i]);
        dy = iRe * (mu_yp * g.idy_v[j] + mu_ym * g.idy_v[j - 1]) * g.idy_c[j - 1];
    }
    return d_time + dx + dy + 0.0 + 0.0;
}

__device__ __forceinline__ double mom_op(const pvk_grid &g, const pvk_phys &ph, double alpha_k, double dt,
                                         const double *rho, const double *mu, const double *x, int comp, int i,
                                         int j)
{
    const int P = g.P;
    const double iRe = ph.iRe;
    double xc = AT(x, i, j), op = 0.0;
    if (comp == 'u') {
        double rho_face = 0.5 * (AT(rho, i, j) + AT(rho, i - 1, j));
        double factor_time = rho_face / (alpha_k * dt);
        double Ax_val = factor_time * xc;
        double mu_xp = AT(mu, i, j), mu_xm = AT(mu, i - 1, j);
        double mu_yp = harm(avg2(AT(mu, i - 1, j + 1), AT(mu, i, j + 1)), avg2(AT(mu, i - 1, j), AT(mu, i, j)));
        double mu_ym = harm(avg2(AT(mu, i - 1, j), AT(mu, i, j)), avg2(AT(mu, i - 1, j - 1), AT(mu, i, j - 1)));
        double flux_xp = mu_xp * (AT(x, i + 1, j) - xc) * g.idx_u[i];
        double flux_xm = mu_xm * (xc - AT(x, i - 1, j)) * g.idx_u[i - 1];
        double flux_yp = mu_yp * (AT(x, i, j + 1) - xc) * g.idy_c[j];
        double flux_ym = mu_ym * (xc - AT(x, i, j - 1)) * g.idy_c[j - 1];
        op += iRe * ((flux_xp - flux_xm) * g.idx_c[i - 1]);
        op += iRe * (flux_yp - flux_ym) * g.idy_v[j];
        op += 0.0;
        return Ax_val - op;
    } else {
        double rho_face = 0.5 * (AT(rho, i, j) + AT(rho, i, j - 1));
        double factor_time = rho_face / (alpha_k * dt);
        double Ax_val = factor_time * xc;
        double mu_yp = AT(mu, i, j), mu_ym = AT(mu, i, j - 1);
        double mu_xp = harm(avg2(AT(mu, i + 1, j - 1), AT(mu, i + 1, j)), avg2(AT(mu, i, j - 1), AT(mu, i, j)));
        double mu_xm = harm(avg2(AT(mu, i, j - 1), AT(mu, i, j)), avg2(AT(mu, i - 1, j - 1), AT(mu, i - 1, j)));
        double flux_xp = mu_xp * (AT(x, i + 1, j) - xc) * g.idx_c[i];
        double flux_xm = mu_xm * (xc - AT(x, i - 1, j)) * g.idx_c[i - 1];
        double flux_yp = mu_yp * (AT(x, i, j + 1) - xc) * g.idy_v[j];
        double flux_ym = mu_ym * (xc - AT(x, i, j - 1)) * g.idy_v[j - 1];
        op += iRe * ((flux_xp - flux_xm) * g.idx_u[i]);
        op += iRe * (flux_yp - flux_ym) * g.idy_c[j - 1];
        return Ax_val - op;
    }
}
#endif
