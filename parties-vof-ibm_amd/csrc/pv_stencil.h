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

// Arithmetic type R: double for the FP64 path, float for the mixed path's
// FP32 stencils.  Inputs are FP64 fields converted on load; every literal is
// written R(...) so a float instantiation never promotes to double.  With
// R = double each expression is the FP64 one, operand for operand.

// Same formula as lap_at applied to five already-evaluated values.
template <class R = double>
__device__ __forceinline__ R lap_vals(const pvk_grid &g, int i, int j, R c, R E, R W, R N, R S)
{
    R dfdxE = (E - c) * R(g.idx_c[i]);
    R dfdxW = (c - W) * R((i != 0) ? g.idx_c[i - 1] : g.idx_c[i]);
    R dfdyN = (N - c) * R(g.idy_c[j]);
    R dfdyS = (c - S) * R((j != 0) ? g.idy_c[j - 1] : g.idy_c[j]);
    return (dfdxE - dfdxW) * R(g.idx_u[i]) + (dfdyN - dfdyS) * R(g.idy_v[j]);
}

// TwodOps_scalar_laplacian, planar (TwodOps.c:159-199)
template <class R = double>
__device__ __forceinline__ R lap_at(const pvk_grid &g, const double *f, int i, int j)
{
    const int P = g.P;
    return lap_vals<R>(g, i, j, R(AT(f, i, j)), R(AT(f, i + 1, j)), R(AT(f, i - 1, j)), R(AT(f, i, j + 1)),
                       R(AT(f, i, j - 1)));
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
template <class R = double>
__device__ __forceinline__ R bulkS(R c)
{
    const R cs = R(0.0);
    return R(0.5) * c - R(1.5) * c * c + c * c * c - c * cs * (R(1.0) - c - cs);
}

// diffuse_weno5_left (VOF_DIFFUSE.c:431-453); pow(x, 2.0) evaluated as x*x
template <class R = double>
__device__ __forceinline__ R weno5(R v0, R v1, R v2, R v3, R v4)
{
    const R eps = R(1e-6);
    R q0 = (R(2.0) * v0 - R(7.0) * v1 + R(11.0) * v2) / R(6.0);
    R q1 = (-v1 + R(5.0) * v2 + R(2.0) * v3) / R(6.0);
    R q2 = (R(2.0) * v2 + R(5.0) * v3 - v4) / R(6.0);
    R t0 = v0 - R(2.0) * v1 + v2, t1 = v0 - R(4.0) * v1 + R(3.0) * v2;
    R t2 = v1 - R(2.0) * v2 + v3, t3 = v1 - v3;
    R t4 = v2 - R(2.0) * v3 + v4, t5 = R(3.0) * v2 - R(4.0) * v3 + v4;
    R b0 = R(13.0 / 12.0) * (t0 * t0) + R(0.25) * (t1 * t1);
    R b1 = R(13.0 / 12.0) * (t2 * t2) + R(0.25) * (t3 * t3);
    R b2 = R(13.0 / 12.0) * (t4 * t4) + R(0.25) * (t5 * t5);
    R a0 = R(0.1) / ((eps + b0) * (eps + b0));
    R a1 = R(0.6) / ((eps + b1) * (eps + b1));
    R a2 = R(0.3) / ((eps + b2) * (eps + b2));
    R asum = a0 + a1 + a2;
    return (a0 * q0 + a1 * q1 + a2 * q2) / asum;
}

// diffuse_face_state_x / _y (VOF_DIFFUSE.c:455-483) on the six cells around
// a face, cm3 .. cp2 = C at face - 3 .. face + 2 along the direction (the
// face sits between cells face - 1 and face)
template <class R = double>
__device__ __forceinline__ R face_vals(double cm3, double cm2, double cm1, double c0, double cp1, double cp2,
                                       double vel, int wo)
{
    if (wo == 5) {
        if (vel >= 0.0) return weno5<R>(R(cm3), R(cm2), R(cm1), R(c0), R(cp1));
        return weno5<R>(R(cp2), R(cp1), R(c0), R(cm1), R(cm2));
    }
    return R((vel >= 0.0) ? cm1 : c0);
}
template <class R = double>
__device__ __forceinline__ R face_x(const pvk_grid &g, const double *C, int i, int j, double vel, int wo)
{
    const int P = g.P;
    if (wo == 5)
        return face_vals<R>(AT(C, i - 3, j), AT(C, i - 2, j), AT(C, i - 1, j), AT(C, i, j), AT(C, i + 1, j),
                            AT(C, i + 2, j), vel, wo);
    return face_vals<R>(0.0, 0.0, AT(C, i - 1, j), AT(C, i, j), 0.0, 0.0, vel, wo);
}
template <class R = double>
__device__ __forceinline__ R face_y(const pvk_grid &g, const double *C, int i, int j, double vel, int wo)
{
    const int P = g.P;
    if (wo == 5)
        return face_vals<R>(AT(C, i, j - 3), AT(C, i, j - 2), AT(C, i, j - 1), AT(C, i, j), AT(C, i, j + 1),
                            AT(C, i, j + 2), vel, wo);
    return face_vals<R>(0.0, 0.0, AT(C, i, j - 1), AT(C, i, j), 0.0, 0.0, vel, wo);
}

// -(div F) at cell (i,j) from its four face fluxes (TwodOps.c:201-229)
__device__ __forceinline__ double weno_div(const pvk_grid &g, int i, int j, double fxW, double fxE, double fyS,
                                           double fyN)
{
    double div_x = (fxE - fxW) / g.dx_c[i];
    double div_y = (fyN - fyS) / g.dy_c[j];
    double div = div_x + div_y;
    return -div;
}

// -(div(u C)) at interior cell (i,j): VOF_DIFFUSE.c:931-996 + TwodOps.c:201-229.
// The face states are reconstructed in R; fluxes u C_face and their
// divergence are always accumulated in FP64 (conservative flux form).
template <class R = double>
__device__ __forceinline__ double weno_rhs_at(const pvk_grid &g, const pvk_phys &ph, const double *C,
                                              const double *u, const double *v, int i, int j)
{
    const int P = g.P;
    double uw = AT(u, i, j), ue = AT(u, i + 1, j);
    double vs = AT(v, i, j), vn = AT(v, i, j + 1);
    double fxW = uw * (double)face_x<R>(g, C, i, j, uw, ph.weno_order);
    double fxE = ue * (double)face_x<R>(g, C, i + 1, j, ue, ph.weno_order);
    double fyS = vs * (double)face_y<R>(g, C, i, j, vs, ph.weno_order);
    double fyN = vn * (double)face_y<R>(g, C, i, j + 1, vn, ph.weno_order);
    return weno_div(g, i, j, fxW, fxE, fyS, fyN);
}

// L(S(C)) / Pe at interior (i,j) (VOF_DIFFUSE.c:998-1016); C ghosts valid.
// S and the face gradients are formed in R; the flux differences are FP64,
// so a face flux enters its two cells as the same double and the sum of the
// term over the domain telescopes to FP64 round-off (phase volume kept).
template <class R = double>
__device__ __forceinline__ double diffusion_vals(const pvk_grid &g, const pvk_phys &ph, int i, int j, R S0, R SE,
                                                 R SW, R SN, R SS)
{
    const double dfdxE = (double)((SE - S0) * R(g.idx_c[i]));
    const double dfdxW = (double)((S0 - SW) * R((i != 0) ? g.idx_c[i - 1] : g.idx_c[i]));
    const double dfdyN = (double)((SN - S0) * R(g.idy_c[j]));
    const double dfdyS = (double)((S0 - SS) * R((j != 0) ? g.idy_c[j - 1] : g.idy_c[j]));
    const double lapS = (dfdxE - dfdxW) * g.idx_u[i] + (dfdyN - dfdyS) * g.idy_v[j];
    return lapS / ph.Pe_CH;
}
template <class R = double>
__device__ __forceinline__ double diffusion_at(const pvk_grid &g, const pvk_phys &ph, const double *C, int i, int j)
{
    const int P = g.P;
    return diffusion_vals<R>(g, ph, i, j, bulkS<R>(R(AT(C, i, j))), bulkS<R>(R(AT(C, i + 1, j))),
                             bulkS<R>(R(AT(C, i - 1, j))), bulkS<R>(R(AT(C, i, j + 1))), bulkS<R>(R(AT(C, i, j - 1))));
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
template <class R = double>
__device__ __forceinline__ R harm(R a, R b) { return R(2.0) * (a) * (b) / ((a) + (b) + R(1e-12)); }
template <class R = double>
__device__ __forceinline__ R avg2(R a, R b) { return R(0.5) * ((a) + (b)); }

// Jacobi diagonal (lsolver/msolve_cg.c:103-249) and operator (:255-493).
template <class R = double>
__device__ __forceinline__ R mom_diag(const pvk_grid &g, const pvk_phys &ph, double alpha_k, double dt,
                                      const double *rho, const double *mu, int comp, int i, int j)
{
    const int P = g.P;
    const R iRe = R(ph.iRe);
    R rho_face, d_time, dx, dy;
    if (comp == 'u') {
        rho_face = R(0.5) * (R(AT(rho, i, j)) + R(AT(rho, i - 1, j)));
        d_time = rho_face / (R(alpha_k) * R(dt));
        R mu_xp = R(AT(mu, i, j)), mu_xm = R(AT(mu, i - 1, j));
        R mu_yp = harm<R>(avg2<R>(R(AT(mu, i - 1, j + 1)), R(AT(mu, i, j + 1))), avg2<R>(R(AT(mu, i - 1, j)), R(AT(mu, i, j))));
        R mu_ym = harm<R>(avg2<R>(R(AT(mu, i - 1, j)), R(AT(mu, i, j))), avg2<R>(R(AT(mu, i - 1, j - 1)), R(AT(mu, i, j - 1))));
        dx = iRe * ((mu_xp * R(g.idx_u[i]) + mu_xm * R(g.idx_u[i - 1])) * R(g.idx_c[i - 1]));
        dy = iRe * (mu_yp * R(g.idy_c[j]) + mu_ym * R(g.idy_c[j - 1])) * R(g.idy_v[j]);
    } else {
        rho_face = R(0.5) * (R(AT(rho, i, j)) + R(AT(rho, i, j - 1)));
        d_time = rho_face / (R(alpha_k) * R(dt));
        R mu_yp = R(AT(mu, i, j)), mu_ym = R(AT(mu, i, j - 1));
        R mu_xp = harm<R>(avg2<R>(R(AT(mu, i + 1, j - 1)), R(AT(mu, i + 1, j))), avg2<R>(R(AT(mu, i, j - 1)), R(AT(mu, i, j))));
        R mu_xm = harm<R>(avg2<R>(R(AT(mu, i, j - 1)), R(AT(mu, i, j))), avg2<R>(R(AT(mu, i - 1, j - 1)), R(AT(mu, i - 1, j))));
        dx = iRe * ((mu_xp * R(g.idx_c[i]) + mu_xm * R(g.idx_c[i - 1])) * R(g.idx_u[i]));
        dy = iRe * (mu_yp * R(g.idy_v[j]) + mu_ym * R(g.idy_v[j - 1])) * R(g.idy_c[j - 1]);
    }
    return d_time + dx + dy + R(0.0) + R(0.0);
}

// The operator with the five x values supplied (centre, E, W, N, S).
template <class R = double>
__device__ __forceinline__ R mom_op_vals(const pvk_grid &g, const pvk_phys &ph, double alpha_k, double dt,
                                         const double *rho, const double *mu, int comp, int i, int j, R xc, R xE,
                                         R xW, R xN, R xS)
{
    const int P = g.P;
    const R iRe = R(ph.iRe);
    R op = R(0.0);
    if (comp == 0) {
        R rho_face = R(0.5) * (R(AT(rho, i, j)) + R(AT(rho, i - 1, j)));
        R factor_time = rho_face / (R(alpha_k) * R(dt));
        R Ax_val = factor_time * xc;
        R mu_xp = R(AT(mu, i, j)), mu_xm = R(AT(mu, i - 1, j));
        R mu_yp = harm<R>(avg2<R>(R(AT(mu, i - 1, j + 1)), R(AT(mu, i, j + 1))), avg2<R>(R(AT(mu, i - 1, j)), R(AT(mu, i, j))));
        R mu_ym = harm<R>(avg2<R>(R(AT(mu, i - 1, j)), R(AT(mu, i, j))), avg2<R>(R(AT(mu, i - 1, j - 1)), R(AT(mu, i, j - 1))));
        R flux_xp = mu_xp * (xE - xc) * R(g.idx_u[i]);
        R flux_xm = mu_xm * (xc - xW) * R(g.idx_u[i - 1]);
        R flux_yp = mu_yp * (xN - xc) * R(g.idy_c[j]);
        R flux_ym = mu_ym * (xc - xS) * R(g.idy_c[j - 1]);
        op += iRe * ((flux_xp - flux_xm) * R(g.idx_c[i - 1]));
        op += iRe * (flux_yp - flux_ym) * R(g.idy_v[j]);
        op += R(0.0);
        return Ax_val - op;
    }
    R rho_face = R(0.5) * (R(AT(rho, i, j)) + R(AT(rho, i, j - 1)));
    R factor_time = rho_face / (R(alpha_k) * R(dt));
    R Ax_val = factor_time * xc;
    R mu_yp = R(AT(mu, i, j)), mu_ym = R(AT(mu, i, j - 1));
    R mu_xp = harm<R>(avg2<R>(R(AT(mu, i + 1, j - 1)), R(AT(mu, i + 1, j))), avg2<R>(R(AT(mu, i, j - 1)), R(AT(mu, i, j))));
    R mu_xm = harm<R>(avg2<R>(R(AT(mu, i, j - 1)), R(AT(mu, i, j))), avg2<R>(R(AT(mu, i - 1, j - 1)), R(AT(mu, i - 1, j))));
    R flux_xp = mu_xp * (xE - xc) * R(g.idx_c[i]);
    R flux_xm = mu_xm * (xc - xW) * R(g.idx_c[i - 1]);
    R flux_yp = mu_yp * (xN - xc) * R(g.idy_v[j]);
    R flux_ym = mu_ym * (xc - xS) * R(g.idy_v[j - 1]);
    op += iRe * ((flux_xp - flux_xm) * R(g.idx_u[i]));
    op += iRe * (flux_yp - flux_ym) * R(g.idy_c[j - 1]);
    return Ax_val - op;
}

// The operator's coefficients at (i, j) in FP32 (centre, E, W, N, S):
// A x = c0 x - ((cE (xE - x) - cW (x - xW)) + (cN (xN - x) - cS (x - xS))),
// the products of mom_op_vals with the x differences factored out.
__device__ __forceinline__ void mom_coeffs(const pvk_grid &g, const pvk_phys &ph, double alpha_k, double dt,
                                           const double *rho, const double *mu, int comp, int i, int j, float c[5])
{
    const int P = g.P;
    const float iRe = float(ph.iRe);
    if (comp == 0) {
        const float rho_face = 0.5f * (float(AT(rho, i, j)) + float(AT(rho, i - 1, j)));
        c[0] = rho_face / (float(alpha_k) * float(dt));
        const float mu_xp = float(AT(mu, i, j)), mu_xm = float(AT(mu, i - 1, j));
        const float mu_yp = harm<float>(avg2<float>(float(AT(mu, i - 1, j + 1)), float(AT(mu, i, j + 1))),
                                        avg2<float>(float(AT(mu, i - 1, j)), float(AT(mu, i, j))));
        const float mu_ym = harm<float>(avg2<float>(float(AT(mu, i - 1, j)), float(AT(mu, i, j))),
                                        avg2<float>(float(AT(mu, i - 1, j - 1)), float(AT(mu, i, j - 1))));
        const float sx = iRe * float(g.idx_c[i - 1]), sy = iRe * float(g.idy_v[j]);
        c[1] = sx * (mu_xp * float(g.idx_u[i]));
        c[2] = sx * (mu_xm * float(g.idx_u[i - 1]));
        c[3] = sy * (mu_yp * float(g.idy_c[j]));
        c[4] = sy * (mu_ym * float(g.idy_c[j - 1]));
        return;
    }
    const float rho_face = 0.5f * (float(AT(rho, i, j)) + float(AT(rho, i, j - 1)));
    c[0] = rho_face / (float(alpha_k) * float(dt));
    const float mu_yp = float(AT(mu, i, j)), mu_ym = float(AT(mu, i, j - 1));
    const float mu_xp = harm<float>(avg2<float>(float(AT(mu, i + 1, j - 1)), float(AT(mu, i + 1, j))),
                                    avg2<float>(float(AT(mu, i, j - 1)), float(AT(mu, i, j))));
    const float mu_xm = harm<float>(avg2<float>(float(AT(mu, i, j - 1)), float(AT(mu, i, j))),
                                    avg2<float>(float(AT(mu, i - 1, j - 1)), float(AT(mu, i - 1, j))));
    const float sx = iRe * float(g.idx_u[i]), sy = iRe * float(g.idy_c[j - 1]);
    c[1] = sx * (mu_xp * float(g.idx_c[i]));
    c[2] = sx * (mu_xm * float(g.idx_c[i - 1]));
    c[3] = sy * (mu_yp * float(g.idy_v[j]));
    c[4] = sy * (mu_ym * float(g.idy_v[j - 1]));
}

__device__ __forceinline__ float mom_op_coeffs(const float c[5], float xc, float xE, float xW, float xN, float xS)
{
    return c[0] * xc - ((c[1] * (xE - xc) - c[2] * (xc - xW)) + (c[3] * (xN - xc) - c[4] * (xc - xS)));
}

template <class R = double>
__device__ __forceinline__ R mom_op(const pvk_grid &g, const pvk_phys &ph, double alpha_k, double dt,
                                    const double *rho, const double *mu, const double *x, int comp, int i, int j)
{
    const int P = g.P;
    return mom_op_vals<R>(g, ph, alpha_k, dt, rho, mu, comp == 'u' ? 0 : 1, i, j, R(AT(x, i, j)),
                          R(AT(x, i + 1, j)), R(AT(x, i - 1, j)), R(AT(x, i, j + 1)), R(AT(x, i, j - 1)));
}
#endif
