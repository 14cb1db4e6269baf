"""CPU checks that pin the oracle (oracle/pvo.c) without a GPU.

The reference ships no golden vectors or known-answer tests for this path and
cannot be built here (its DataTypes.h includes fftw3.h, absent from the image),
so each oracle routine is pinned by what the mathematics fixes: discrete
eigenpairs, exactness classes, an independent spectral solve of the same
discrete system, symmetry/definiteness, conservation identities and the RK3
coefficient identities.  Each check is chosen so that a dropped term, a wrong
sign or a wrong index in the oracle makes it fail.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib
import pvof
from helpers import oracle_state, regions


@pytest.fixture(scope="module")
def olib():
    return oracle_lib.load_oracle()


def make(olib, **kw):
    base = dict(NXM=32, NYM=24, xmax=1.0, ymax=0.75)
    base.update(kw)
    return oracle_lib.Oracle(pvof.config1_params(**base), olib)


def cell_xy(o):
    p = o.params
    dx = (p.xmax - p.xmin) / p.NXM
    dy = (p.ymax - p.ymin) / p.NYM
    x = p.xmin + (np.arange(p.NXM) + 0.5) * dx
    y = p.ymin + (np.arange(p.NYM) + 0.5) * dy
    return x, y, dx, dy


# ---------------------------------------------------------------- RK3 tables
def test_rk3_identities(olib):
    t = (C.c_double * 12)()
    olib.pvo_rk3_tables(t)
    bet, bet2, gamb, zetb = np.array(t[0:3]), np.array(t[3:6]), np.array(t[6:9]), np.array(t[9:12])
    gamma = np.array([8 / 15, 5 / 12, 3 / 4])
    zeta = np.array([0.0, -17 / 60, -5 / 12])
    np.testing.assert_allclose(gamma + zeta, 2 * bet, rtol=1e-15)      # alpha_k = 2 beta_k
    np.testing.assert_allclose(bet2, 2 * bet, rtol=1e-15)
    assert abs(bet2.sum() - 1.0) < 1e-15                                 # sum alpha = 1
    np.testing.assert_allclose(gamb, gamma / bet, rtol=1e-15)
    np.testing.assert_allclose(zetb, zeta / bet, rtol=1e-15)


# ---------------------------------------------------------------- Laplacian
@pytest.mark.parametrize("k", [1, 3, 7])
def test_laplacian_periodic_x_eigenpair(olib, k):
    o = make(olib)
    x, y, dx, dy = cell_xy(o)
    p = o.params
    R = regions(p)
    f = np.zeros(o.shape)
    f[R["cell"]] = np.cos(2 * np.pi * k * x)[None, :] * np.ones((p.NYM, 1))
    o.set("SCRATCH0", f)
    o.call("VOF_DIFFUSE_compute_laplacian", pvof.F["SCRATCH0"], pvof.F["SCRATCH1"])
    lam = -(4.0 / dx**2) * np.sin(np.pi * k / p.NXM) ** 2
    np.testing.assert_allclose(o.get("SCRATCH1")[R["cell"]], lam * f[R["cell"]], atol=1e-9 * abs(lam))


@pytest.mark.parametrize("k", [1, 2, 5])
def test_laplacian_neumann_y_eigenpair(olib, k):
    o = make(olib)
    p = o.params
    x, y, dx, dy = cell_xy(o)
    R = regions(p)
    j = np.arange(p.NYM)
    f = np.zeros(o.shape)
    f[R["cell"]] = np.cos(np.pi * k * (j + 0.5) / p.NYM)[:, None] * np.ones((1, p.NXM))
    o.set("SCRATCH0", f)
    o.call("VOF_DIFFUSE_compute_laplacian", pvof.F["SCRATCH0"], pvof.F["SCRATCH1"])
    lam = -(4.0 / dy**2) * np.sin(np.pi * k / (2 * p.NYM)) ** 2
    np.testing.assert_allclose(o.get("SCRATCH1")[R["cell"]], lam * f[R["cell"]], atol=1e-9 * abs(lam))


def test_bulk_S_and_psi_pointwise(olib):
    o = make(olib)
    p = o.params
    R = regions(p)
    rng = np.random.default_rng(3)
    c = np.zeros(o.shape)
    c[R["cell"]] = rng.uniform(-0.1, 1.1, (p.NYM, p.NXM))
    o.set("C_L", c)
    o.call("VOF_DIFFUSE_compute_bulk_S")
    cc = c[R["cell"]]
    np.testing.assert_allclose(o.get("BULK_S")[R["cell"]], cc * (cc - 0.5) * (cc - 1.0), atol=1e-15)
    o.call("VOF_DIFFUSE_compute_psi_LG")
    lap = o.get("LAP_C")[R["cell"]]
    np.testing.assert_allclose(o.get("PSI_LG")[R["cell"]],
                               cc * (cc - 0.5) * (cc - 1.0) - p.Cn**2 * lap, atol=1e-13)


# ---------------------------------------------------------------- WENO5
def quad_avg(a, b, c, x0, x1):
    """Cell average of a + b x + c x^2 over [x0, x1]."""
    return a + b * (x0 + x1) / 2 + c * (x0**2 + x0 * x1 + x1**2) / 3


def test_weno_exact_for_quadratic_cell_averages(olib):
    o = make(olib, NXM=64, NYM=16)
    p = o.params
    R = regions(p)
    x, y, dx, dy = cell_xy(o)
    a, b, c, u0 = 0.3, 0.7, -0.4, 0.8
    xe = p.xmin + np.arange(p.NXM + 1) * dx
    C_ = np.zeros(o.shape)
    C_[R["cell"]] = quad_avg(a, b, c, xe[:-1], xe[1:])[None, :]
    o.set("C_L", C_)
    U = np.zeros(o.shape)
    U[:, :] = u0
    o.set("U", U)
    o.set("V", np.zeros(o.shape))
    o.call("VOF_DIFFUSE_advect_WENO5", pvof.F["CH_AUX2"])
    rhs = o.get("CH_AUX2")[R["cell"]]
    pface = a + b * xe + c * xe**2
    expect = -u0 * (pface[1:] - pface[:-1]) / dx
    # cells whose 7-point stencil stays clear of the periodic seam
    np.testing.assert_allclose(rhs[:, 4:-4], np.broadcast_to(expect[4:-4], rhs[:, 4:-4].shape), rtol=0,
                               atol=1e-9)


def test_weno_constant_state_and_telescoping(olib):
    o = make(olib)
    p = o.params
    R = regions(p)
    rng = np.random.default_rng(7)
    o.set("C_L", np.full(o.shape, 0.37))
    U = np.zeros(o.shape)
    U[R["G"]] = rng.uniform(-1, 1, (p.NYM + 1, p.NXM + 1))
    o.set("U", U)
    o.call("Velocity_update_boundaries", pvof.F["U"], pvof.PV_VEL_TYPE_NORMAL)
    V = np.zeros(o.shape)
    V[R["G"]] = rng.uniform(-1, 1, (p.NYM + 1, p.NXM + 1))
    o.set("V", V)
    o.call("Velocity_update_boundaries", pvof.F["V"], pvof.PV_VEL_TYPE_NORMAL)
    o.call("VOF_DIFFUSE_advect_WENO5", pvof.F["CH_AUX2"])
    u = o.get("U")
    v = o.get("V")
    x, y, dx, dy = cell_xy(o)
    # constant state: face value = 0.37, rhs = -0.37 div(u)
    divu = (u[3:3 + p.NYM, 4:4 + p.NXM] - u[3:3 + p.NYM, 3:3 + p.NXM]) / dx + \
           (v[4:4 + p.NYM, 3:3 + p.NXM] - v[3:3 + p.NYM, 3:3 + p.NXM]) / dy
    np.testing.assert_allclose(o.get("CH_AUX2")[R["cell"]], -0.37 * divu, atol=1e-12)
    # random field: the flux divergence telescopes to zero (periodic x, v = 0 on walls)
    C_ = np.zeros(o.shape)
    C_[R["cell"]] = rng.uniform(0, 1, (p.NYM, p.NXM))
    o.set("C_L", C_)
    o.call("VOF_DIFFUSE_advect_WENO5", pvof.F["CH_AUX2"])
    assert abs(o.get("CH_AUX2")[R["cell"]].sum()) < 1e-11 * np.abs(o.get("CH_AUX2")).max() * p.NXM * p.NYM


def test_weno_mirror_symmetry(olib):
    o1, o2 = make(olib), make(olib)
    p = o1.params
    R = regions(p)
    rng = np.random.default_rng(11)
    c = rng.uniform(0, 1, (p.NYM, p.NXM))
    u = rng.uniform(-1, 1, (p.NYM, p.NXM))          # u at faces i = 0..NXM-1
    for o, cc, uu in ((o1, c, u), (o2, c[:, ::-1], -np.roll(u[:, ::-1], 1, axis=1))):
        C_ = np.zeros(o.shape)
        C_[R["cell"]] = cc
        o.set("C_L", C_)
        U = np.zeros(o.shape)
        U[3:3 + p.NYM, 3:3 + p.NXM] = uu
        U[3:3 + p.NYM, 3 + p.NXM] = uu[:, 0]
        o.set("U", U)
        o.call("Velocity_update_boundaries", pvof.F["U"], pvof.PV_VEL_TYPE_NORMAL)
        o.set("V", np.zeros(o.shape))
        o.call("VOF_DIFFUSE_advect_WENO5", pvof.F["CH_AUX2"])
    r1 = o1.get("CH_AUX2")[R["cell"]]
    r2 = o2.get("CH_AUX2")[R["cell"]]
    np.testing.assert_allclose(r2, r1[:, ::-1], atol=1e-12 * np.abs(r1).max())


def weno_smooth_error(olib, n, amp=0.25, phase=0.3):
    """Max error of the WENO5 flux divergence for u = 1 on cell averages of
    0.5 + amp sin(2 pi x + phase); exact: -(c(x_{i+1/2}) - c(x_{i-1/2})) / dx."""
    o = make(olib, NXM=n, NYM=8, ymax=8.0 / n)
    p = o.params
    R = regions(p)
    dx = 1.0 / n
    xe = np.arange(n + 1) * dx
    prim = 0.5 * xe - amp / (2 * np.pi) * np.cos(2 * np.pi * xe + phase)
    cface = 0.5 + amp * np.sin(2 * np.pi * xe + phase)
    C_ = np.zeros(o.shape)
    C_[R["cell"]] = ((prim[1:] - prim[:-1]) / dx)[None, :]
    o.set("C_L", C_)
    o.set("U", np.ones(o.shape))
    o.set("V", np.zeros(o.shape))
    o.call("VOF_DIFFUSE_advect_WENO5", pvof.F["CH_AUX2"])
    rhs = o.get("CH_AUX2")[R["cell"]]
    return float(np.abs(rhs + (cface[1:] - cface[:-1])[None, :] / dx).max())


def test_weno_fifth_order_on_smooth_data(olib):
    """Only the linear weights 0.1/0.6/0.3 make the three 3-point candidates a
    5th-order scheme on smooth data; any other convex weighting (or a wrong
    smoothness indicator) leaves 3rd order, which the quadratic test cannot see."""
    e = [weno_smooth_error(olib, n) for n in (64, 128, 256)]
    orders = [np.log2(e[k] / e[k + 1]) for k in range(2)]
    assert min(orders) > 4.6, (e, orders)


# ---------------------------------------------------------------- biharmonic
def spectral_biharmonic_solve(b, a_dt, coef, dx, dy):
    """Exact inverse of a_dt I + coef L^2 (periodic x, Neumann y) by FFT x DCT."""
    from scipy.fft import dct, idct
    ny, nx = b.shape
    kx = np.fft.fftfreq(nx) * nx
    lx = -(4 / dx**2) * np.sin(np.pi * kx / nx) ** 2
    ly = -(4 / dy**2) * np.sin(np.pi * np.arange(ny) / (2 * ny)) ** 2
    lam = (lx[None, :] + ly[:, None])
    B = np.fft.fft(dct(b, type=2, axis=0, norm="ortho"), axis=1)
    X = B / (a_dt + coef * lam**2)
    return idct(np.real(np.fft.ifft(X, axis=1)), type=2, axis=0, norm="ortho")


def test_biharmonic_cg_matches_spectral_solve(olib):
    o = make(olib, ch_tol=1e-14, ch_iter_max=2000)
    p = o.params
    R = regions(p)
    x, y, dx, dy = cell_xy(o)
    rng = np.random.default_rng(13)
    b = rng.standard_normal((p.NYM, p.NXM))
    dt, stage = 2e-4, 1
    o.call("set_time_state", 0.0, dt, dt, stage)
    o.set("C_L", np.zeros(o.shape))
    B = np.zeros(o.shape)
    B[R["cell"]] = b
    o.set("SCRATCH0", B)
    o.call("VOF_DIFFUSE_set_boundary_values", pvof.F["SCRATCH0"])
    rel = C.c_double()
    o.lib.pvo_VOF_DIFFUSE_solve_implicit_biharmonic(o.h, pvof.F["SCRATCH0"], C.byref(rel))
    a_dt = 1.0 / (dt * (1.0 / 15.0))
    coef = p.Cn**2 / p.Pe_CH
    xs = spectral_biharmonic_solve(b, a_dt, coef, dx, dy)
    got = o.get("C_L")[R["cell"]]
    np.testing.assert_allclose(got, xs, atol=1e-11 * np.abs(xs).max())
    # the mean mode is solved exactly by construction: mean(x) = mean(b) / a_dt
    assert abs(got.mean() - b.mean() / a_dt) < 1e-12 * np.abs(xs).max()


# ---------------------------------------------------------------- clamp
def test_clamp_bounds_idempotence_and_delta(olib):
    o = make(olib)
    p = o.params
    R = regions(p)
    rng = np.random.default_rng(17)
    c = np.zeros(o.shape)
    c[R["G"]] = rng.uniform(-0.2, 1.2, (p.NYM + 1, p.NXM + 1))
    o.set("C_L", c)
    o.call("VOF_DIFFUSE_set_boundary_values", pvof.F["C_L"])
    before = o.get("C_L")
    d1 = o.lib.pvo_VOF_DIFFUSE_bound_liquid_fraction(o.h)
    after = o.get("C_L")
    assert after[R["G"]].min() >= 0 and after[R["G"]].max() <= 1
    np.testing.assert_array_equal(o.get("C_G")[R["G"]], 1.0 - after[R["G"]])
    x, y, dx, dy = cell_xy(o)
    meas = dx * dy * o.params.twod_slab_thickness
    expect = 2 * ((after - before)[R["G"]] * meas / 2).sum()      # two z slabs, half measure each
    assert abs(d1 - expect) < 1e-15 * p.NXM * p.NYM
    assert o.lib.pvo_VOF_DIFFUSE_bound_liquid_fraction(o.h) == 0.0    # idempotent


# ---------------------------------------------------------------- pressure
def random_rho(o, rng):
    p = o.params
    R = regions(p)
    rho = np.zeros(o.shape)
    rho[R["G"]] = rng.uniform(1e-3, 1.0, (p.NYM + 1, p.NXM + 1))
    o.set("RHO", rho)
    o.call("VOF_DIFFUSE_set_boundary_values", pvof.F["RHO"])


def apply_p(o, phi):
    R = regions(o.params)
    f = np.zeros(o.shape)
    f[R["cell"]] = phi
    o.set("SCRATCH0", f)
    o.call("Pressure_apply_BCs", pvof.F["SCRATCH0"])
    o.call("Pressure_operator_variableCoeff", pvof.F["SCRATCH1"], pvof.F["SCRATCH0"])
    return o.get("SCRATCH1")[R["cell"]]


def test_pressure_operator_symmetric_nsd_nullspace(olib):
    o = make(olib)
    p = o.params
    rng = np.random.default_rng(19)
    random_rho(o, rng)
    a, b = rng.standard_normal((2, p.NYM, p.NXM))
    Aa, Ab = apply_p(o, a), apply_p(o, b)
    assert abs((b * Aa).sum() - (a * Ab).sum()) < 1e-10 * abs((b * Aa).sum())
    assert (a * Aa).sum() < 0
    assert np.abs(apply_p(o, np.ones((p.NYM, p.NXM)))).max() < 1e-9
    # Jacobi preconditioner = 1 / (-diag + 1e-12); diag from A e_i
    o.call("Pressure_compute_preconditioner")
    M = o.get("M_INV")[regions(p)["cell"]]
    for (jj, ii) in ((0, 0), (5, 7), (p.NYM - 1, p.NXM - 1)):
        e = np.zeros((p.NYM, p.NXM))
        e[jj, ii] = 1.0
        diag = apply_p(o, e)[jj, ii]
        wall = (jj == 0) or (jj == p.NYM - 1)
        if not wall:
            assert abs(M[jj, ii] - 1.0 / (-diag + 1e-12)) < 1e-13 * M[jj, ii]


def test_pressure_uniform_rho_eigenpair(olib):
    o = make(olib)
    p = o.params
    o.set("RHO", np.ones(o.shape))
    x, y, dx, dy = cell_xy(o)
    j = np.arange(p.NYM)
    kx, ky = 2, 3
    phi = np.cos(2 * np.pi * kx * x)[None, :] * np.cos(np.pi * ky * (j + 0.5) / p.NYM)[:, None]
    lam = -(4 / dx**2) * np.sin(np.pi * kx / p.NXM) ** 2 - (4 / dy**2) * np.sin(np.pi * ky / (2 * p.NYM)) ** 2
    np.testing.assert_allclose(apply_p(o, phi), lam * phi, atol=1e-9 * abs(lam))


def test_fft_solve_residual_and_pin(olib):
    o = make(olib, NXM=64, NYM=32, ymax=0.5)
    p = o.params
    R = regions(p)
    rng = np.random.default_rng(23)
    rhs = rng.standard_normal((p.NYM, p.NXM))
    rhs -= rhs.mean()
    f = np.zeros(o.shape)
    f[R["cell"]] = rhs
    o.set("SCRATCH0", f)
    o.call("Pressure_solve_fft", pvof.F["SCRATCH0"], pvof.F["SCRATCH1"])
    z = o.get("SCRATCH1")[R["cell"]]
    o.set("RHO", np.ones(o.shape))
    Az = apply_p(o, z)
    assert np.abs(Az - rhs).max() < 1e-12 * np.abs(rhs).max() * p.NXM
    assert abs(z[0].mean()) < 1e-14 * np.abs(z).max()            # (0,0) mode pinned at j = 0


# ---------------------------------------------------------------- momentum
def test_momentum_operator_spd(olib):
    p = pvof.config1_params(NXM=32, NYM=64)
    F = pvof.ic(pvof.load_lib(), "rising_bubble", p)
    o = oracle_state(p, F, 2, olib)
    R = regions(o.params)
    rng = np.random.default_rng(29)
    for comp, fld, reg in (("u", "U_D", "u"), ("v", "V_D", "v")):
        vals = []
        xs = []
        for _ in range(2):
            x = np.zeros(o.shape)
            x[R[reg]] = rng.standard_normal(x[R[reg]].shape)
            o.set(fld, x)
            o.call("Velocity_update_boundaries", pvof.F[fld], pvof.PV_VEL_TYPE_CG)
            o.call("matVec", pvof.F["SCRATCH0"], pvof.F[fld], ord(comp))
            vals.append(o.get("SCRATCH0")[R[reg]])
            xs.append(x[R[reg]])
        assert abs((xs[1] * vals[0]).sum() - (xs[0] * vals[1]).sum()) < 1e-10 * abs((xs[1] * vals[0]).sum())
        assert (xs[0] * vals[0]).sum() > 0


# ---------------------------------------------------------------- time step
def test_dt_limiter_truth_table(olib):
    p = pvof.config1_params(NXM=32, NYM=64, max_dt=1e-3)
    F = pvof.ic(pvof.load_lib(), "rising_bubble", p)
    o = oracle_state(p, F, 2, olib)
    o.call("Velocity_cell_center")
    cfl_dt = o.lib.pvo_Dtime_cfl
    o.call("set_time_state", 0.1, 1.0, 1.0, 0)            # dt_cfl << 0.95 dt_old: floor, max_dt skipped
    assert cfl_dt(o.h) == 0.95 * 1.0
    o.call("set_time_state", 0.1, 1e-9, 1e-9, 0)          # dt_cfl >> 1.05 dt_old: cap at 1.05 dt_old
    assert cfl_dt(o.h) == 1.05 * 1e-9


# ---------------------------------------------------------------- substep invariants
def test_substep_mass_and_divergence(olib):
    p = pvof.config1_params(NXM=32, NYM=64, ch_tol=1e-14, ch_iter_max=1000, CG_ETOL=1e-7,
                            P_CG_ETOL=1e-12, P_CG_MAXIT=6000)
    F = pvof.ic(pvof.load_lib(), "rising_bubble", p)
    o = oracle_lib.Oracle(p, olib)
    o.init_from_F(F)
    R = regions(o.params)
    x, y, dx, dy = cell_xy(o)
    meas = dx * dy * o.params.twod_slab_thickness
    for _ in range(3):
        m0 = o.get("C_L")[R["cell"]].sum() * meas
        assert o.call("Temporal_int_rk3", 1) == 0
        st = o.stats()
        m1 = o.get("C_L")[R["cell"]].sum() * meas
        assert abs((m1 - m0) - sum(st["clamp_delta"])) < 1e-13 * m0
        assert max(st["div_max"]) <= 1e-6
