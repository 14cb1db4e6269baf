"""CPU checks that pin the momentum, surface-tension, density and time-step
routines of the oracle (oracle/pvo.c) by what the mathematics or the physics
fixes, without a GPU and without the reference.

Each check uses a closed form that does not restate the oracle's expression:
exactness of centred differences on linear/quadratic data, discrete
eigenpairs, the grad-div identity, the identity between the implicit viscous
term and the operator the momentum CG inverts, the diagonal of that operator,
hydrostatic and capillary force balances, the surface tension of the
equilibrium tanh profile, the published low-storage RK3 coefficients
(Spalart, Moser & Rogers 1991: gamma = 8/15, 5/12, 3/4; zeta = 0, -17/60,
-5/12) and Brackbill's capillary time-step limit (Brackbill, Kothe & Zemach
1992).  test_oracle_mutants.py plants plausible mistakes in the oracle and
checks that these tests fail on them.
"""
import ctypes as C

import numpy as np
import pytest
from scipy.integrate import quad

import oracle_lib
import pvof
from helpers import regions

GAMMA = np.array([8 / 15, 5 / 12, 3 / 4])
ZETA = np.array([0.0, -17 / 60, -5 / 12])
BETA = (GAMMA + ZETA) / 2


@pytest.fixture(scope="module")
def olib():
    return oracle_lib.load_oracle()


def make(olib, **kw):
    base = dict(NXM=32, NYM=24, xmax=1.0, ymax=0.75)
    base.update(kw)
    return oracle_lib.Oracle(pvof.config1_params(**base), olib)


def coords(o):
    """Coordinates over the whole canonical frame: u faces (x_u, y_c), v faces
    (x_c, y_v) and cell centres (x_c, y_c); column c holds i = c - 3, row r
    holds j = r - 3."""
    p = o.params
    dx = (p.xmax - p.xmin) / p.NXM
    dy = (p.ymax - p.ymin) / p.NYM
    i = np.arange(p.NXM + 7) - 3
    j = np.arange(p.NYM + 7) - 3
    xu, xc = p.xmin + i * dx, p.xmin + (i + 0.5) * dx
    yv, yc = p.ymin + j * dy, p.ymin + (j + 0.5) * dy
    return {"xu": np.broadcast_to(xu[None, :], o.shape), "xc": np.broadcast_to(xc[None, :], o.shape),
            "yv": np.broadcast_to(yv[:, None], o.shape), "yc": np.broadcast_to(yc[:, None], o.shape),
            "dx": dx, "dy": dy, "i": i, "j": j}


def interior(p, kind):
    """u faces i = 1..NXM, j = 0..NYM-1; v faces i = 0..NXM-1, j = 1..NYM-1."""
    return regions(p)[kind]


# ---------------------------------------------------------------- cell centre (Velocity.c:278-357)
def test_cell_center_exact_for_linear_velocity(olib):
    o = make(olib)
    g = coords(o)
    p = o.params
    o.set("U", 0.2 + 0.7 * g["xu"] - 0.4 * g["yc"])
    o.set("V", -0.1 + 0.3 * g["xc"] + 0.9 * g["yv"])
    o.call("Velocity_cell_center")
    # the periodic wrap of u replaces u(0) by u(NXM): compare columns i = 1..NXM-2
    cols = slice(4, 3 + p.NXM - 1)
    rows = slice(3, 3 + p.NYM)
    ub = o.get("U_BC")[rows, cols]
    vb = o.get("V_BC")[rows, 3:3 + p.NXM]
    np.testing.assert_allclose(ub, (0.2 + 0.7 * g["xc"] - 0.4 * g["yc"])[rows, cols], atol=1e-14)
    np.testing.assert_allclose(vb, (-0.1 + 0.3 * g["xc"] + 0.9 * g["yc"])[rows, 3:3 + p.NXM], atol=1e-14)


# ---------------------------------------------------------------- momentum terms (Velocity.c:366-1144)
def linear_flow(o, a, b):
    g = coords(o)
    u = lambda x, y: a[0] + a[1] * x + a[2] * y          # noqa: E731
    v = lambda x, y: b[0] + b[1] * x + b[2] * y          # noqa: E731
    o.set("U", u(g["xu"], g["yc"]))
    o.set("V", v(g["xc"], g["yv"]))
    o.set("U_BC", u(g["xc"], g["yc"]))
    o.set("V_BC", v(g["xc"], g["yc"]))
    return g, u, v


def test_convective_terms_exact_for_linear_velocity(olib):
    """For linear u, v the products u^2, uv, v^2 are quadratic and the
    centred flux differences give their exact derivatives at the faces:
    E_u = -(d(uu)/dx + d(uv)/dy), E_v = -(d(vu)/dx + d(vv)/dy)."""
    o = make(olib)
    p = o.params
    a, b = (0.3, 0.8, -0.6), (-0.2, 0.5, 0.4)
    g, u, v = linear_flow(o, a, b)
    o.set("MU", np.full(o.shape, 1.3))
    o.call("Velocity_u_set_implicit_explicit")
    o.call("Velocity_v_set_implicit_explicit")
    Ru, Rv = interior(p, "u"), interior(p, "v")
    xu, yc = g["xu"][Ru], g["yc"][Ru]
    uu, vu = u(xu, yc), v(xu, yc)
    Eu = -(2 * uu * a[1] + a[2] * vu + uu * b[2])
    np.testing.assert_allclose(o.get("U_EXPL")[Ru], Eu, atol=1e-12)
    xc, yv = g["xc"][Rv], g["yv"][Rv]
    uv_, vv = u(xc, yv), v(xc, yv)
    Ev = -(b[1] * uv_ + vv * a[1] + 2 * vv * b[2])
    np.testing.assert_allclose(o.get("V_EXPL")[Rv], Ev, atol=1e-12)
    # uniform viscosity, linear velocity: no viscous force, implicit or transposed
    for f in ("U_IMPL", "U_VISC"):
        assert np.abs(o.get(f)[Ru]).max() < 1e-12, f
    for f in ("V_IMPL", "V_VISC"):
        assert np.abs(o.get(f)[Rv]).max() < 1e-12, f


@pytest.mark.parametrize("k,m", [(1, 1), (3, 2), (5, 4)])
def test_implicit_viscous_uniform_mu_is_laplacian_eigenpair(olib, k, m):
    """Uniform mu: L_imp = (mu/Re) times the 5-point Laplacian; on the wall
    modes (u odd about the walls, v zero on them) the discrete eigenvalues are
    -(4/dx^2) sin^2(pi k/NXM) - (4/dy^2) sin^2(pi m/(2 NYM))."""
    o = make(olib)
    p = o.params
    g = coords(o)
    mu0 = 1.7
    o.set("MU", np.full(o.shape, mu0))
    NX, NY = p.NXM, p.NYM
    u = np.cos(2 * np.pi * k * g["xu"]) * np.sin(np.pi * m * (g["j"][:, None] + 0.5) / NY)
    v = np.cos(2 * np.pi * k * g["xc"]) * np.sin(np.pi * m * g["j"][:, None] / NY)
    o.set("U", u)
    o.set("V", v)
    o.call("Velocity_update_boundaries", pvof.F["U"], pvof.PV_VEL_TYPE_NORMAL)
    o.call("Velocity_update_boundaries", pvof.F["V"], pvof.PV_VEL_TYPE_NORMAL)
    o.call("Velocity_u_set_implicit_explicit")
    o.call("Velocity_v_set_implicit_explicit")
    lam = -(4 / g["dx"] ** 2) * np.sin(np.pi * k / NX) ** 2 - (4 / g["dy"] ** 2) * np.sin(np.pi * m / (2 * NY)) ** 2
    Ru, Rv = interior(p, "u"), interior(p, "v")
    scale = mu0 / p.Re * abs(lam)
    np.testing.assert_allclose(o.get("U_IMPL")[Ru], mu0 / p.Re * lam * u[Ru], atol=1e-10 * scale)
    np.testing.assert_allclose(o.get("V_IMPL")[Rv], mu0 / p.Re * lam * v[Rv], atol=1e-10 * scale)


def test_transposed_viscous_uniform_mu_is_grad_div(olib):
    """Uniform mu: the transposed part d/dx(mu du/dx) + d/dy(mu dv/dx) is
    (mu/Re) times the x-difference of the cell divergence D (and likewise for
    v), because centred differences commute on a uniform grid."""
    o = make(olib)
    p = o.params
    g = coords(o)
    rng = np.random.default_rng(41)
    mu0 = 0.8
    o.set("MU", np.full(o.shape, mu0))
    o.set("U", rng.standard_normal(o.shape))
    o.set("V", rng.standard_normal(o.shape))
    o.call("Velocity_update_boundaries", pvof.F["U"], pvof.PV_VEL_TYPE_NORMAL)
    o.call("Velocity_update_boundaries", pvof.F["V"], pvof.PV_VEL_TYPE_NORMAL)
    o.call("Velocity_cell_center")
    o.call("Velocity_u_set_implicit_explicit")
    o.call("Velocity_v_set_implicit_explicit")
    u, v = o.get("U"), o.get("V")
    D = np.zeros(o.shape)
    D[:-1, :-1] = (u[:-1, 1:] - u[:-1, :-1]) / g["dx"] + (v[1:, :-1] - v[:-1, :-1]) / g["dy"]
    Ru, Rv = interior(p, "u"), interior(p, "v")
    gx = np.zeros(o.shape)
    gx[:, 1:] = (D[:, 1:] - D[:, :-1]) / g["dx"]
    gy = np.zeros(o.shape)
    gy[1:, :] = (D[1:, :] - D[:-1, :]) / g["dy"]
    scale = mu0 / p.Re * np.abs(gx[Ru]).max()
    np.testing.assert_allclose(o.get("U_VISC")[Ru], mu0 / p.Re * gx[Ru], atol=1e-10 * scale)
    np.testing.assert_allclose(o.get("V_VISC")[Rv], mu0 / p.Re * gy[Rv], atol=1e-10 * scale)


def momentum_state(olib, seed, stage=1, dt=2e-4):
    o = make(olib)
    rng = np.random.default_rng(seed)
    o.set("MU", rng.uniform(0.01, 1.0, o.shape))
    o.set("RHO", np.full(o.shape, 0.6))
    o.call("set_time_state", 0.0, dt, dt, stage)
    return o, rng


@pytest.mark.parametrize("comp", ["u", "v"])
def test_implicit_viscous_term_is_the_cg_operator(olib, comp):
    """The implicit viscous term of the right-hand side and the operator the
    momentum CG inverts are the same L_imp: matVec(x) = rho_f/(beta dt) x - L_imp x
    (uniform rho, random variable mu)."""
    o, rng = momentum_state(olib, 43)
    p = o.params
    R = interior(p, comp)
    x = np.zeros(o.shape)
    x[R] = rng.standard_normal(x[R].shape)
    fld, dfl = ("U", "U_D") if comp == "u" else ("V", "V_D")
    o.set(dfl, x)
    o.call("Velocity_update_boundaries", pvof.F[dfl], pvof.PV_VEL_TYPE_NORMAL)
    o.set(fld, o.get(dfl))
    o.call("matVec", pvof.F["SCRATCH0"], pvof.F[dfl], ord(comp))
    o.call("Velocity_cell_center")
    o.call(f"Velocity_{comp}_set_implicit_explicit")
    st = o.stats()
    a = 0.6 / (BETA[1] * st["dt"])
    L = a * x[R] - o.get("SCRATCH0")[R]
    np.testing.assert_allclose(o.get(f"{comp.upper()}_IMPL")[R], L, atol=1e-9 * np.abs(L).max())


@pytest.mark.parametrize("comp", ["u", "v"])
def test_jacobi_diagonal_is_operator_diagonal(olib, comp):
    """The momentum CG's Jacobi diagonal equals matVec(e_ij)_ij away from the
    wall rows (where the mirrored ghost adds to the operator but not to the
    reference's diagonal).  The diagonal is read from the CG's first direction
    d = r / (diag + 1e-14) with x0 = 0, rhs = 1 and no iteration."""
    o, rng = momentum_state(olib, 47)
    o.close()
    p = pvof.config1_params(NXM=32, NYM=24, xmax=1.0, ymax=0.75, CG_MAXIT=0)
    o = oracle_lib.Oracle(p, olib)
    o.set("MU", rng.uniform(0.01, 1.0, o.shape))
    o.set("RHO", rng.uniform(1e-3, 1.0, o.shape))
    o.call("set_time_state", 0.0, 2e-4, 2e-4, 2)
    pp = o.params
    R = interior(pp, comp)
    fld, rhs, dfl = ("U", "U_RHS", "U_D") if comp == "u" else ("V", "V_RHS", "V_D")
    o.set(fld, np.zeros(o.shape))
    o.set(rhs, np.ones(o.shape))
    assert o.call("Velocity_solve_cg", ord(comp)) == 0
    with np.errstate(divide="ignore"):
        diag = 1.0 / o.get(dfl) - 1e-14
    checked = 0
    for _ in range(24):
        i = int(rng.integers(1, pp.NXM + 1)) if comp == "u" else int(rng.integers(0, pp.NXM))
        j = int(rng.integers(1, pp.NYM - 1)) if comp == "u" else int(rng.integers(2, pp.NYM - 1))
        e = np.zeros(o.shape)
        e[j + 3, i + 3] = 1.0
        o.set(dfl, e)
        o.call("Velocity_update_boundaries", pvof.F[dfl], pvof.PV_VEL_TYPE_CG)
        o.call("matVec", pvof.F["SCRATCH0"], pvof.F[dfl], ord(comp))
        a_ii = o.get("SCRATCH0")[j + 3, i + 3]
        assert abs(diag[j + 3, i + 3] - a_ii) < 1e-9 * a_ii, (comp, i, j, diag[j + 3, i + 3], a_ii)
        checked += 1
    assert checked == 24


# ---------------------------------------------------------------- RHS (Velocity.c:1494-1947)
@pytest.mark.parametrize("stage", [0, 1, 2])
def test_momentum_rhs_low_storage_rk3(olib, stage):
    """rhs = rho_f u/(beta dt) + rho_f (gamma E + zeta E_old)/beta + L_imp
    + (gamma V + zeta V_old)/beta - 2 grad p (+ 2 (-dp_dx) in x), with the
    published RK3 coefficients; rho linear so rho_f is rho at the face; then
    E_old <- E and V_old <- V."""
    p = pvof.config1_params(NXM=32, NYM=24, xmax=1.0, ymax=0.75, dp_dx=0.3)
    o = oracle_lib.Oracle(p, olib)
    g = coords(o)
    dt = 3e-4
    o.call("set_time_state", 0.0, dt, dt, stage)
    rng = np.random.default_rng(53 + stage)
    rho = lambda x, y: 0.4 + 0.3 * x + 0.5 * y          # noqa: E731
    o.set("RHO", rho(g["xc"], g["yc"]))
    c1, c2 = 0.7, -1.1
    o.set("P", c1 * g["xc"] + c2 * g["yc"])
    fields = {}
    for name in ("U", "V", "U_EXPL", "U_EXPL_OLD", "U_VISC", "U_VISC_OLD", "U_IMPL",
                 "V_EXPL", "V_EXPL_OLD", "V_VISC", "V_VISC_OLD", "V_IMPL"):
        fields[name] = rng.standard_normal(o.shape)
        o.set(name, fields[name])
    o.call("Velocity_u_set_RHS")
    o.call("Velocity_v_set_RHS")
    b, gm, z = BETA[stage], GAMMA[stage], ZETA[stage]
    for comp, R, rf, grad, src in (("U", interior(o.params, "u"), rho(g["xu"], g["yc"]), c1, -2 * 0.3),
                                   ("V", interior(o.params, "v"), rho(g["xc"], g["yv"]), c2, 0.0)):
        f = {k: v[R] for k, v in fields.items() if k.startswith(comp)}
        expect = (rf[R] * f[comp] / (b * dt) + rf[R] * (gm * f[comp + "_EXPL"] + z * f[comp + "_EXPL_OLD"]) / b +
                  f[comp + "_IMPL"] + (gm * f[comp + "_VISC"] + z * f[comp + "_VISC_OLD"]) / b - 2 * grad + src)
        got = o.get(comp + "_RHS")[R]
        np.testing.assert_allclose(got, expect, atol=1e-12 * np.abs(expect).max(), err_msg=comp)
        np.testing.assert_array_equal(o.get(comp + "_EXPL_OLD")[R], f[comp + "_EXPL"])
        np.testing.assert_array_equal(o.get(comp + "_VISC_OLD")[R], f[comp + "_VISC"])


# ---------------------------------------------------------------- gravity (VOF_SurfaceTension.c:2075-2154)
@pytest.mark.parametrize("grav,rho_kind", [((0.0, -1.0, 0.0), "linear_y"), ((3.0, -4.0, 0.0), "uniform"),
                                           ((0.0, -2.5, 0.0), "uniform")])
def test_gravity_hydrostatic_balance(olib, grav, rho_kind):
    """A fluid at rest in its hydrostatic pressure (grad p = rho Ri g_hat,
    g_hat the unit gravity vector) has a zero momentum right-hand side: the
    gravity force balances the -2 grad p of the right-hand side."""
    Ri = 1.7
    p = pvof.config1_params(NXM=32, NYM=24, xmax=1.0, ymax=0.75, grav=grav, richardson=Ri)
    o = oracle_lib.Oracle(p, olib)
    g = coords(o)
    gh = np.array(grav[:2]) / np.hypot(*grav[:2])
    if rho_kind == "linear_y":
        a, b = 0.9, -0.5
        o.set("RHO", a + b * g["yc"])
        pres = Ri * gh[1] * (a * g["yc"] + b * g["yc"] ** 2 / 2)
    else:
        r0 = 0.45
        o.set("RHO", np.full(o.shape, r0))
        pres = Ri * r0 * (gh[0] * g["xc"] + gh[1] * g["yc"])
    o.set("P", pres)
    for name in ("U", "V", "U_EXPL", "U_EXPL_OLD", "U_VISC", "U_VISC_OLD", "U_IMPL",
                 "V_EXPL", "V_EXPL_OLD", "V_VISC", "V_VISC_OLD", "V_IMPL"):
        o.set(name, np.zeros(o.shape))
    o.call("set_time_state", 0.0, 1e-4, 1e-4, 1)
    o.call("Velocity_u_set_RHS")
    o.call("Velocity_v_set_RHS")
    ru0, rv0 = o.get("U_RHS"), o.get("V_RHS")
    o.call("Velocity_add_gravity_2_RHS")
    Ru, Rv = interior(o.params, "u"), interior(o.params, "v")
    scale = 2 * Ri * 1.0
    if abs(gh[0]) > 0:
        assert np.abs(ru0[Ru]).max() > 0.1 * scale
    assert np.abs(rv0[Rv]).max() > 0.1 * scale
    assert np.abs(o.get("U_RHS")[Ru]).max() < 1e-12 * scale
    assert np.abs(o.get("V_RHS")[Rv]).max() < 1e-12 * scale


# ---------------------------------------------------------------- density, viscosity (VOF_DIFFUSE.c:1453-1488)
def test_density_viscosity_mixture_rule(olib):
    """Pure liquid (C_L = 1) has rho = mu = 1 (scaled by rho1, mu1), pure gas
    has rho2/rho1 and mu2/mu1, and both are linear in the volume fraction."""
    p = pvof.config1_params(NXM=32, NYM=24, xmax=1.0, ymax=0.75, rho2=0.003, mu2=0.07)
    o = oracle_lib.Oracle(p, olib)
    R = regions(o.params)["cell"]
    rng = np.random.default_rng(59)
    c = rng.uniform(0, 1, o.shape)
    c[R][0, :4] = [0.0, 1.0, 0.5, 0.25]
    cc = c.copy()
    cc[3, 3:7] = [0.0, 1.0, 0.5, 0.25]
    o.set("C_L", cc)
    o.call("VOF_DIFFUSE_set_boundary_values", pvof.F["C_L"])
    o.call("VOF_DIFFUSE_update_phase_cache")
    o.call("VOF_update_density_viscosity")
    rho, mu = o.get("RHO")[3, 3:7], o.get("MU")[3, 3:7]
    np.testing.assert_allclose(rho[:2], [0.003, 1.0], rtol=1e-15)
    np.testing.assert_allclose(mu[:2], [0.07, 1.0], rtol=1e-15)
    cl = o.get("C_L")[R]
    np.testing.assert_allclose(o.get("RHO")[R], 0.003 + cl * (1 - 0.003), atol=1e-15)
    np.testing.assert_allclose(o.get("MU")[R], 0.07 + cl * (1 - 0.07), atol=1e-15)


# ---------------------------------------------------------------- explicit CH diffusion (VOF_DIFFUSE.c:998-1016)
@pytest.mark.parametrize("k", [1, 2])
def test_explicit_diffusion_on_a_cosine(olib, k):
    """C = 1/2 + e cos(theta) gives S(C) = C(C-1/2)(C-1) = (3e^3/4 - e/4) cos(theta)
    + (e^3/4) cos(3 theta), a sum of two discrete Laplacian eigenmodes; the
    explicit diffusion is (1/Pe) L S."""
    o = make(olib, NXM=48, NYM=16, ymax=1.0 / 3.0, Pe_CH=37.0)
    p = o.params
    g = coords(o)
    eps = 0.3
    th = 2 * np.pi * k * g["xc"]
    o.set("C_L", 0.5 + eps * np.cos(th))
    o.call("VOF_DIFFUSE_explicit_diffusion", pvof.F["CH_AUX2"])
    lam = lambda kk: -(4 / g["dx"] ** 2) * np.sin(np.pi * kk / p.NXM) ** 2    # noqa: E731
    expect = ((3 * eps**3 / 4 - eps / 4) * lam(k) * np.cos(th) + eps**3 / 4 * lam(3 * k) * np.cos(3 * th)) / p.Pe_CH
    R = regions(p)["cell"]
    np.testing.assert_allclose(o.get("CH_AUX2")[R], expect[R], atol=1e-10 * np.abs(expect).max())


# ---------------------------------------------------------------- CH step (VOF_DIFFUSE.c:1262-1359)
@pytest.mark.parametrize("stage", [0, 1, 2])
def test_ch_step_is_low_storage_rk3(olib, stage):
    """Uniform C = c0 in a flow of uniform divergence D (v = D y) has R^k = -c0 D
    (WENO of a constant is the constant, L S(c0) = 0) and L^2 C = 0, so the step
    is the low-storage RK3 update C <- c0 + dt (gamma R^k + zeta R^{k-1});
    afterwards R^{k-1} <- R^k and the one before is kept."""
    o = make(olib)
    g = coords(o)
    c0, D, Rold, dt = 0.4, 0.8, 0.15, 1e-3
    o.set("C_L", np.full(o.shape, c0))
    o.call("VOF_DIFFUSE_update_phase_cache")
    o.set("U", np.zeros(o.shape))
    o.set("V", D * g["yv"])
    o.set("CH_RHS_N", np.full(o.shape, Rold))
    o.set("CH_RHS_NM1", np.full(o.shape, -7.0))
    o.call("set_time_state", 0.0, dt, dt, stage)
    o.call("VOF_DIFFUSE_step")
    R = regions(o.params)["cell"]
    Rk = -c0 * D
    expect = c0 + dt * (GAMMA[stage] * Rk + ZETA[stage] * Rold)
    np.testing.assert_allclose(o.get("C_L")[R], expect, rtol=1e-13)
    np.testing.assert_allclose(o.get("C_G")[R], 1 - expect, rtol=1e-13)
    np.testing.assert_allclose(o.get("CH_RHS_N")[R], Rk, rtol=1e-12)
    np.testing.assert_allclose(o.get("CH_RHS_NM1")[R], Rold, rtol=0)


@pytest.mark.parametrize("k,m", [(1, 1), (4, 3)])
def test_biharmonic_step_crank_nicolson_factor(olib, k, m):
    """The implicit biharmonic step (a_dt + c L^2) C^k = a_dt C^{k-1} - c L^2 C^{k-1}
    (a_dt = 1/(beta dt), c = Cn^2/Pe) multiplies an eigenmode of L (eigenvalue
    lam) by the trapezoidal factor (a_dt - c lam^2) / (a_dt + c lam^2)."""
    o = make(olib, ch_tol=1e-14, ch_iter_max=500, Cn=0.05, Pe_CH=1.0)
    p = o.params
    g = coords(o)
    e = np.cos(2 * np.pi * k * g["xc"]) * np.cos(np.pi * m * (g["j"][:, None] + 0.5) / p.NYM)
    dt, stage = 0.2, 2
    o.call("set_time_state", 0.0, dt, dt, stage)
    a_dt = 1.0 / (dt * BETA[stage])
    o.set("C_L", e)
    o.set("SCRATCH0", a_dt * e)
    o.call("VOF_DIFFUSE_set_boundary_values", pvof.F["C_L"])
    o.call("VOF_DIFFUSE_set_boundary_values", pvof.F["SCRATCH0"])
    rel = C.c_double()
    o.lib.pvo_VOF_DIFFUSE_solve_implicit_biharmonic(o.h, pvof.F["SCRATCH0"], C.byref(rel))
    lam = -(4 / g["dx"] ** 2) * np.sin(np.pi * k / p.NXM) ** 2 - (4 / g["dy"] ** 2) * np.sin(np.pi * m / (2 * p.NYM)) ** 2
    cc = p.Cn**2 / p.Pe_CH
    fac = (a_dt - cc * lam**2) / (a_dt + cc * lam**2)
    assert abs(fac - 1) > 0.1       # the step changes the mode visibly
    R = regions(p)["cell"]
    np.testing.assert_allclose(o.get("C_L")[R], fac * e[R], atol=1e-11)


# ---------------------------------------------------------------- surface tension (VOF_DIFFUSE.c:1363-1445)
def sigma_coefficient(p):
    """lambda with lambda * Cn^2 * int (dc/dx)^2 dx = 1/We for the equilibrium
    profile c = (1 + tanh(x / (2 sqrt 2 Cn))) / 2 of the double well
    c^2 (1-c)^2 / 4 (its interface energy is Cn^2 int c'^2 by equipartition);
    the integral by quadrature."""
    w = 2 * np.sqrt(2) * p.Cn
    dc = lambda x: 0.5 / w / np.cosh(x / w) ** 2          # noqa: E731
    val, _ = quad(lambda x: dc(x) ** 2, -40 * w, 40 * w, epsabs=0, epsrel=1e-13, limit=400)
    return 1.0 / (p.We * p.Cn**2 * val)


@pytest.mark.parametrize("direction", ["x", "y"])
def test_f_sigma_balances_its_pressure_at_constant_potential(olib, direction):
    """With a constant chemical potential psi0 the potential-form force is a
    pure gradient, 2 lambda psi0 grad C, and the pressure p = lambda psi0 C
    (the Laplace pressure of the interface energy, lambda fixed by the surface
    tension 1/We) cancels it in the momentum right-hand side (-2 grad p + f)."""
    p = pvof.config1_params(NXM=32, NYM=24, xmax=1.0, ymax=0.75, Cn=0.02, We=37.0)
    o = oracle_lib.Oracle(p, olib)
    g = coords(o)
    lam = sigma_coefficient(o.params)
    psi0 = 0.05
    Cf = 0.3 + 0.4 * (g["xc"] if direction == "x" else g["yc"] / 0.75)
    o.set("C_L", Cf)
    o.call("VOF_DIFFUSE_update_phase_cache")
    o.set("PSI_LG", np.full(o.shape, psi0))
    o.call("VOF_DIFFUSE_compute_f_sigma")
    o.call("VOF_swap_f_sigma")
    o.set("P", lam * psi0 * Cf)
    for name in ("U", "V", "U_EXPL", "U_EXPL_OLD", "U_VISC", "U_VISC_OLD", "U_IMPL",
                 "V_EXPL", "V_EXPL_OLD", "V_VISC", "V_VISC_OLD", "V_IMPL"):
        o.set(name, np.zeros(o.shape))
    o.set("RHO", np.ones(o.shape))
    o.call("set_time_state", 0.0, 1e-4, 1e-4, 0)
    o.call("Velocity_u_set_RHS")
    o.call("Velocity_v_set_RHS")
    comp, R = ("U", interior(o.params, "u")) if direction == "x" else ("V", interior(o.params, "v"))
    before = o.get(comp + "_RHS")[R]
    o.call("VOF_apply_f_sigma_old")
    after = o.get(comp + "_RHS")[R]
    assert np.abs(before).max() > 0.1 * lam * psi0
    assert np.abs(after).max() < 1e-12 * np.abs(before).max()


@pytest.mark.parametrize("direction", ["x", "y"])
def test_f_sigma_on_a_tanh_profile(olib, direction):
    """A tanh profile of width W = 2 Cn (not the equilibrium one) has the closed
    forms c' = c(1-c)/(sqrt 2 W) and psi = c(1-c)(1-2c)/2 (1 - (Cn/W)^2); the
    face force is 2 lambda psi c'.  It is zero on faces whose mean liquid or
    gas fraction lies outside [0.005, 0.995] (the band of VOF_DIFFUSE.c:1363-1445)
    and non-zero on every other face where psi does not vanish."""
    n = 256
    if direction == "x":
        p = pvof.config1_params(NXM=n, NYM=8, xmax=1.0, ymax=8.0 / n, Cn=3.0 / n, We=11.0)
    else:
        p = pvof.config1_params(NXM=8, NYM=n, xmax=8.0 / n, ymax=1.0, Cn=3.0 / n, We=11.0)
    o = oracle_lib.Oracle(p, olib)
    g = coords(o)
    pp = o.params
    W = 2 * pp.Cn
    s = (g["xc"] if direction == "x" else g["yc"]) - 0.5
    c = 0.5 * (1 + np.tanh(s / (2 * np.sqrt(2) * W)))
    o.set("C_L", c)
    o.call("VOF_DIFFUSE_set_boundary_values", pvof.F["C_L"])
    o.call("VOF_DIFFUSE_update_phase_cache")
    o.call("VOF_DIFFUSE_compute_psi_LG")
    o.call("VOF_DIFFUSE_compute_f_sigma")
    lam = sigma_coefficient(pp)
    if direction == "x":
        f = o.get("FSX_NEW")[3:3 + pp.NYM, 4:3 + pp.NXM]           # faces i = 1..NXM-1
        sf = (g["xu"] - 0.5)[3:3 + pp.NYM, 4:3 + pp.NXM]
        cl = o.get("C_L")
        cbar = 0.5 * (cl[3:3 + pp.NYM, 4:3 + pp.NXM] + cl[3:3 + pp.NYM, 3:2 + pp.NXM])
    else:
        f = o.get("FSY_NEW")[4:3 + pp.NYM, 3:3 + pp.NXM - 1]       # faces j = 1..NYM-1
        sf = (g["yv"] - 0.5)[4:3 + pp.NYM, 3:3 + pp.NXM - 1]
        cl = o.get("C_L")
        cbar = 0.5 * (cl[4:3 + pp.NYM, 3:3 + pp.NXM - 1] + cl[3:2 + pp.NYM, 3:3 + pp.NXM - 1])
    cf = 0.5 * (1 + np.tanh(sf / (2 * np.sqrt(2) * W)))
    psi = cf * (1 - cf) * (1 - 2 * cf) / 2 * (1 - (pp.Cn / W) ** 2)
    dcf = cf * (1 - cf) / (np.sqrt(2) * W)
    exact = 2 * lam * psi * dcf
    band = (cbar >= 0.005) & (cbar <= 0.995) & (1 - cbar >= 0.005) & (1 - cbar <= 0.995)
    assert band.sum() > 20 and (~band).sum() > 20
    fmax = np.abs(exact).max()
    assert np.all(f[~band] == 0.0)
    # psi vanishes at the interface centre (c = 1/2); every other band face carries force
    assert np.all(f[band & (np.abs(exact) > 1e-3 * fmax)] != 0.0)
    np.testing.assert_allclose(f[band], exact[band], atol=0.01 * fmax)


# ---------------------------------------------------------------- time step (Dtime.c:24-249)
def test_dt_capillary_limit_at_rest(olib):
    """A fluid at rest is limited by the capillary time step of Brackbill et al.
    (1992), dt = cfl sqrt((rho1 + rho2) dx^3 / (4 pi sigma)) with sigma = 1/We."""
    p = pvof.config1_params(NXM=32, NYM=24, xmax=1.0, ymax=0.75, We=150.0, cfl=0.3, max_dt=10.0)
    o = oracle_lib.Oracle(p, olib)
    pp = o.params
    dx = 1.0 / 32
    dts = pp.cfl * np.sqrt((pp.rho1 + pp.rho2) * dx**3 / (4 * np.pi / pp.We))
    o.set("U_BC", np.zeros(o.shape))
    o.set("V_BC", np.zeros(o.shape))
    o.call("set_time_state", 0.5, dts, dts, 0)
    got = o.lib.pvo_Dtime_cfl(o.h)
    assert abs(got - dts) < 1e-14 * dts, (got, dts)


def test_dt_convective_limit(olib):
    """A uniform flow (U0, V0) faster than the capillary limit gives the CFL
    time step cfl / (|U0|/dx + |V0|/dy)."""
    p = pvof.config1_params(NXM=32, NYM=24, xmax=1.0, ymax=0.5, We=1e6, cfl=0.25, max_dt=10.0)
    o = oracle_lib.Oracle(p, olib)
    dx, dy = 1.0 / 32, 1.0 / 48
    U0, V0 = -3.0, 5.0
    o.set("U_BC", np.full(o.shape, U0))
    o.set("V_BC", np.full(o.shape, V0))
    dtc = 0.25 / (abs(U0) / dx + abs(V0) / dy)
    o.call("set_time_state", 0.5, dtc, dtc, 0)
    got = o.lib.pvo_Dtime_cfl(o.h)
    assert abs(got - dtc) < 1e-14 * dtc, (got, dtc)


# ---------------------------------------------------------------- solves and the pressure right-hand side
@pytest.mark.parametrize("comp", ["u", "v"])
def test_velocity_cg_solves_its_system(olib, comp):
    """Velocity_solve_cg (msolve_cg.c:525-722) returns x with A x = rhs to its
    RMS tolerance, A the operator of matVec.  (Its EPS = 1e-14 in the step
    length stalls it once <d, A d> falls to that size, so the tolerance is an
    absolute RMS the right-hand side's scale allows.)"""
    p = pvof.config1_params(NXM=32, NYM=24, xmax=1.0, ymax=0.75, CG_ETOL=1e-8, CG_MAXIT=2000)
    o = oracle_lib.Oracle(p, olib)
    rng = np.random.default_rng(61)
    o.set("MU", rng.uniform(0.01, 1.0, o.shape))
    o.set("RHO", rng.uniform(1e-3, 1.0, o.shape))
    o.call("set_time_state", 0.0, 2e-3, 2e-3, 1)
    R = interior(o.params, comp)
    fld, rhs, dfl = ("U", "U_RHS", "U_D") if comp == "u" else ("V", "V_RHS", "V_D")
    b = np.zeros(o.shape)
    b[R] = 100.0 * rng.standard_normal(b[R].shape)
    o.set(rhs, b)
    o.set(fld, np.zeros(o.shape))
    it = o.call("Velocity_solve_cg", ord(comp))
    assert 0 < it < 2000
    o.set(dfl, o.get(fld))
    o.call("Velocity_update_boundaries", pvof.F[dfl], pvof.PV_VEL_TYPE_CG)
    o.call("matVec", pvof.F["SCRATCH0"], pvof.F[dfl], ord(comp))
    res = o.get("SCRATCH0")[R] - b[R]
    assert np.sqrt((res**2).mean()) < 2e-8


@pytest.mark.parametrize("comp", ["u", "v"])
def test_velocity_cg_exact_steps(olib, comp):
    """The momentum CG with its step lengths taken without EPS (the library's
    pv_set_momentum_cg_steps(1) variant) meets the same stop test on the
    operator of matVec, checked by an independent residual, in a few
    iterations; with the reference's EPS = 1e-14 (msolve_cg.c:674, :705) the
    same system needs more than ten times as many, and on a small right-hand
    side it stalls at CG_MAXIT above the tolerance.  With dt = 10 the
    operator is nearly the viscous one (condition number ~ grid size squared):
    there conjugate directions matter, and the count stays at CG's (steepest
    descent needs about seven times as many)."""
    out = {}
    for exact in (0, 1):
        for scale, tol, dt in ((100.0, 1e-8, 2e-3), (1e-2, 1e-10, 2e-3), (100.0, 1e-8, 10.0)):
            if dt > 1.0 and not exact:
                continue
            p = pvof.config1_params(NXM=32, NYM=24, xmax=1.0, ymax=0.75, CG_ETOL=tol, CG_MAXIT=2000)
            o = oracle_lib.Oracle(p, olib)
            o.lib.pvo_set_momentum_cg_steps(o.h, exact)
            rng = np.random.default_rng(61)
            o.set("MU", rng.uniform(0.01, 1.0, o.shape))
            o.set("RHO", rng.uniform(1e-3, 1.0, o.shape))
            o.call("set_time_state", 0.0, dt, dt, 1)
            R = interior(o.params, comp)
            fld, rhs, dfl = ("U", "U_RHS", "U_D") if comp == "u" else ("V", "V_RHS", "V_D")
            b = np.zeros(o.shape)
            b[R] = scale * rng.standard_normal(b[R].shape)
            o.set(rhs, b)
            o.set(fld, np.zeros(o.shape))
            it = o.call("Velocity_solve_cg", ord(comp))
            o.set(dfl, o.get(fld))
            o.call("Velocity_update_boundaries", pvof.F[dfl], pvof.PV_VEL_TYPE_CG)
            o.call("matVec", pvof.F["SCRATCH0"], pvof.F[dfl], ord(comp))
            rms = float(np.sqrt(((o.get("SCRATCH0")[R] - b[R]) ** 2).mean()))
            out[exact, scale, dt] = (it, rms, tol)
    for scale in (100.0, 1e-2):
        it, rms, tol = out[1, scale, 2e-3]
        assert it <= 10 and rms < tol, (scale, it, rms)
        assert out[0, scale, 2e-3][0] > 10 * it, (scale, out[0, scale, 2e-3])
    it, rms, tol = out[0, 1e-2, 2e-3]
    assert it == 2000 and rms > tol
    it, rms, tol = out[1, 100.0, 10.0]
    assert it <= 120 and rms < tol, (it, rms)


def test_pressure_rhs_is_scaled_divergence(olib):
    """Pressure_set_RHS (Pressure.c:100-196) is the cell divergence of the
    predicted velocity over 2 beta dt (alpha dt of the substep), and
    Pressure_compute_velocity_divergence returns its largest magnitude."""
    o = make(olib)
    g = coords(o)
    rng = np.random.default_rng(67)
    o.set("U", rng.standard_normal(o.shape))
    o.set("V", rng.standard_normal(o.shape))
    dt, stage = 7e-4, 2
    o.call("set_time_state", 0.0, dt, dt, stage)
    o.call("Pressure_set_RHS")
    u, v = o.get("U"), o.get("V")
    R = regions(o.params)["cell"]
    D = (u[:-1, 1:] - u[:-1, :-1])[R] / g["dx"] + (v[1:, :-1] - v[:-1, :-1])[R] / g["dy"]
    np.testing.assert_allclose(o.get("P_RHS")[R], D / (2 * BETA[stage] * dt), rtol=1e-12,
                               atol=1e-12 * np.abs(D).max() / dt)
    assert abs(o.lib.pvo_Pressure_compute_velocity_divergence(o.h) - np.abs(D).max()) < 1e-12 * np.abs(D).max()


@pytest.mark.parametrize("solver", [pvof.PV_PSOLVE_JACOBI_PCG, pvof.PV_PSOLVE_FFT_PCG])
def test_pressure_solve_satisfies_poisson(olib, solver):
    """Both pressure solves return phi with div(grad phi / rho_f) = rhs for a
    right-hand side of zero mean (what a divergence sums to; the constant is the
    null space of the Neumann/periodic operator), phi of zero mean (the gauge of
    psolve_cg.c:255-279)."""
    p = pvof.config1_params(NXM=32, NYM=24, xmax=1.0, ymax=0.75, pressure_solver=solver, P_CG_ETOL=1e-9,
                            P_CG_MAXIT=5000, p_rel_tol=1e-12, p_max_iter=5000)
    o = oracle_lib.Oracle(p, olib)
    rng = np.random.default_rng(71)
    rho = np.zeros(o.shape)
    R = regions(o.params)
    rho[R["G"]] = rng.uniform(0.05, 1.0, (o.params.NYM + 1, o.params.NXM + 1))
    o.set("RHO", rho)
    o.call("VOF_DIFFUSE_set_boundary_values", pvof.F["RHO"])
    b = np.zeros(o.shape)
    b[R["cell"]] = rng.standard_normal((o.params.NYM, o.params.NXM))
    b[R["cell"]] -= b[R["cell"]].mean()
    o.set("P_RHS", b)
    o.set("PHI", np.zeros(o.shape))
    o.call("Pressure_compute_preconditioner")
    it = o.call("Pressure_solve_cg")
    assert 0 < it < 5000
    phi = o.get("PHI")[R["cell"]]
    assert abs(phi.mean()) < 1e-13 * np.abs(phi).max()
    f = np.zeros(o.shape)
    f[R["cell"]] = phi
    o.set("SCRATCH0", f)
    o.call("Pressure_apply_BCs", pvof.F["SCRATCH0"])
    o.call("Pressure_operator_variableCoeff", pvof.F["SCRATCH1"], pvof.F["SCRATCH0"])
    bb = b[R["cell"]]
    assert np.abs(o.get("SCRATCH1")[R["cell"]] - bb).max() < 1e-6 * np.abs(bb).max()


def test_init_from_F_bootstrap(olib):
    """pvo_init_from_F (Cart3d.c:1094-1130, VOF_DIFFUSE.c:524-568): C_L is F
    clamped to [0, 1], C_G = 1 - C_L, rho and mu follow them, the lagged
    surface-tension force starts equal to the new one, the flow is at rest and
    the time state is reset."""
    p = pvof.config1_params(NXM=32, NYM=64)
    o = oracle_lib.Oracle(p, olib)
    rng = np.random.default_rng(73)
    F = rng.uniform(-0.2, 1.2, o.shape)
    o.init_from_F(F)
    R = regions(o.params)["cell"]
    np.testing.assert_array_equal(o.get("C_L")[R], np.clip(F[R], 0, 1))
    np.testing.assert_array_equal(o.get("C_G")[R], 1 - np.clip(F[R], 0, 1))
    np.testing.assert_allclose(o.get("RHO")[R], 1e-3 + np.clip(F[R], 0, 1) * (1 - 1e-3), atol=1e-15)
    np.testing.assert_array_equal(o.get("FSX_OLD"), o.get("FSX_NEW"))
    np.testing.assert_array_equal(o.get("FSY_OLD"), o.get("FSY_NEW"))
    assert np.abs(o.get("FSX_NEW")).max() > 0
    for f in ("U", "V", "P"):
        assert not o.get(f).any(), f
    st = o.stats()
    assert st["time"] == 0.0 and st["dt"] == p.default_dt and st["which_stage"] == 0


# ---------------------------------------------------------------- clamp accounting (VOF_DIFFUSE.c:1093-1127)
def test_clamp_sum_counts_the_periodic_image_and_the_ghost_row(olib):
    """The reference sums the clamp's correction over its G range, which holds
    the periodic image column i = NXM (a copy of column 0) and the ghost row
    j = NYM above the top wall (a copy of row NYM-1).  So the reported
    correction exceeds the change of the interior volume by exactly the
    corrections of column 0 and row NYM-1 once more (the corner cell three
    times more); the interior sum is the change of the interior volume."""
    o = make(olib)
    p = o.params
    g = coords(o)
    rng = np.random.default_rng(79)
    o.set("C_L", rng.uniform(-0.3, 1.3, o.shape))
    o.call("VOF_DIFFUSE_set_boundary_values", pvof.F["C_L"])
    before = o.get("C_L")
    dG = o.lib.pvo_VOF_DIFFUSE_bound_liquid_fraction(o.h)
    dI = o.lib.pvo_clamp_interior(o.h)
    meas = g["dx"] * g["dy"] * p.twod_slab_thickness
    d = (np.clip(before, 0, 1) - before) * meas
    nx, ny = p.NXM, p.NYM
    interior_ = d[3:3 + ny, 3:3 + nx].sum()
    col0, row_top, corner = d[3:3 + ny, 3].sum(), d[2 + ny, 3:3 + nx].sum(), d[2 + ny, 3]
    vol_change = (o.get("C_L") - before)[3:3 + ny, 3:3 + nx].sum() * meas
    assert abs(dI - interior_) < 1e-15 * nx * ny
    assert abs(dI - vol_change) < 1e-15 * nx * ny
    assert abs(dG - (interior_ + col0 + row_top + corner)) < 1e-15 * nx * ny
    assert abs(dG - dI) > 1e-3 * abs(dI)


# ---------------------------------------------------------------------------
# ghost exchange (Communication.c:128-295) and VOF_DIFFUSE_init (VOF_DIFFUSE.c:524-568)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("component", ["c", "a", "u", "d", "p", "q"])
@pytest.mark.parametrize("pnodes", [1, 2, 3])
def test_ghost_exchange_is_the_periodic_extension(olib, component, pnodes):
    """One rank that is its own x neighbour: after the exchange every ghost
    column it fills holds the periodic image of its column (period NXM in
    the index: cells 0..NXM-1, and for u-type data faces 1..NXM with face 0
    the image of face NXM), on rows -pnodes .. NYM+pnodes; everything else is
    untouched."""
    o = make(olib)
    p = o.params
    nx = p.NXM
    rng = np.random.default_rng(7 + pnodes)
    a = rng.standard_normal(o.shape)
    o.set("SCRATCH0", a)
    o.call("Communication_update_ghost_nodes_flow_variable", pvof.F["SCRATCH0"], ord(component), pnodes)
    b = o.get("SCRATCH0")
    i = np.arange(nx + 7) - 3
    ulike = component in "udep"
    lo = 1 if ulike else 0                        # first owned column
    own = (i >= lo) & (i < lo + nx)
    filled_lo = (i < lo) & (i >= lo - (pnodes + 1 if ulike else pnodes))
    filled_hi = (i >= lo + nx) & (i < lo + nx + pnodes)
    rows = slice(3 - pnodes, 3 + p.NYM + 1 + pnodes)
    src_col = lo + np.mod(i - lo, nx) + 3          # canonical column of the owned image
    expect = a.copy()
    for c in np.nonzero(filled_lo | filled_hi)[0]:
        expect[rows, c] = a[rows, src_col[c]]
    assert np.array_equal(b, expect)
    assert np.array_equal(b[:, own.nonzero()[0]], a[:, own.nonzero()[0]])


def test_vof_diffuse_init_clamps_and_completes_the_phases(olib):
    """VOF_DIFFUSE_init: C_L in [0, 1] everywhere, equal to F where F is
    already in [0, 1] inside the domain, C_L + C_G = 1 (no solid) over the
    L range, zero-gradient wall ghosts and periodic ghost columns."""
    o = make(olib)
    p = o.params
    rng = np.random.default_rng(3)
    F = rng.uniform(-0.3, 1.3, size=o.shape)
    o.set("F", F)
    o.call("VOF_DIFFUSE_init")
    CL, CG = o.get("C_L"), o.get("C_G")
    R = regions(p)["cell"]
    assert CL.min() >= 0.0 and CL.max() <= 1.0
    inside = (F[R] >= 0) & (F[R] <= 1)
    assert np.array_equal(CL[R][inside], F[R][inside])
    assert np.array_equal(CL[R][F[R] < 0], np.zeros((F[R] < 0).sum()))
    assert np.array_equal(CL[R][F[R] > 1], np.ones((F[R] > 1).sum()))
    frame = regions(p)["frame"]
    assert np.allclose(CL[frame] + CG[frame], 1.0, rtol=0, atol=1e-15)
    nx, ny = p.NXM, p.NYM
    for g in (1, 2, 3):                             # walls: zero gradient
        assert np.array_equal(CL[3 - g, 3:3 + nx], CL[3, 3:3 + nx])
    assert np.array_equal(CL[3 + ny - 1 + 1, 3:3 + nx], CL[3 + ny - 1, 3:3 + nx])
    assert np.array_equal(CL[3:3 + ny, 3 + nx], CL[3:3 + ny, 3])     # periodic image column
    assert np.array_equal(CL[3:3 + ny, 2], CL[3:3 + ny, 3 + nx - 1])
