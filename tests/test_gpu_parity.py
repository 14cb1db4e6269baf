"""GPU parity: each HIP entry point against the CPU oracle on the same state.

Stencil kernels are compared bit for bit (both sides are built without FMA
contraction and follow the reference's operand order), except WENO5, whose
smoothness indicators use pow(x, 2.0) in the oracle (glibc, not always
correctly rounded) and x*x on the GPU: 1e-13 relative there.  Iterative
solves are compared on their converged solutions.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import oracle_lib
import pvof
from helpers import copy_state, oracle_state, regions, relerr

pytestmark = pytest.mark.gpu


def params_small(**kw):
    base = dict(ch_tol=1e-13, CG_ETOL=1e-7, CG_MAXIT=400, P_CG_ETOL=1e-12, P_CG_MAXIT=6000)
    base.update(kw)
    return pvof.config1_params(NXM=32, NYM=64, **base)


@pytest.fixture(scope="module")
def lib():
    return pvof.load_lib()


@pytest.fixture(scope="module")
def state(gpu, lib):
    p = params_small()
    F = pvof.ic(lib, "rising_bubble", p)
    o = oracle_state(p, F, 3)
    return p, F, o


def fresh(state, lib, stage=1):
    p, F, o_base = state
    o = oracle_lib.Oracle(p)
    for name in pvof.FIELDS:
        o.set(name, o_base.get(name))
    st = o_base.stats()
    o.call("set_time_state", st["time"], st["dt"], st["dt_old"], stage)
    s = pvof.Solver(p, lib)
    copy_state(o, s, stage)
    return p, o, s


def test_init_matches(gpu, lib):
    p = params_small()
    F = pvof.ic(lib, "rising_bubble", p)
    o = oracle_lib.Oracle(p)
    o.init_from_F(F)
    s = pvof.Solver(p, lib)
    s.init_from_F(F)
    R = regions(p)
    for name in ("C_L", "C_G", "RHO", "MU"):
        assert np.array_equal(s.get(name)[R["frame"]], o.get(name)[R["frame"]]), name
    for name in ("FSX_OLD", "FSY_OLD"):
        assert np.array_equal(s.get(name)[R["G"]], o.get(name)[R["G"]]), name


@pytest.mark.parametrize("stage", [0, 1, 2])
def test_weno_advect(state, lib, stage):
    p, o, s = fresh(state, lib, stage)
    R = regions(p)
    o.call("VOF_DIFFUSE_advect_WENO5", pvof.F["CH_AUX2"])
    s.call("VOF_DIFFUSE_advect_WENO5", pvof.F["CH_AUX2"])
    assert relerr(s.get("CH_AUX2")[R["frame"]], o.get("CH_AUX2")[R["frame"]]) < 1e-13


def test_explicit_diffusion(state, lib):
    p, o, s = fresh(state, lib)
    R = regions(p)
    o.call("VOF_DIFFUSE_explicit_diffusion", pvof.F["CH_AUX1"])
    s.call("VOF_DIFFUSE_explicit_diffusion", pvof.F["CH_AUX1"])
    assert np.array_equal(s.get("CH_AUX1")[R["frame"]], o.get("CH_AUX1")[R["frame"]])


def test_laplacian_and_psi(state, lib):
    p, o, s = fresh(state, lib)
    R = regions(p)
    o.call("VOF_DIFFUSE_compute_laplacian", pvof.F["C_L"], pvof.F["SCRATCH0"])
    s.call("VOF_DIFFUSE_compute_laplacian", pvof.F["C_L"], pvof.F["SCRATCH0"])
    assert np.array_equal(s.get("SCRATCH0")[R["frame"]], o.get("SCRATCH0")[R["frame"]])
    o.call("VOF_DIFFUSE_compute_psi_LG")
    s.call("VOF_DIFFUSE_compute_psi_LG")
    assert np.array_equal(s.get("PSI_LG")[R["frame"]], o.get("PSI_LG")[R["frame"]])


def test_f_sigma(state, lib):
    p, o, s = fresh(state, lib)
    R = regions(p)
    o.call("VOF_DIFFUSE_compute_psi_LG")
    o.call("VOF_DIFFUSE_compute_f_sigma")
    s.call("VOF_DIFFUSE_compute_f_sigma")
    assert np.array_equal(s.get("FSX_NEW")[R["u"]], o.get("FSX_NEW")[R["u"]])
    assert np.array_equal(s.get("FSY_NEW")[R["v"]], o.get("FSY_NEW")[R["v"]])
    assert np.abs(o.get("FSX_NEW")[R["u"]]).max() > 0


@pytest.mark.parametrize("stage", [0, 1, 2])
def test_momentum_terms_and_rhs(state, lib, stage):
    p, o, s = fresh(state, lib, stage)
    R = regions(p)
    # the substep refreshes the cell-centre velocity first (Temporal_int.c:625-628)
    o.call("Velocity_cell_center")
    s.call("Velocity_cell_center")
    for c in ("u", "v"):
        o.call(f"Velocity_{c}_set_implicit_explicit")
        s.call(f"Velocity_{c}_set_implicit_explicit")
        C_ = c.upper()
        for fld in ("EXPL", "IMPL", "VISC"):
            assert np.array_equal(s.get(f"{C_}_{fld}")[R[c]], o.get(f"{C_}_{fld}")[R[c]]), (c, fld)
        o.call(f"Velocity_{c}_set_RHS")
        s.call(f"Velocity_{c}_set_RHS")
        for fld in ("RHS", "EXPL_OLD", "VISC_OLD"):
            assert np.array_equal(s.get(f"{C_}_{fld}")[R[c]], o.get(f"{C_}_{fld}")[R[c]]), (c, fld)
    o.call("VOF_apply_f_sigma_old")
    s.call("VOF_apply_f_sigma_old")
    o.call("Velocity_add_gravity_2_RHS")
    s.call("Velocity_add_gravity_2_RHS")
    assert np.array_equal(s.get("U_RHS")[R["u"]], o.get("U_RHS")[R["u"]])
    assert np.array_equal(s.get("V_RHS")[R["v"]], o.get("V_RHS")[R["v"]])


def test_matvec(state, lib):
    p, o, s = fresh(state, lib)
    R = regions(p)
    for c, fld in (("u", "U"), ("v", "V")):
        o.call("matVec", pvof.F["SCRATCH1"], pvof.F[fld], ord(c))
        s.call("matVec", pvof.F["SCRATCH1"], pvof.F[fld], ord(c))
        assert np.array_equal(s.get("SCRATCH1")[R[c]], o.get("SCRATCH1")[R[c]]), c


def test_pressure_rhs_precond_operator(state, lib):
    p, o, s = fresh(state, lib)
    R = regions(p)
    for name in ("Pressure_set_RHS", "Pressure_compute_preconditioner"):
        o.call(name)
        s.call(name)
    assert np.array_equal(s.get("P_RHS")[R["cell"]], o.get("P_RHS")[R["cell"]])
    assert np.array_equal(s.get("M_INV")[R["cell"]], o.get("M_INV")[R["cell"]])
    o.call("Pressure_apply_BCs", pvof.F["P"])
    s.call("Pressure_apply_BCs", pvof.F["P"])
    o.call("Pressure_operator_variableCoeff", pvof.F["SCRATCH0"], pvof.F["P"])
    s.call("Pressure_operator_variableCoeff", pvof.F["SCRATCH0"], pvof.F["P"])
    assert np.array_equal(s.get("SCRATCH0")[R["cell"]], o.get("SCRATCH0")[R["cell"]])


def test_cfl(state, lib):
    p, o, s = fresh(state, lib, 0)
    st = o.stats()
    o.call("set_time_state", st["time"], st["dt"], st["dt"], 0)
    dt_o = o.lib.pvo_Dtime_cfl(o.h)
    out = C.c_double()
    s.call("Dtime_cfl", C.byref(out))
    assert out.value == dt_o


def test_biharmonic_solve(state, lib):
    p, o, s = fresh(state, lib, 1)
    R = regions(p)
    o.call("VOF_DIFFUSE_step")
    s.call("VOF_DIFFUSE_step")
    assert relerr(s.get("C_L")[R["frame"]], o.get("C_L")[R["frame"]]) < 1e-12
    assert relerr(s.get("CH_RHS_N")[R["cell"]], o.get("CH_RHS_N")[R["cell"]]) < 1e-12
    so, ss = o.stats(), s.stats()
    assert abs(so["n_ch"][1] - ss["n_ch"][1]) <= 1


def test_velocity_solve(state, lib):
    p, o, s = fresh(state, lib, 2)
    R = regions(p)
    o.call("Velocity_cell_center")
    s.call("Velocity_cell_center")
    for c in ("u", "v"):
        o.call(f"Velocity_{c}_set_implicit_explicit")
        o.call(f"Velocity_{c}_set_RHS")
        s.call(f"Velocity_{c}_set_RHS")
        o.call("Velocity_solve_cg", ord(c))
        it = C.c_int()
        s.call("Velocity_solve_cg", ord(c), C.byref(it))
        fld = c.upper()
        assert relerr(s.get(fld)[R[c]], o.get(fld)[R[c]]) < 1e-10, c


def test_pressure_solve_project(state, lib):
    p, o, s = fresh(state, lib, 1)
    R = regions(p)
    for x in (o, s):
        x.set("PHI", np.zeros(x.shape))
        x.call("Pressure_set_RHS")
        x.call("Pressure_compute_preconditioner")
    o.call("Pressure_solve_cg")
    it = C.c_int()
    s.call("Pressure_solve_cg", C.byref(it))
    assert relerr(s.get("PHI")[R["cell"]], o.get("PHI")[R["cell"]]) < 1e-9
    o.call("Pressure_project_velocity")
    s.call("Pressure_project_velocity")
    for fld, r in (("U", "u"), ("V", "v"), ("P", "cell")):
        assert relerr(s.get(fld)[R[r]], o.get(fld)[R[r]]) < 1e-9, fld


def test_fft_solve(state, lib):
    p, o, s = fresh(state, lib)
    R = regions(p)
    rng = np.random.default_rng(5)
    rhs = np.zeros(o.shape)
    rhs[R["cell"]] = rng.standard_normal((p.NYM, p.NXM))
    rhs[R["cell"]] -= rhs[R["cell"]].mean()
    o.set("SCRATCH0", rhs)
    s.set("SCRATCH0", rhs)
    o.call("Pressure_solve_fft", pvof.F["SCRATCH0"], pvof.F["SCRATCH1"])
    s.call("Pressure_solve_fft", pvof.F["SCRATCH0"], pvof.F["SCRATCH1"])
    assert relerr(s.get("SCRATCH1")[R["cell"]], o.get("SCRATCH1")[R["cell"]]) < 1e-11


@pytest.mark.parametrize("solver", [pvof.PV_PSOLVE_JACOBI_PCG, pvof.PV_PSOLVE_FFT_PCG])
def test_trajectory_10_steps(gpu, lib, solver):
    p = params_small(pressure_solver=solver, p_rel_tol=1e-13)
    F = pvof.ic(lib, "rising_bubble", p)
    o = oracle_lib.Oracle(p)
    o.init_from_F(F)
    s = pvof.Solver(p, lib)
    s.init_from_F(F)
    assert o.call("Temporal_int_rk3", 10) == 0
    s.call("Temporal_int_rk3", 10)
    R = regions(p)
    so, ss = o.stats(), s.stats()
    assert abs(so["time"] - ss["time"]) <= 1e-15 * so["time"]
    for fld, r in (("C_L", "cell"), ("U", "u"), ("V", "v"), ("P", "cell")):
        assert relerr(s.get(fld)[R[r]], o.get(fld)[R[r]]) < 1e-11, fld


def test_trajectory_momentum_exact_steps(gpu, lib):
    """The momentum CG without the reference's EPS in its step lengths
    (pv_set_momentum_cg_steps(1)) against the oracle's same variant: 5 RK3
    steps of the rising bubble at a CG_ETOL (1e-9) where the reference's
    steps stall, with the same iteration counts per solve."""
    p = params_small(pressure_solver=pvof.PV_PSOLVE_FFT_PCG, p_rel_tol=1e-13, CG_ETOL=1e-9)
    F = pvof.ic(lib, "rising_bubble", p)
    o = oracle_lib.Oracle(p)
    o.init_from_F(F)
    o.lib.pvo_set_momentum_cg_steps(o.h, 1)
    s = pvof.Solver(p, lib)
    s.init_from_F(F)
    s.check(lib.pv_set_momentum_cg_steps(s.h, 1))
    for _ in range(5):
        assert o.call("Temporal_int_rk3", 1) == 0
        s.call("Temporal_int_rk3", 1)
        so, ss = o.stats(), s.stats()
        for k in ("n_u", "n_v"):
            assert ss[k] == so[k] and max(so[k]) < p.CG_MAXIT, (k, ss[k], so[k])
    R = regions(p)
    for fld, r in (("C_L", "cell"), ("U", "u"), ("V", "v"), ("P", "cell")):
        assert relerr(s.get(fld)[R[r]], o.get(fld)[R[r]]) < 1e-11, fld
    with pytest.raises(RuntimeError, match="momentum CG steps"):
        s.check(lib.pv_set_momentum_cg_steps(s.h, 2))


@pytest.mark.parametrize("n,m", [(48, 96), (96, 80)])
def test_trajectory_fft_pcg_mixed_radix(gpu, lib, n, m):
    """FFT-PCG on grids the power-of-two transforms cannot take: rows of
    48 = 2^4 x 3 and 96 = 2^5 x 3 (radix-3 passes) and 80 columns (five
    segments of the y solve); 3 RK3 steps of the rising bubble against the
    oracle's FFT-PCG (its DFT in x, Thomas in y)."""
    p = pvof.config1_params(NXM=n, NYM=m, ymax=m / n,
                            ch_tol=1e-13, CG_ETOL=1e-7, CG_MAXIT=400, pressure_solver=pvof.PV_PSOLVE_FFT_PCG,
                            p_rel_tol=1e-13)
    F = pvof.ic(lib, "rising_bubble", p)
    o = oracle_lib.Oracle(p)
    o.init_from_F(F)
    s = pvof.Solver(p, lib)
    s.init_from_F(F)
    assert o.call("Temporal_int_rk3", 3) == 0
    s.call("Temporal_int_rk3", 3)
    R = regions(p)
    for fld, r in (("C_L", "cell"), ("U", "u"), ("V", "v"), ("P", "cell")):
        assert relerr(s.get(fld)[R[r]], o.get(fld)[R[r]]) < 1e-10, fld
    s.close()


def test_config1_200_steps(gpu, lib):
    """The plain FP64 path over the whole small plain run: 128 x 256, one gas
    bubble of radius 0.25, 200 RK3 steps (FFT-PCG to 1e-13, CH and momentum
    solves converged far below the bound), against the oracle's state stored
    in tests/golden/config1_200_steps.npz by tests/golden/make_config1_200.py
    (the oracle needs about 11 minutes for it): every interior value of C_L,
    u, v and p, and the maximum and L2 norm of each field, to relative 1e-11
    (L-infinity over the field, scaled by its maximum); every step changes
    sum(C_L) by the clamp correction only, to the round-off of the sum
    (N eps); the time agrees to round-off."""
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.join(here, "golden"))
    import make_config1_200 as mk
    g = np.load(os.path.join(here, "golden", "config1_200_steps.npz"))
    p = mk.config1()
    s = pvof.Solver(p, lib)
    s.init_from_F(mk.initial_field(lib, p))
    R = regions(s.params)
    meas = (s.params.xmax - s.params.xmin) / p.NXM * (s.params.ymax - s.params.ymin) / p.NYM * \
        s.params.twod_slab_thickness
    vol0 = s.get("C_L")[R["cell"]].sum()
    assert abs(vol0 - g["vol0"]) <= 1e-13 * abs(g["vol0"])
    # round-off bound of an N-term sum of positive values: N eps
    bound, drift = p.NXM * p.NYM * np.finfo(float).eps, 0.0
    for _ in range(int(g["steps"])):
        m0 = s.get("C_L")[R["cell"]].sum() * meas
        s.call("Temporal_int_rk3", 1)
        m1 = s.get("C_L")[R["cell"]].sum() * meas
        drift = max(drift, abs((m1 - m0) - sum(s.stats()["clamp_delta"])) / m0)
        assert drift < bound, drift
    assert abs(s.stats()["time"] - g["time"]) <= 1e-13 * g["time"]
    report = []
    for f, r in mk.FIELDS:
        a = s.get(f)[R[r]]
        assert g[f].shape == a.shape, (f, g[f].shape, a.shape)
        e_inf = np.abs(a - g[f]).max() / g[f + "_max"]
        e_max = abs(np.abs(a).max() - g[f + "_max"]) / g[f + "_max"]
        e_l2 = abs(np.sqrt((a * a).sum()) - g[f + "_l2"]) / g[f + "_l2"]
        report.append(f"{f}: Linf {e_inf:.1e}, max {e_max:.1e}, L2 {e_l2:.1e}")
        assert max(e_inf, e_max, e_l2) < 1e-11, f
    print("\n" + "\n".join(report) + f"\nlargest per-step phase-volume change beyond the clamp: {drift:.1e} "
          f"(bound {bound:.1e})")


def test_config2_10_steps(gpu, lib):
    """The plain FP64 path at the scale configuration's size (512 x 1024
    square cells, 8-mode Rayleigh-Taylor interface, Atwood 0.5, Re 3000)
    against the oracle's state after 10 RK3 steps stored in
    tests/golden/config2_10_steps.npz by tests/golden/make_config2_10.py (the
    oracle needs about 4 minutes for it): every 16th cell of C_L, u, v and p,
    and the maximum and L2 norm of each full field, to relative 1e-11; the
    time agrees to round-off."""
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.join(here, "golden"))
    import make_config2_10 as mk
    g = np.load(os.path.join(here, "golden", "config2_10_steps.npz"))
    p = mk.config2()
    s = pvof.Solver(p, lib)
    s.init_from_F(mk.initial_field(lib, p))
    s.call("Temporal_int_rk3", int(g["steps"]))
    R = regions(s.params)
    assert abs(s.stats()["time"] - g["time"]) <= 1e-13 * g["time"]
    report = []
    for f, r in mk.FIELDS:
        a = s.get(f)[R[r]]
        e_sub = np.abs(a[::mk.STRIDE, ::mk.STRIDE] - g[f + "_sub"]).max() / g[f + "_max"]
        e_max = abs(np.abs(a).max() - g[f + "_max"]) / g[f + "_max"]
        e_l2 = abs(np.sqrt((a * a).sum()) - g[f + "_l2"]) / g[f + "_l2"]
        report.append(f"{f}: sampled Linf {e_sub:.1e}, max {e_max:.1e}, L2 {e_l2:.1e}")
        assert max(e_sub, e_max, e_l2) < 1e-11, f
    print("\n" + "\n".join(report))


@pytest.mark.parametrize("n,m", [(64, 32), (512, 256), (1024, 512), (96, 96), (192, 96), (384, 768), (96, 100),
                                 (48, 17), (144, 1000)])
def test_fft_solve_sizes(gpu, lib, n, m):
    """Standalone FFT-x / tridiagonal-y solve against the oracle's FFT + Thomas
    solve: power-of-two rows, rows of 2^a 3^b (radix-3 passes), columns of any
    length (the partitioned Thomas solve: one segment up to 16 rows, ragged
    last segments)."""
    p = pvof.default_params(NXM=n, NYM=m, ymax=m / n)
    o = oracle_lib.Oracle(p)
    s = pvof.Solver(p, lib)
    R = regions(p)
    rng = np.random.default_rng(n)
    rhs = np.zeros(o.shape)
    rhs[R["cell"]] = rng.standard_normal((p.NYM, p.NXM))
    rhs[R["cell"]] -= rhs[R["cell"]].mean()
    for x in (o, s):
        x.set("SCRATCH0", rhs)
        x.call("Pressure_solve_fft", pvof.F["SCRATCH0"], pvof.F["SCRATCH1"])
    got, ref = s.get("SCRATCH1")[R["cell"]], o.get("SCRATCH1")[R["cell"]]
    assert np.isfinite(got).all()
    assert relerr(got, ref) < 1e-10


@pytest.mark.parametrize("component", ["c", "u", "p", "f"])
@pytest.mark.parametrize("pnodes", [1, 2, 3])
def test_ghost_exchange(gpu, lib, component, pnodes):
    p = params_small()
    o = oracle_lib.Oracle(p)
    s = pvof.Solver(p, lib)
    a = np.random.default_rng(pnodes).standard_normal(s.shape)
    o.set("SCRATCH1", a)
    s.set("SCRATCH1", a)
    o.call("Communication_update_ghost_nodes_flow_variable", pvof.F["SCRATCH1"], ord(component), pnodes)
    s.call("Communication_update_ghost_nodes_flow_variable", pvof.F["SCRATCH1"], ord(component), pnodes)
    assert np.array_equal(s.get("SCRATCH1"), o.get("SCRATCH1"))
    s.close()


def test_vof_diffuse_init(gpu, lib):
    p = params_small()
    o = oracle_lib.Oracle(p)
    s = pvof.Solver(p, lib)
    F = np.random.default_rng(11).uniform(-0.3, 1.3, size=s.shape)
    o.set("F", F)
    s.set("F", F)
    o.call("VOF_DIFFUSE_init")
    s.call("VOF_DIFFUSE_init")
    # over the frame every stencil reads, as everywhere in these tests (the
    # outermost ghost row and column i = NXM+3, j = NYM+3 are not written by
    # the device's ghost kernels)
    fr = regions(p)["frame"]
    for name in ("C_L", "C_G"):
        assert np.array_equal(s.get(name)[fr], o.get(name)[fr]), name
    assert np.array_equal(s.get("F")[fr], o.get("C_L")[fr])
    s.close()
