"""The reference-signature layer on the GPU: routines called by their
reference names and prototypes, on reference-layout double*** arrays bound
to a device subdomain, give what the oracle gives for the same routine.

The arrays follow Memory.c:36-96 / :270-330 (helpers.RefArray); values move
between them and the device only through pvs_push / pvs_pull, as a PARTIES
build would do at init, resume and output."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib
import pvof
from helpers import RefArray, copy_state, oracle_state, regions, relerr

pytestmark = pytest.mark.gpu


def params_small(**kw):
    # CH and pressure solves converged far below the compared tolerance; the
    # momentum CG stops at CG_ETOL = 1e-7 after a few iterations (below about
    # 1e-8 its EPS = 1e-14 guards, msolve_cg.c:525-722, stall it on this small
    # grid until CG_MAXIT, where GPU and oracle iterates drift apart by their
    # summation order), so both stop at the same iterate
    base = dict(ch_tol=1e-13, CG_ETOL=1e-7, CG_MAXIT=400, pressure_solver=pvof.PV_PSOLVE_FFT_PCG,
                p_rel_tol=1e-13, p_max_iter=3000)
    base.update(kw)
    return pvof.config1_params(NXM=32, NYM=64, **base)


class Bag:
    """Stand-ins for the reference's Cart3d_bag and its members (only their
    addresses matter to the shim), bound to one device subdomain."""

    def __init__(self, shim, lib, p):
        self.shim = shim
        self.s = pvof.Solver(p, lib)
        self.mem = {k: C.create_string_buffer(64) for k in ("db", "grid", "params", "p", "u", "v")}
        self.ptr = {k: C.addressof(b) for k, b in self.mem.items()}
        assert shim.pvs_bind(self.ptr["db"], self.s.h, self.ptr["grid"], self.ptr["params"], self.ptr["p"],
                             self.ptr["u"], self.ptr["v"]) == 0, shim.pvs_last_error()
        self.arrays = {}

    def array(self, name, ghost=True):
        if name not in self.arrays:
            a = RefArray(self.s.params.NXM, self.s.params.NYM, ghost)
            rng = pvof.PVS_L_RANGE if ghost else pvof.PVS_G_RANGE
            assert self.shim.pvs_bind_field(self.ptr["db"], a.ptr, pvof.F[name], rng) == 0, \
                self.shim.pvs_last_error()
            self.arrays[name] = a
        return self.arrays[name]

    def pull(self, name):
        a = self.arrays[name]
        assert self.shim.pvs_pull(self.ptr["db"], a.ptr) == 0, self.shim.pvs_last_error()
        return a

    def close(self):
        self.shim.pvs_unbind(self.ptr["db"])
        self.s.close()


@pytest.fixture(scope="module")
def lib(gpu):
    return pvof.load_lib()


@pytest.fixture(scope="module")
def shim(lib):
    return pvof.load_shim()


@pytest.fixture(scope="module")
def state(lib):
    p = params_small()
    F = pvof.ic(lib, "rising_bubble", p)
    return p, oracle_state(p, F, 3)


def fresh(state, lib, shim, stage=1):
    p, base = state
    o = oracle_lib.Oracle(p)
    for name in pvof.FIELDS:
        o.set(name, base.get(name))
    bag = Bag(shim, lib, p)
    copy_state(base, bag.s, stage)
    st = base.stats()
    o.call("set_time_state", st["time"], st["dt"], st["dt_old"], stage)
    return p, o, bag


def test_push_pull_round_trip(state, lib, shim):
    p, o, bag = fresh(state, lib, shim)
    a = bag.array("C_L")
    g = bag.array("U_EXPL_OLD", ghost=False)
    bag.pull("C_L")
    bag.pull("U_EXPL_OLD")
    for k in range(-3, 5):                  # every slab holds the device's k = 0 values
        assert np.array_equal(a.canonical(k), o.get("C_L")), k
    R = regions(p)["G"]
    assert np.array_equal(g.canonical(1)[R], o.get("U_EXPL_OLD")[R])
    new = np.random.default_rng(1).uniform(0, 1, size=bag.s.shape)
    a.set_canonical(new)
    assert shim.pvs_push(bag.ptr["db"], a.ptr) == 0
    assert np.array_equal(bag.s.get("C_L"), new)
    bag.close()


def test_weno_through_reference_signature(state, lib, shim):
    p, o, bag = fresh(state, lib, shim)
    rhs = bag.array("CH_AUX2")
    shim.VOF_DIFFUSE_advect_WENO5(bag.ptr["db"], rhs.ptr)
    o.call("VOF_DIFFUSE_advect_WENO5", pvof.F["CH_AUX2"])
    bag.pull("CH_AUX2")
    R = regions(p)["frame"]
    assert relerr(rhs.canonical(0)[R], o.get("CH_AUX2")[R]) < 1e-13
    bag.close()


def test_routines_by_reference_name(state, lib, shim):
    """Laplacian, ghost exchange, velocity boundaries, pressure BCs (found by
    its grid pointer), the pressure operator and the divergence."""
    p, o, bag = fresh(state, lib, shim)
    db = bag.ptr["db"]
    cl, lap = bag.array("C_L"), bag.array("SCRATCH0")
    shim.VOF_DIFFUSE_compute_laplacian(cl.ptr, lap.ptr, db)
    o.call("VOF_DIFFUSE_compute_laplacian", pvof.F["C_L"], pvof.F["SCRATCH0"])
    R = regions(p)["frame"]
    assert np.array_equal(bag.pull("SCRATCH0").canonical()[R], o.get("SCRATCH0")[R])

    u = bag.array("U")
    shim.Communication_update_ghost_nodes_flow_variable(u.ptr, b"d", 3, db)      # U_VELOCITY
    o.call("Communication_update_ghost_nodes_flow_variable", pvof.F["U"], ord("d"), 3)
    assert np.array_equal(bag.pull("U").canonical(), o.get("U"))

    v = bag.array("V")
    shim.Velocity_update_boundaries(v.ptr, b"v", pvof.PV_VEL_TYPE_NORMAL, db)
    o.call("Velocity_update_boundaries", pvof.F["V"], pvof.PV_VEL_TYPE_NORMAL)
    assert np.array_equal(bag.pull("V").canonical()[R], o.get("V")[R])

    phi = bag.array("PHI")
    shim.Pressure_apply_BCs(phi.ptr, bag.ptr["grid"], None)
    o.call("Pressure_apply_BCs", pvof.F["PHI"])
    assert np.array_equal(bag.pull("PHI").canonical()[R], o.get("PHI")[R])

    aphi, rho = bag.array("P_AD"), bag.array("RHO")
    shim.Pressure_operator_variableCoeff(aphi.ptr, phi.ptr, rho.ptr, bag.ptr["grid"], bag.ptr["params"], db)
    o.call("Pressure_operator_variableCoeff", pvof.F["P_AD"], pvof.F["PHI"])
    Rc = regions(p)["cell"]
    assert np.array_equal(bag.pull("P_AD").canonical()[Rc], o.get("P_AD")[Rc])

    d = shim.Pressure_compute_velocity_divergence(db)
    assert d == o.call("Pressure_compute_velocity_divergence")
    bag.close()


def test_substep_and_rk3_loop_through_the_shim(state, lib, shim):
    """A whole substep by its reference name matches the oracle's; then the
    reference's time loop runs until time_max (whole steps only)."""
    p, o, bag = fresh(state, lib, shim, stage=0)
    db = bag.ptr["db"]
    names = ("C_L", "U", "V", "P")
    arrs = {n: bag.array(n) for n in names}
    shim.Temporal_int_all_the_equations(db, None)
    # the RK3 loop forms the cell-centre velocities before stage 0
    # (Temporal_int.c:399-448); the stored ones are a stage old.  The device
    # forms them on the fly inside its momentum kernel.
    o.call("Velocity_cell_center")
    assert o.call("Temporal_int_all_the_equations") == 0
    so, sg = o.stats(), bag.s.stats()
    for k in ("n_u", "n_v"):
        assert sg[k][0] == so[k][0] < p.CG_MAXIT, (k, sg[k], so[k])
    R = regions(p)["cell"]
    for n in names:
        bag.pull(n)
        assert relerr(arrs[n].canonical()[R], o.get(n)[R]) < 1e-11, n
    st0 = bag.s.stats()
    dt = st0["dt"]
    # the loop (Temporal_int.c:270) runs while time - time_max < dt: from
    # time 0 with time_max = 1.5 dt and dt changing by at most 5 % a step,
    # that is three steps
    bag.s.call("set_time_state", 0.0, dt, dt, 0)
    assert shim.pvs_set_time_max(db, 1.5 * dt) == 0
    assert shim.Temporal_int_rk3(db, None) == 0
    st = bag.s.stats()
    assert st["ntime"] - st0["ntime"] == 3
    assert st["time"] - 1.5 * dt >= st["dt"]
    for n in names:
        assert np.isfinite(bag.pull(n).canonical()[R]).all()
    bag.close()


def test_solvers_by_reference_name(state, lib, shim):
    p, o, bag = fresh(state, lib, shim, stage=2)
    db = bag.ptr["db"]
    shim.Velocity_cell_center(db)
    o.call("Velocity_cell_center")
    for c in ("u", "v"):
        o.call(f"Velocity_{c}_set_implicit_explicit")
        o.call(f"Velocity_{c}_set_RHS")
        getattr(shim, f"Velocity_{c}_set_implicit_explicit")(db)
        getattr(shim, f"Velocity_{c}_set_RHS")(db)
    its_u = shim.Velocity_solve_cg(bag.ptr["u"], db)
    its_v = shim.Velocity_solve_cg(bag.ptr["v"], db)
    ou = o.call("Velocity_solve_cg", ord("u"))
    ov = o.call("Velocity_solve_cg", ord("v"))
    assert its_u > 0 and its_v > 0 and abs(its_u - ou) <= 1 and abs(its_v - ov) <= 1
    R = regions(p)
    assert relerr(bag.s.get("U")[R["u"]], o.get("U")[R["u"]]) < 1e-10
    assert relerr(bag.s.get("V")[R["v"]], o.get("V")[R["v"]]) < 1e-10
    shim.Pressure_set_RHS(db)
    o.call("Pressure_set_RHS")
    rhs = bag.array("P_RHS")
    phi = bag.array("PHI")
    ok = shim.Pressure_solve_x_period(bag.ptr["p"], bag.ptr["grid"], bag.ptr["params"])
    o.call("Pressure_solve_fft", pvof.F["P_RHS"], pvof.F["PHI"])
    assert ok == 0
    assert relerr(bag.pull("PHI").canonical()[R["cell"]], o.get("PHI")[R["cell"]]) < 1e-10
    bag.close()


def test_fatal_path_on_a_bound_bag(state, lib, shim):
    """An array that is not bound, a Velocity that is neither u nor v, and a
    rho that is not the bag's density go to the fatal handler; the bag keeps
    working afterwards."""
    p, o, bag = fresh(state, lib, shim)
    seen = []

    @pvof.FATAL_FN
    def handler(msg, ctx):
        seen.append(msg.decode())

    shim.pvs_set_fatal(handler, None)
    try:
        db = bag.ptr["db"]
        stray = RefArray(p.NXM, p.NYM)
        shim.VOF_DIFFUSE_explicit_diffusion(db, stray.ptr)
        assert shim.Velocity_solve_cg(bag.ptr["params"], db) == -1
        phi, lap = bag.array("PHI"), bag.array("SCRATCH1")
        shim.Pressure_operator_variableCoeff(lap.ptr, phi.ptr, phi.ptr, None, None, db)
        assert len(seen) == 3, seen
        assert "not bound" in seen[0] and "neither" in seen[1] and "density" in seen[2]
        shim.Temporal_int_all_the_equations(db, None)
        assert len(seen) == 3
    finally:
        shim.pvs_set_fatal(pvof.FATAL_FN(), None)
    bag.close()
