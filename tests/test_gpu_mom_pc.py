"""The momentum CG's polynomial preconditioner (pv_set_momentum_preconditioner
1: z = D^-1 (r + N D^-1 r) for A = D - N, the Jacobi splitting) on the GPU.

The preconditioner changes the path of the CG, not the system it solves or
its stop test (the RMS residual against CG_ETOL), so it is checked against
the Jacobi-preconditioned CG of the same handle kind:

* a routine-level momentum solve to a tight tolerance gives the Jacobi
  solve's velocity to that tolerance, in fewer iterations, and leaves the
  final residual (which the preconditioned CG keeps in two alternating
  buffers) in the public U_R / V_R field: its RMS is below CG_ETOL, as the
  stop test saw it;
* whole RK3 steps of the fast form agree with Jacobi to the solves'
  tolerance with fewer momentum iterations;
* the option is refused on FP64 handles (it reads the FP32 operator
  coefficients) and for unknown kinds.
"""
import ctypes as C

import numpy as np
import pytest

import pvof
from helpers import regions, relerr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib(gpu):
    return pvof.load_lib()


def params(nx, ny, **kw):
    base = dict(NXM=nx, NYM=ny, xmax=1.0, ymax=ny / nx, Re=100.0, We=200.0, rho1=1.0, rho2=1e-3, mu1=1.0,
                mu2=1e-2, Cn=2.0 / nx, Pe_CH=nx / 2.0, cfl=0.2, default_dt=1e-4, max_dt=1e-3, ch_tol=1e-12,
                CG_ETOL=1e-9, CG_MAXIT=200, pressure_solver=pvof.PV_PSOLVE_MG_PCG, p_rel_tol=1e-10,
                p_max_iter=3000, precision=pvof.PV_PREC_MIXED)
    base.update(kw)
    return pvof.default_params(**base)


def stepped(lib, p, F, nsteps, pc, exact=1):
    s = pvof.Solver(p, lib)
    s.init_from_F(F)
    s.check(lib.pv_set_momentum_cg_steps(s.h, exact))
    s.check(lib.pv_set_momentum_preconditioner(s.h, pc))
    if nsteps:
        s.call("Temporal_int_rk3", nsteps)
    return s


def test_refused_on_fp64_and_unknown_kinds(lib):
    p = params(64, 64, precision=pvof.PV_PREC_FP64)
    s = pvof.Solver(p, lib)
    with pytest.raises(RuntimeError, match="mixed-precision"):
        s.check(lib.pv_set_momentum_preconditioner(s.h, 1))
    s.check(lib.pv_set_momentum_preconditioner(s.h, 0))
    s.close()
    s = pvof.Solver(params(64, 64), lib)
    for bad in (-1, 2):
        with pytest.raises(RuntimeError, match="momentum preconditioner"):
            s.check(lib.pv_set_momentum_preconditioner(s.h, bad))
    s.close()


def test_solve_matches_jacobi_and_keeps_the_residual(lib):
    """Both CGs from the same state and right-hand sides, exact steps (with
    the reference's EPS in the step lengths neither reaches 1e-12: the
    recurrences stall near rz ~ EPS).  The mixed form's iterations apply the
    FP32 operator, so the two paths agree to FP32 rounding of A d, not to
    the tolerance."""
    n = 128
    p = params(n, 2 * n, CG_ETOL=1e-12, CG_MAXIT=400)
    F = pvof.ic(lib, "bubbles", p, seed=61, nmin=6, nmax=16, rmin=0.03, rmax=0.12, gas_inside=1)
    R = regions(p)
    s = stepped(lib, p, F, 2, 0)                      # a developed state
    s.call("Velocity_cell_center")
    for c in ("u", "v"):
        s.call(f"Velocity_{c}_set_implicit_explicit")
    for c in ("u", "v"):
        s.call(f"Velocity_{c}_set_RHS")
    start = {f: s.get(f).copy() for f in ("U", "V")}
    out = {}
    for pc in (0, 1):
        for f, a in start.items():
            s.set(f, a)
        s.check(lib.pv_set_momentum_preconditioner(s.h, pc))
        its = {}
        for c in ("u", "v"):
            it = C.c_int()
            s.call("Velocity_solve_cg", ord(c), C.byref(it))
            its[c] = it.value
            r = s.get(c.upper() + "_R")[R[c]]
            rms = float(np.sqrt((r * r).mean()))
            assert 0 < its[c] < p.CG_MAXIT, (pc, c, its[c])
            assert rms < p.CG_ETOL, (pc, c, rms)        # the residual the stop test saw, in the public field
        out[pc] = (its, {f: s.get(f).copy() for f in ("U", "V")})
    s.close()
    print(f"\nmomentum CG to 1e-12: Jacobi {out[0][0]}, polynomial {out[1][0]}")
    for c in ("u", "v"):
        assert out[1][0][c] < out[0][0][c], c
        f = c.upper()
        assert relerr(out[1][1][f][R[c]], out[0][1][f][R[c]]) < 1e-6, (c, relerr(out[1][1][f][R[c]], out[0][1][f][R[c]]))


def test_steps_match_jacobi(lib):
    n = 256
    p = params(n, n)
    F = pvof.ic(lib, "bubbles", p, seed=67, nmin=8, nmax=24, rmin=0.02, rmax=0.1, gas_inside=1)
    R = regions(p)
    runs = {}
    for pc in (0, 1):
        s = stepped(lib, p, F, 0, pc)
        its = 0
        for _ in range(4):
            s.call("Temporal_int_rk3", 1)
            st = s.stats()
            assert max(st["p_res"]) <= 1e-10
            assert max(st["n_u"]) < p.CG_MAXIT and max(st["n_v"]) < p.CG_MAXIT
            its += sum(st["n_u"]) + sum(st["n_v"])
        runs[pc] = (s, its)
    print(f"\nmomentum iterations over 4 steps: Jacobi {runs[0][1]}, polynomial {runs[1][1]}")
    assert runs[1][1] < 0.8 * runs[0][1]
    for f in ("U", "V", "P", "C_L"):
        assert relerr(runs[1][0].get(f)[R["cell"]], runs[0][0].get(f)[R["cell"]]) < 1e-6, f
    for s, _ in runs.values():
        s.close()
