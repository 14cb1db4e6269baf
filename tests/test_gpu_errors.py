"""Error behaviour of the C ABI with a device present.

Every rejected call returns PV_EINVAL (-1) with a message from
pv_last_error, launches nothing, and leaves the handle usable: an RK3 step
on it afterwards meets the pressure tolerance.  Size limits: MG-PCG needs a
grid that halves to <= 4096 cells and on to a last level of <= 256 cells
(rejected at pv_create), FFT-PCG needs NXM = 2^a 3^b, a multiple of 16
(rejected by the pressure solve, the only routine that needs the transform).
"""
import ctypes as C

import numpy as np
import pytest

import pvof

pytestmark = pytest.mark.gpu

PV_EINVAL = -1


@pytest.fixture(scope="module")
def lib(gpu):
    return pvof.load_lib()


def params(n, solver):
    return pvof.default_params(
        NXM=n, NYM=n, xmax=1.0, ymax=1.0, Re=100.0, We=200.0, rho1=1.0, rho2=1e-3, mu1=1.0, mu2=1e-2,
        Cn=2.0 / n, Pe_CH=n / 2.0, cfl=0.2, default_dt=1e-4, max_dt=1e-3, ch_tol=1e-12, CG_ETOL=1e-7,
        CG_MAXIT=120, pressure_solver=solver, p_rel_tol=1e-10, p_max_iter=3000)


def message(lib):
    return lib.pv_last_error().decode()


def test_mg_pcg_rejects_grid_that_does_not_halve(lib):
    """98 x 98 halves once, to 49 x 49 (2401 cells): no last level of <= 256 cells."""
    h = C.c_void_p()
    p = params(98, pvof.PV_PSOLVE_MG_PCG)
    assert lib.pv_create(C.byref(p), 0, C.byref(h)) == PV_EINVAL
    assert not h.value
    assert "does not halve" in message(lib)
    # the same grid with FFT-PCG is created, and asking it for the MG cycle fails
    s = pvof.Solver(params(98, pvof.PV_PSOLVE_FFT_PCG), lib)
    assert lib.pv_set_mg_cycle(s.h, 1) == PV_EINVAL
    assert "not available" in message(lib)
    s.close()


@pytest.mark.parametrize("n", [98, 100, 40])
def test_fft_pcg_rejects_unsupported_row_lengths(lib, n):
    """98 = 2 x 7^2 and 100 = 2^2 x 5^2 have other prime factors; 40 = 2^3 x 5
    also; the transforms take NXM = 2^a 3^b with 16 | NXM."""
    s = pvof.Solver(params(n, pvof.PV_PSOLVE_FFT_PCG), lib)
    it = C.c_int(-7)
    assert lib.pv_Pressure_solve_cg(s.h, C.byref(it)) == PV_EINVAL
    assert "2^a 3^b" in message(lib)
    assert it.value == -7
    s.close()


def test_bad_arguments_leave_handle_usable(lib):
    n = 64
    p = params(n, pvof.PV_PSOLVE_MG_PCG)
    s = pvof.Solver(p, lib)
    buf = np.zeros(s.shape)
    bad = [
        (lambda: lib.pv_set_mg_cycle(s.h, 9), "W levels"),
        (lambda: lib.pv_set_mg_sweeps(s.h, 3), "sweeps"),
        (lambda: lib.pv_set_mg_precision(s.h, 2), "precision"),
        (lambda: lib.pv_Velocity_solve_cg(s.h, ord("w"), None), "component"),
        (lambda: lib.pv_get_field(s.h, 999, buf.ctypes.data_as(C.POINTER(C.c_double))), "field"),
    ]
    for call, word in bad:
        assert call() == PV_EINVAL
        assert word in message(lib), message(lib)
    F = pvof.ic(lib, "bubbles", p, seed=31, nmin=4, nmax=8, rmin=0.05, rmax=0.12, gas_inside=1)
    s.init_from_F(F)
    s.call("Temporal_int_rk3", 1)
    st = s.stats()
    assert max(st["p_res"]) <= 1e-10
    assert np.isfinite(s.get("P")).all()
    s.close()
