"""Cahn-Hilliard implicit solve started from the exact spectral solution
(pv_set_ch_solver(bag, 1)) on the GPU.

On the uniform grid the CH operator a_dt + (Cn^2/Pe) L^2 (periodic x, the
CH ghost rules in y) is diagonal in the FFT-x / DCT-y basis, so the spectral
start is the solution up to rounding and the reference CG's own stop test
(relative weighted residual < ch_tol) accepts it at iteration 0.  Trajectories
agree with the plain CG ones to the CG tolerance.
"""
import numpy as np
import pytest

import pvof
from helpers import regions, relerr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib(gpu):
    return pvof.load_lib()


def params(n, precision):
    return pvof.default_params(
        NXM=n, NYM=n, xmax=1.0, ymax=1.0, Re=100.0, We=200.0, rho1=1.0, rho2=1e-3, mu1=1.0, mu2=1e-2,
        Cn=2.0 / n, Pe_CH=n / 2.0, cfl=0.2, default_dt=1e-4, max_dt=1e-3, ch_tol=1e-12, CG_ETOL=1e-7,
        CG_MAXIT=120, pressure_solver=pvof.PV_PSOLVE_FFT_PCG, p_rel_tol=1e-10, p_max_iter=3000,
        precision=precision)


@pytest.mark.parametrize("precision", [pvof.PV_PREC_FP64, pvof.PV_PREC_MIXED])
def test_spectral_start_matches_cg(lib, precision):
    n = 128
    p = params(n, precision)
    F = pvof.ic(lib, "bubbles", p, seed=31, nmin=8, nmax=24, rmin=0.02, rmax=0.1, gas_inside=1)
    out = {}
    for spec in (0, 1):
        s = pvof.Solver(p, lib)
        s.init_from_F(F)
        s.check(lib.pv_set_ch_solver(s.h, spec))
        n_ch = []
        for _ in range(3):
            s.call("Temporal_int_rk3", 1)
            st = s.stats()
            n_ch += list(st["n_ch"])
            assert max(st["ch_res"]) < p.ch_tol
        out[spec] = (s, n_ch)
    print(f"\nbiharmonic CG iterations per substep: from C {out[0][1]}, from the spectral solve {out[1][1]}")
    assert max(out[1][1]) <= 1
    R = regions(p)
    # C_L to well inside the CH tolerance; u, v, p to the momentum CG's stop
    # (RMS residual CG_ETOL = 1e-7): the two runs' momentum solves stop at
    # different iterates once their inputs differ at round-off
    for f, tol in (("C_L", 1e-9), ("U", 5e-7), ("V", 5e-7), ("P", 5e-7)):
        e = relerr(out[1][0].get(f)[R["cell"]], out[0][0].get(f)[R["cell"]])
        print(f"  {f}: {e:.2e}")
        assert e < tol, f
