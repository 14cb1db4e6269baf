"""Mixed-precision substep (precision = PV_PREC_MIXED) against the FP64 oracle.

The mixed path forms the WENO5 face states, the explicit chemical potential
(S, its Laplacian, psi_LG and the capillary force) and the momentum stencils
(convective/viscous terms and the momentum CG operator) in FP32.  The
conservative phase-field flux accumulation, the biharmonic and pressure
solves and every inner product stay FP64.  The tests state the per-field
error bound of the mixed path against the oracle, check that the FP64 path
run side by side still meets its 1e-11 parity bound, and that phase volume is
conserved to FP64 round-off.
"""
import numpy as np
import pytest

import oracle_lib
import pvof
from helpers import regions, relerr

pytestmark = pytest.mark.gpu

# per-field bounds of the mixed path after 10 RK3 steps (relative L-inf)
MIXED_BOUND = {"C_L": 1e-6, "U": 1e-4, "V": 1e-4, "P": 1e-4}
FIELDS = (("C_L", "cell"), ("U", "u"), ("V", "v"), ("P", "cell"))


def params(precision, **kw):
    base = dict(ch_tol=1e-13, CG_ETOL=1e-7, CG_MAXIT=400, pressure_solver=pvof.PV_PSOLVE_FFT_PCG,
                p_rel_tol=1e-13, precision=precision)
    base.update(kw)
    return pvof.config1_params(NXM=32, NYM=64, **base)


@pytest.fixture(scope="module")
def lib():
    return pvof.load_lib()


def l2rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_mixed_per_field_error_vs_oracle(gpu, lib):
    p64 = params(pvof.PV_PREC_FP64)
    F = pvof.ic(lib, "rising_bubble", p64)
    o = oracle_lib.Oracle(p64)
    o.init_from_F(F)
    assert o.call("Temporal_int_rk3", 10) == 0
    R = regions(p64)
    runs = {}
    for prec in (pvof.PV_PREC_FP64, pvof.PV_PREC_MIXED):
        s = pvof.Solver(params(prec), lib)
        s.init_from_F(F)
        s.call("Temporal_int_rk3", 10)
        runs[prec] = s
    report = []
    for fld, r in FIELDS:
        ref = o.get(fld)[R[r]]
        e64 = relerr(runs[pvof.PV_PREC_FP64].get(fld)[R[r]], ref)
        em = relerr(runs[pvof.PV_PREC_MIXED].get(fld)[R[r]], ref)
        em2 = l2rel(runs[pvof.PV_PREC_MIXED].get(fld)[R[r]], ref)
        report.append(f"{fld}: fp64 Linf {e64:.2e}  mixed Linf {em:.2e} L2 {em2:.2e}")
        assert e64 < 1e-11, fld
        assert em < MIXED_BOUND[fld], (fld, em)
    print("\n".join(report))


def test_mixed_conserves_phase_volume(gpu, lib):
    """Per step, the change of sum(C_L) equals the clamp correction to FP64
    round-off: the FP32 stencils enter only through FP64 face fluxes."""
    p = params(pvof.PV_PREC_MIXED, ch_tol=1e-14, ch_iter_max=1000)
    F = pvof.ic(lib, "rising_bubble", p)
    s = pvof.Solver(p, lib)
    s.init_from_F(F)
    R = regions(s.params)
    meas = (s.params.xmax - s.params.xmin) / p.NXM * (s.params.ymax - s.params.ymin) / p.NYM * \
        s.params.twod_slab_thickness
    for _ in range(3):
        m0 = s.get("C_L")[R["cell"]].sum() * meas
        s.call("Temporal_int_rk3", 1)
        st = s.stats()
        m1 = s.get("C_L")[R["cell"]].sum() * meas
        assert abs((m1 - m0) - sum(st["clamp_delta"])) < 1e-13 * m0
        assert max(st["div_max"]) <= 1e-6
