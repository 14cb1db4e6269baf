"""The kernel timers' working-launch readout (pv_timers_read_working): of
the timed launches it keeps those that ran at least a quarter of the class's
longest, so the MG-PCG's CG launches queued past convergence (they exit on
the device stop flag) drop out of the average that the bandwidth figures
divide by.  Checked on a short MG-PCG run against the all-launch readout:
a subset (counts and milliseconds no larger), averaging no less, and for a
class whose launches include such exits, strictly fewer launches.
"""
import ctypes as C

import pytest

import pvof

pytestmark = pytest.mark.gpu

NT = 16          # PV_NTIMERS


def test_working_launches_are_a_subset(gpu):
    lib = pvof.load_lib()
    n = 128
    p = pvof.default_params(NXM=n, NYM=n, xmax=1.0, ymax=1.0, Re=100.0, We=200.0, rho1=1.0, rho2=1e-3, mu1=1.0,
                            mu2=1e-2, Cn=2.0 / n, Pe_CH=n / 2.0, cfl=0.2, default_dt=1e-4, max_dt=1e-3,
                            ch_tol=1e-12, CG_ETOL=1e-9, CG_MAXIT=200, pressure_solver=pvof.PV_PSOLVE_MG_PCG,
                            p_rel_tol=1e-10, p_max_iter=3000, precision=pvof.PV_PREC_MIXED)
    F = pvof.ic(lib, "bubbles", p, seed=71, nmin=4, nmax=8, rmin=0.05, rmax=0.12, gas_inside=1)
    s = pvof.Solver(p, lib)
    s.init_from_F(F)
    s.call("timers_enable", 1)
    s.call("timers_reset")
    s.call("Temporal_int_rk3", 3)
    ms, cnt = (C.c_double * NT)(), (C.c_longlong * NT)()
    wms, wcnt = (C.c_double * NT)(), (C.c_longlong * NT)()
    s.check(lib.pv_timers_read(s.h, ms, cnt))
    s.check(lib.pv_timers_read_working(s.h, wms, wcnt))
    fewer = 0
    for t in range(NT):
        assert 0 <= wcnt[t] <= cnt[t] and wms[t] <= ms[t] * (1 + 1e-12), t
        if cnt[t]:
            assert wcnt[t] >= 1, t               # the longest launch itself is working
            assert wms[t] / wcnt[t] >= ms[t] / cnt[t] * (1 - 1e-12), t
            fewer += wcnt[t] < cnt[t]
    assert fewer >= 1                            # some CG class had launches that exited at once
    s.call("timers_enable", 0)
    wcnt2 = (C.c_longlong * NT)()
    s.check(lib.pv_timers_read_working(s.h, None, wcnt2))
    assert all(v == 0 for v in wcnt2)            # no samples kept once the timers are off
    s.close()
