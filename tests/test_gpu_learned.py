"""Learned pressure solve: what the trained network gains over the method
without it (ablation), on the GPU.

The network is trained on the GPU by the project's own kernels on crops of
the plain FFT-PCG solves of one seeded bubble field, then evaluated on a
held-out seed: the same RK3 steps are run from the same state with each
pressure solver with and without the network (FFT-PCG / learned warm start /
learned flexible CG, and MG-PCG / learned warm start + MG-PCG).  All solve to
the same relative residual, so the pressure fields agree to the solve
tolerance; the test records the pressure iterations of each, and that a warm
start is kept only when its residual energy is below the cold start's.
"""
import ctypes as C
import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402
import pvof  # noqa: E402
from helpers import regions, relerr  # noqa: E402

pytestmark = pytest.mark.gpu

N = 256


@pytest.fixture(scope="module")
def trained(gpu):
    lib = pvof.load_lib()
    p = bench.workload_params(N)
    F = pvof.ic(lib, "bubbles", p, seed=111, nmin=16, nmax=64, rmin=0.01, rmax=0.08, gas_inside=1)
    s = pvof.Solver(p, lib)
    s.init_from_F(F)
    net = pvof.Network(s, seed=3)
    net.collect([0, 8, 32, 96], crops_per_snap=48, crop_size=64, max_samples=20000, seed=5)
    s.call("Temporal_int_rk3", 6)
    nsamp = net.collect_end()
    l0, l1 = net.train(1500, batch=32, lr=1e-3, seed=1)
    _, l2 = net.train(800, batch=32, lr=3e-4, seed=2)
    print(f"\nsamples {nsamp}  loss {l0:.3e} -> {l2:.3e}")
    return lib, p, net.params(), nsamp, l0, l2


def run_mode(lib, p, prm, solver, steps=3):
    F = pvof.ic(lib, "bubbles", p, seed=222, nmin=16, nmax=64, rmin=0.01, rmax=0.08, gas_inside=1)
    s = pvof.Solver(p, lib)
    s.init_from_F(F)
    s.call("Temporal_int_rk3", 2)
    if solver not in (pvof.PV_PSOLVE_FFT_PCG, pvof.PV_PSOLVE_MG_PCG):
        pvof.Network(s).set_params(prm)
    s.check(lib.pv_set_pressure_solver(s.h, solver))
    iters, starts, fb = [], [], 0
    for _ in range(steps):
        s.call("Temporal_int_rk3", 1)
        st = s.stats()
        iters.append(sum(st["n_p"]))
        fb += sum(st["p_fallback"])
        assert max(st["p_res"]) <= p.p_rel_tol
        r = C.c_double(0.0)
        lib.pv_learn_start_residual(s.h, C.byref(r))
        starts.append(r.value)
    return s, iters, starts, fb


def test_training_reduces_loss(trained):
    _, _, _, nsamp, l0, l2 = trained
    assert nsamp >= 1000
    assert l2 < 0.5 * l0


def test_learned_modes_vs_plain(trained):
    lib, p, prm, _, _, _ = trained
    base, it0, _, _ = run_mode(lib, p, prm, pvof.PV_PSOLVE_FFT_PCG)
    _, itm, _, _ = run_mode(lib, p, prm, pvof.PV_PSOLVE_MG_PCG)
    R = regions(p)
    report = [f"plain FFT-PCG iterations per step {it0}", f"MG-PCG iterations per step {itm}"]
    for name, solver in (("learned warm start", pvof.PV_PSOLVE_LEARNED),
                         ("learned FCG", pvof.PV_PSOLVE_LEARNED_FCG),
                         ("learned warm start + MG-PCG", pvof.PV_PSOLVE_LEARNED_MG)):
        s, it, starts, fb = run_mode(lib, p, prm, solver)
        report.append(f"{name}: iterations {it}, start residuals {[round(x, 4) for x in starts]}, "
                      f"fallbacks {fb}")
        # same tolerance, same pressure (the iterates differ below the solve tolerance)
        assert relerr(s.get("P")[R["cell"]], base.get("P")[R["cell"]]) < 1e-5
        accepted = [x for x in starts if x < 1.0]
        assert all(0.0 < x < 1.0 for x in accepted)
    print("\n".join(report))
