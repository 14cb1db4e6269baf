"""The scale, mixed-precision and learned configurations on the GPU.

* 512 x 1024 Rayleigh-Taylor (plain FP64 path at scale), 100 RK3 steps:
  fields finite, divergence after every projection below the outer loop's
  1e-6, and every step's change of the interior phase volume equal to the
  clamp's interior correction to a few ulps of the volume.  The reference's
  own clamp sum (over its G range, with the periodic image column and the
  ghost row) is reported beside it.
* 2048^2 with 32 seeded droplets (Cn 0.004): the fast mixed form (FP32
  stencils and multigrid cycle, FP64 phi fluxes, solves and dots) against the
  FP64 GPU path from the same field, 20 RK3 steps, per-field relative L2 and
  Linf, phase volume.
* 1024^2 bubbly field, learned warm start + MG-PCG: the network is trained on
  the GPU on the plain solver's own solves of a held-out seed; every pressure
  solve meets 1e-10 relative, the fields agree with cold-started MG-PCG to
  the solve tolerance, and the network's gain (iterations, start energy) is
  printed.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402
import pvof  # noqa: E402
from helpers import regions, relerr  # noqa: E402

pytestmark = pytest.mark.gpu

EPS = np.finfo(float).eps


@pytest.fixture(scope="module")
def lib(gpu):
    return pvof.load_lib()


def test_config2_rayleigh_taylor_100_steps(lib):
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.join(here, "golden"))
    import make_config2_10 as mk
    p = mk.config2()
    s = pvof.Solver(p, lib)
    s.init_from_F(mk.initial_field(lib, p))
    pp = s.params
    R = regions(pp)["cell"]
    meas = (pp.xmax - pp.xmin) / pp.NXM * (pp.ymax - pp.ymin) / pp.NYM * pp.twod_slab_thickness
    vol = s.get("C_L")[R].sum() * meas
    vol0 = vol
    worst_int, worst_ref, clamped = 0.0, 0.0, 0
    for step in range(100):
        s.call("Temporal_int_rk3", 1)
        st = s.stats()
        assert max(st["div_max"]) <= 1e-6, (step, st["div_max"])
        new = s.get("C_L")[R].sum() * meas
        d_int = (new - vol) - sum(st["clamp_delta_interior"])
        d_ref = (new - vol) - sum(st["clamp_delta"])
        clamped += any(c != 0.0 for c in st["clamp_delta_interior"])
        worst_int = max(worst_int, abs(d_int) / vol0)
        worst_ref = max(worst_ref, abs(d_ref) / vol0)
        # interior volume = transport (telescoping fluxes) + CH solve (mean mode exact) + clamp
        assert abs(d_int) <= 64 * EPS * vol0, (step, d_int / vol0)
        vol = new
    for f in ("C_L", "U", "V", "P"):
        assert np.isfinite(s.get(f)).all(), f
    print(f"\nconfig 2, 100 steps: {clamped} steps clamped; worst |dV - interior clamp| / V = {worst_int:.2e}, "
          f"against the reference's G-range clamp sum {worst_ref:.2e}; time {s.stats()['time']:.4f}")
    assert clamped > 0
    s.close()


def test_config4_mixed_against_fp64_2048(lib):
    n, steps = 2048, 20
    runs = {}
    for name, prec in (("fp64", pvof.PV_PREC_FP64), ("mixed", pvof.PV_PREC_MIXED)):
        p = bench.workload_params(n, pvof.PV_PSOLVE_MG_PCG, prec)
        F = pvof.ic(lib, "bubbles", p, seed=bench.SEED, nmin=32, nmax=32, rmin=0.02, rmax=0.06, gas_inside=0)
        s = pvof.Solver(p, lib)
        s.init_from_F(F)
        if name == "mixed":          # the fast form: spectral CH start, temporal starts, W cycle,
            s.check(lib.pv_set_ch_solver(s.h, 1))              # momentum CG exact steps, polynomial
            s.check(lib.pv_set_pressure_warm_start(s.h, 1))    # preconditioner
            s.check(lib.pv_set_mg_cycle(s.h, 2))
            s.check(lib.pv_set_momentum_warm_start(s.h, 1))
            s.check(lib.pv_set_momentum_cg_steps(s.h, 1))
            s.check(lib.pv_set_momentum_preconditioner(s.h, 1))
        runs[name] = s
    for step in range(steps):
        for name, s in runs.items():
            s.call("Temporal_int_rk3", 1)
            st = s.stats()
            assert max(st["p_res"]) <= 1e-10, (name, step, st["p_res"])
    R = regions(runs["fp64"].params)["cell"]
    report, bounds = [], {"C_L": 1e-6, "U": 1e-4, "V": 1e-4, "P": 1e-4}
    for f, bound in bounds.items():
        a, b = runs["mixed"].get(f)[R], runs["fp64"].get(f)[R]
        assert np.isfinite(a).all()
        d = a - b
        l2 = float(np.sqrt((d * d).sum() / (b * b).sum()))
        linf = float(np.abs(d).max() / np.abs(b).max())
        report.append(f"{f}: rel L2 {l2:.2e}, rel Linf {linf:.2e} (bound {bound:g})")
        assert l2 <= bound and linf <= bound, f
    vol = [runs[k].get("C_L")[R].sum() for k in ("fp64", "mixed")]
    assert abs(vol[1] - vol[0]) <= 1e-12 * vol[0]
    assert runs["fp64"].stats()["time"] == pytest.approx(runs["mixed"].stats()["time"], rel=1e-4)
    print(f"\n2048^2, 32 droplets, {steps} steps, mixed vs FP64: " + "; ".join(report) +
          f"; phase volume diff {abs(vol[1] - vol[0]) / vol[0]:.1e}")
    for s in runs.values():
        s.close()


def test_config5_learned_mg_1024(lib):
    n = 1024
    prec = pvof.PV_PREC_FP64
    pt = bench.workload_params(n, pvof.PV_PSOLVE_FFT_PCG, prec)
    Ft = pvof.ic(lib, "bubbles", pt, seed=bench.SEED + 1, nmin=16, nmax=64, rmin=0.01, rmax=0.08, gas_inside=1)
    st = pvof.Solver(pt, lib)
    st.init_from_F(Ft)
    net = pvof.Network(st, seed=3)
    net.collect([0, 8, 32, 96], crops_per_snap=300, crop_size=64, max_samples=20000, seed=5)
    st.call("Temporal_int_rk3", 3)
    nsamp = net.collect_end()
    l0, _ = net.train(1500, batch=32, lr=1e-3, seed=11)
    _, l1 = net.train(750, batch=32, lr=3e-4, seed=12)
    prm = net.params()
    st.close()
    assert nsamp > 1000 and l1 < l0
    p = bench.workload_params(n, pvof.PV_PSOLVE_MG_PCG, prec)
    F = pvof.ic(lib, "bubbles", p, seed=bench.SEED, nmin=16, nmax=64, rmin=0.01, rmax=0.08, gas_inside=1)
    out = {}
    for name, solver in (("cold", pvof.PV_PSOLVE_MG_PCG), ("learned", pvof.PV_PSOLVE_LEARNED_MG)):
        p.pressure_solver = solver
        s = pvof.Solver(p, lib)
        s.init_from_F(F)
        if solver == pvof.PV_PSOLVE_LEARNED_MG:
            pvof.Network(s).set_params(prm)
        iters, ratios = 0, []
        for step in range(3):
            s.call("Temporal_int_rk3", 1)
            stt = s.stats()
            assert max(stt["p_res"]) <= 1e-10, (name, step, stt["p_res"])
            iters += sum(stt["n_p"])
            if solver == pvof.PV_PSOLVE_LEARNED_MG:
                import ctypes as C
                r = C.c_double(0.0)
                lib.pv_learn_start_residual(s.h, C.byref(r))
                ratios.append(r.value)
        out[name] = {"iters": iters, "fields": {f: s.get(f) for f in ("C_L", "U", "V", "P")}, "ratios": ratios}
        s.close()
    R = regions(p)["cell"]
    errs = {f: relerr(out["learned"]["fields"][f][R], out["cold"]["fields"][f][R]) for f in ("C_L", "U", "V", "P")}
    for f, e in errs.items():
        assert e < 1e-6, (f, e)
    print(f"\n1024^2 learned MG-PCG: {nsamp} crops, loss {l0:.2e} -> {l1:.2e}; pressure iterations over 3 steps "
          f"{out['learned']['iters']} (cold MG-PCG {out['cold']['iters']}); last start energy ratios "
          f"{['%.3g' % r for r in out['learned']['ratios']]}; field differences " +
          ", ".join(f"{f} {e:.1e}" for f, e in errs.items()))
