#!/usr/bin/env python3
"""Throughput of PARTIES' two-phase RK3 step on one MI355X (gfx950).

Workload (one GPU): 1024 x 1024 planar cells on [0,1]^2, x periodic, no-slip
walls in y, 16-64 seeded gas bubbles of radius 0.01-0.08 (seed 20261202,
splitmix64), density ratio 1e-3, viscosity ratio 1e-2, Re = 100, We = 200,
Ri = 1, gravity -y, Cn = 0.004, Pe_CH = 1/Cn, CFL 0.2, default_dt 1e-5.
Pressure: FFT-x/DCT-y preconditioned PCG to ||r||/||b|| <= 1e-10.  The data are
synthetic (generated in-library); all fields are resident in HBM before the
timed region starts.

A "step" is one full RK3 step (three substeps of Temporal_int_all_the_equations
including CFL).  value = NXM * NYM * steps / seconds / 1e6 (Mcell-steps/s).

The roofline block reports the kernel with the largest share of device time,
timed live with HIP events on the library's stream over the timed region.
The cpu_baseline block times the CPU oracle (serial C restatement of the
reference algorithm) on a bounded sample of the same workload.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "parties-vof-ibm_amd")
sys.path.insert(0, PKG)
import pvof  # noqa: E402

HBM_PEAK_GBS = 8000.0
TIMERS = ["CH_APPLY", "CH_UPDATE", "MOM_APPLY", "MOM_UPDATE", "P_APPLY", "P_UPDATE", "P_DUPDATE",
          "FFT_ROWS_FWD", "FFT_COLS", "FFT_ROWS_INV", "CH_RHS", "MOM_TERMS", "P_TRANSPOSE", "MG_VCYCLE"]
# algorithmic HBM bytes per cell and launch (each distinct array read or
# written once; stencil neighbours come from cache)
BYTES_PER_CELL = {
    "CH_APPLY": 16, "CH_UPDATE": 48, "MOM_APPLY": 32, "MOM_UPDATE": 56,
    "P_APPLY": 24, "P_UPDATE": 48, "P_DUPDATE": 24,
    "FFT_ROWS_FWD": 16, "FFT_COLS": 16, "FFT_ROWS_INV": 24,
    "CH_RHS": 56, "MOM_TERMS": 88, "P_TRANSPOSE": 16,
}
# the fused FFT-PCG kernels behind the P_APPLY / P_UPDATE timers also form the
# new direction (z, d_old, rho -> d, Ad) and the forward x-FFT (-> spectrum)
BYTES_PER_CELL_FFT = dict(BYTES_PER_CELL, P_APPLY=40, P_UPDATE=56)
SEED = 20261202


def workload_params(n=1024, solver=pvof.PV_PSOLVE_FFT_PCG, precision=pvof.PV_PREC_FP64):
    return pvof.default_params(
        NXM=n, NYM=n, xmax=1.0, ymax=1.0, Re=100.0, We=200.0, rho1=1.0, rho2=1e-3, mu1=1.0, mu2=1e-2,
        Cn=0.004, Pe_CH=1.0 / 0.004, cfl=0.2, default_dt=1e-5, max_dt=1e-3,
        ch_tol=1e-12, ch_iter_max=300, CG_ETOL=1e-7, CG_MAXIT=120,
        P_CG_ETOL=1e-7, P_CG_MAXIT=300, pressure_solver=solver, p_rel_tol=1e-10, p_max_iter=3000,
        precision=precision)


def step_bytes_per_cell(st):
    """Algorithmic bytes per cell for one RK3 step from the solver counts
    (fixed 563 B per substep + 88 per biharmonic, 136 per momentum and 152 per
    FFT-preconditioned pressure iteration)."""
    total = 0
    for k in range(3):
        total += 563 + 88 * st["n_ch"][k] + 136 * (st["n_u"][k] + st["n_v"][k]) + 152 * st["n_p"][k]
    return total


def cpu_baseline(params, F):
    """Time the CPU oracle on the first RK3 substep of the same workload (same
    initial field, same solver settings) and scale it to whole steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib

    o = oracle_lib.Oracle(params)
    o.init_from_F(F)
    o.call("set_time_state", 0.0, params.default_dt, params.default_dt, 0)
    t0 = time.perf_counter()
    rc = o.call("Temporal_int_all_the_equations")
    dt = time.perf_counter() - t0
    if rc != 0:
        return None
    cells = params.NXM * params.NYM
    return {"value": round(cells / (3.0 * dt) / 1e6, 5), "unit": "Mcell-steps/s", "cores": 1, "kind": "port",
            "sample": f"first RK3 substep ({params.NXM}x{params.NYM}, same initial field and solver settings) "
                      f"of the serial C oracle: {dt:.1f} s, scaled by 3 substeps per step"}


def train_network(lib, args, prec):
    """Train the preconditioner network (untimed) on the plain solver's own
    pressure solves of a held-out seed of the same workload."""
    t0 = time.perf_counter()
    pt = workload_params(args.n, pvof.PV_PSOLVE_FFT_PCG, prec)
    Ft = pvof.ic(lib, "bubbles", pt, seed=SEED + 1, nmin=16, nmax=64, rmin=0.01, rmax=0.08, gas_inside=1)
    st = pvof.Solver(pt, lib, device=0)
    st.init_from_F(Ft)
    net = pvof.Network(st, seed=3)
    net.collect([0, 8, 32, 96], crops_per_snap=400, crop_size=64, max_samples=20000, seed=5)
    st.call("Temporal_int_rk3", args.train_collect_steps)
    nsamp = net.collect_end()
    t1 = time.perf_counter()
    l0, l1 = net.train(args.train_steps, batch=32, lr=1e-3, seed=11)
    _, l2 = net.train(max(1, args.train_steps // 2), batch=32, lr=3e-4, seed=12)
    prm = net.params()
    st.close()
    return {"params": prm, "samples": nsamp, "collect_s": round(t1 - t0, 2),
            "train_s": round(time.perf_counter() - t1, 2), "train_steps": args.train_steps * 3 // 2,
            "loss_first": l0, "loss_last": l2, "held_out_seed": SEED + 1}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--solver", choices=["fft", "jacobi", "learned", "learned_fcg"], default="fft",
                    help="learned: network warm start + FFT-PCG; learned_fcg: + network-corrected "
                         "preconditioner in flexible CG (network trained first, untimed, on a held-out seed)")
    ap.add_argument("--train-collect-steps", type=int, default=4)
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="mixed: FP32 WENO / chemical-potential / momentum stencils, FP64 fluxes, solves and dots")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=1024)
    ap.add_argument("--timer-period", type=int, default=16,
                    help="time every k-th launch of each kernel class with HIP events")
    ap.add_argument("--no-timers", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs one subdomain on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note, not a code path)", file=sys.stderr)
        sys.exit(2)

    lib = pvof.load_lib()
    solver = {"fft": pvof.PV_PSOLVE_FFT_PCG, "jacobi": pvof.PV_PSOLVE_JACOBI_PCG,
              "learned": pvof.PV_PSOLVE_LEARNED, "learned_fcg": pvof.PV_PSOLVE_LEARNED_FCG}[args.solver]
    prec = pvof.PV_PREC_MIXED if args.precision == "mixed" else pvof.PV_PREC_FP64
    p = workload_params(args.n, solver, prec)
    learned = None
    if args.solver.startswith("learned"):
        learned = train_network(lib, args, prec)
    F = pvof.ic(lib, "bubbles", p, seed=SEED, nmin=16, nmax=64, rmin=0.01, rmax=0.08, gas_inside=1)
    s = pvof.Solver(p, lib, device=0)
    s.init_from_F(F)
    if learned is not None:
        pvof.Network(s).set_params(learned.pop("params"))
    if args.warmup:
        s.call("Temporal_int_rk3", args.warmup)
    s.call("synchronize")
    if not args.no_timers:
        s.call("timers_enable", max(1, args.timer_period))
        s.call("timers_reset")
    stats, starts = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.call("Temporal_int_rk3", 1)
        stats.append(s.stats())
        if learned is not None:
            r = C.c_double(0.0)
            lib.pv_learn_start_residual(s.h, C.byref(r))
            starts.append(r.value)
    s.call("synchronize")
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ms = (C.c_double * len(TIMERS))()
    cnt = (C.c_longlong * len(TIMERS))()
    if not args.no_timers:
        s.check(lib.pv_timers_read(s.h, ms, cnt))
    bpc = BYTES_PER_CELL if args.solver == "jacobi" else BYTES_PER_CELL_FFT

    cells = p.NXM * p.NYM
    value = cells * args.steps / elapsed / 1e6
    bytes_step = sum(step_bytes_per_cell(st) for st in stats) / len(stats)
    path_gbs = bytes_step * cells * args.steps / elapsed / 1e9
    kernels = {}
    for k, name in enumerate(TIMERS):
        if cnt[k]:
            avg_ms = ms[k] / cnt[k]
            kernels[name] = {"timed_launches": int(cnt[k]), "avg_us": round(1e3 * avg_ms, 3),
                             "est_total_ms": round(ms[k] * max(1, args.timer_period), 3),
                             "GBs": round(bpc[name] * cells / (avg_ms * 1e-3) / 1e9, 1)}
    iters = {key: float(np.mean([sum(st[key]) for st in stats])) for key in ("n_ch", "n_u", "n_v", "n_p")}

    out = {
        "metric": "Mcell-steps/s at 1024^2, " + ("plain FP64" if prec == pvof.PV_PREC_FP64 else "mixed FP32/FP64")
                  + {"fft": " (FFT-PCG pressure)", "jacobi": " (Jacobi-PCG pressure)",
                     "learned": " (learned warm start + FFT-PCG pressure)",
                     "learned_fcg": " (learned warm start + learned-preconditioned FCG pressure)"}[args.solver],
        "value": round(value, 3),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == pvof.PV_PREC_FP64 else "f32+f64",
        "data": "synthetic (seeded bubble field generated in-library)",
        "config": {"workload": f"planar {p.NXM}x{p.NYM} bubbly two-phase RK3 step, rho2/rho1=1e-3, "
                               f"{'FFT-PCG' if solver else 'Jacobi-PCG'} pressure",
                   "NXM": p.NXM, "NYM": p.NYM, "pressure_solver": args.solver, "precision": args.precision,
                   "p_rel_tol": p.p_rel_tol, "ch_tol": p.ch_tol, "CG_ETOL": p.CG_ETOL},
        "path_hbm_fraction": round(path_gbs / HBM_PEAK_GBS, 4),
        "path_bytes_per_cell_step": round(bytes_step, 1),
        "iterations_per_step": iters,
        "roofline": None,
        "kernels": kernels,
        "cpu_baseline": None,
    }
    if learned is not None:
        learned["start_residual_last_step"] = starts[-1] if starts else None
        learned["fallbacks"] = int(sum(sum(st["p_fallback"]) for st in stats))
        out["learned"] = learned
    if kernels:
        dom = max(kernels, key=lambda n: kernels[n]["est_total_ms"])
        achieved = kernels[dom]["GBs"]
        out["roofline"] = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                           "bytes_per_cell": bpc[dom], "cells_per_launch": cells,
                           "avg_us": kernels[dom]["avg_us"], "timer_period": args.timer_period}
    if not args.no_cpu_baseline:
        pc = workload_params(args.cpu_n, solver)
        Fc = F if args.cpu_n == args.n else pvof.ic(lib, "bubbles", pc, seed=SEED, nmin=16, nmax=64,
                                                   rmin=0.01, rmax=0.08, gas_inside=1)
        out["cpu_baseline"] = cpu_baseline(pc, Fc)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
