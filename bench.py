#!/usr/bin/env python3
"""Throughput of PARTIES' two-phase RK3 step on one MI355X (gfx950).

Workload (one GPU): 1024 x 1024 planar cells on [0,1]^2, x periodic, no-slip
walls in y, 16-64 seeded gas bubbles of radius 0.01-0.08 (seed 20261202,
splitmix64), density ratio 1e-3, viscosity ratio 1e-2, Re = 100, We = 200,
Ri = 1, gravity -y, Cn = 0.004, Pe_CH = 1/Cn, CFL 0.2, default_dt 1e-5.
Every pressure solve stops at ||r|| / ||b|| <= 1e-10.  The data are synthetic
(generated in-library); all fields are resident in HBM before the timed
region starts.

A "step" is one full RK3 step (three substeps of Temporal_int_all_the_equations
including CFL).  value = NXM * NYM * steps / seconds / 1e6 (Mcell-steps/s).

Default run: the learned, mixed-precision form (network warm start +
multigrid-preconditioned CG pressure, FP32 stencils with FP64 fluxes, solves
and dots; each solve starts from the best of phi = 0, the same RK3 stage's
last correction and the network's correction by residual energy; the network
is trained first, untimed, on a held-out seed), then on
the same GPU the plain FP64 path (FFT-preconditioned CG, the reference
algorithm) and the same form without the network (ablation), each reported
with its own Mcell-steps/s.  --solver / --precision time one configuration.

The roofline block reports the HBM-bound kernel class (one with an
algorithmic byte count per cell) with the largest share of device time,
timed live with HIP events on the library's stream (every timer_period-th
launch) over a separate run of kernel_steps steps of the same configuration,
so that the event records do not perturb the timed run.  The multigrid's
one-workgroup coarse kernel (MG_COARSE in kernels) can take more time; it
runs on one CU, bound by its workgroup barriers rather than by HBM, so it has
no byte count and no roofline.
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
          "FFT_ROWS_FWD", "FFT_COLS", "FFT_ROWS_INV", "CH_RHS", "MOM_TERMS", "P_TRANSPOSE",
          "MG_PRE0", "MG_COARSE", "MG_POST0"]
# algorithmic HBM bytes per cell and launch (each distinct array read or
# written once; stencil neighbours come from cache); MOM_* per velocity
# component (a launch carries u and v together)
BYTES_PER_CELL = {
    "CH_APPLY": 32, "CH_UPDATE": 48, "MOM_APPLY": 56, "MOM_UPDATE": 56,
    "P_APPLY": 24, "P_UPDATE": 48, "P_DUPDATE": 24,
    "FFT_ROWS_FWD": 16, "FFT_COLS": 16, "FFT_ROWS_INV": 24,
    "CH_RHS": 56, "MOM_TERMS": 88, "P_TRANSPOSE": 16,
}
# the fused FFT-PCG kernels behind the P_APPLY / P_UPDATE timers also form the
# new direction (z, d_old, faces -> d, Ad) and the forward x-FFT (-> spectrum)
BYTES_PER_CELL_FFT = dict(BYTES_PER_CELL, P_APPLY=48, P_UPDATE=56)
# MG-PCG: the same direction kernel; the level-0 pre-smoothing also does the
# CG step (r_n, A d, d, phi, faces -> r_{n+1}, phi, pre-smoothed x, coarse f)
# and correction + post-smoothing (x, coarse x, r, faces -> z).  P_UPDATE is
# the separate step kernel of grids without tiled levels.
BYTES_PER_CELL_MG = dict(BYTES_PER_CELL, P_APPLY=48, P_UPDATE=48, MG_PRE0=74, MG_POST0=42)
# the FP32 V-cycle (mixed precision): FP32 faces and pre-smoothed x, FP64 r,
# A d, d, phi and z
BYTES_PER_CELL_MG32 = dict(BYTES_PER_CELL_MG, MG_PRE0=61, MG_POST0=29)
# per pressure iteration, all of its kernels (MG: level 0 + 1/3 of the level-0
# smoothing bytes, without the CG step, for the coarser levels)
BYTES_PER_ITER = {"fft": 144, "mg": 189, "mg32": 155, "jacobi": 112}
# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on
# the 1024^2 MG-PCG mixed form (FETCH_SIZE scaled by 2.0, WRITE_SIZE by 1.0:
# the factors a 256 MiB device copy gives in runs of its own); the summary is
# profiles/mg_pcg_mixed_1024_pmc_traffic.md
TRAFFIC_MB_MG32_1024 = {"MG_PRE0": 66.5, "MG_POST0": 32.0, "P_APPLY": 51.5,
                        "MOM_APPLY": 114.6, "MOM_UPDATE": 226.0, "CH_RHS": 68.8, "MOM_TERMS": 111.0}
SEED = 20261202
# the headline metric: the timed form is the learned-start mixed form, the plain
# FP64 path is reported beside it (compare.plain_fp64), the dominant kernel's
# fraction of the HBM peak in roofline
METRIC = "Mcell-steps/s at 1024\u00b2, plain FP64 vs learned-PCG mixed form; % of HBM roofline"


# momentum CG stop: the reference default (IO/default.inp [lsolve] CG_ETOL)
CG_ETOL = 1e-9


def workload_params(n=1024, solver=pvof.PV_PSOLVE_FFT_PCG, precision=pvof.PV_PREC_FP64, cg_etol=None):
    return pvof.default_params(
        NXM=n, NYM=n, xmax=1.0, ymax=1.0, Re=100.0, We=200.0, rho1=1.0, rho2=1e-3, mu1=1.0, mu2=1e-2,
        Cn=0.004, Pe_CH=1.0 / 0.004, cfl=0.2, default_dt=1e-5, max_dt=1e-3,
        ch_tol=1e-12, ch_iter_max=300, CG_ETOL=CG_ETOL if cg_etol is None else cg_etol, CG_MAXIT=120,
        P_CG_ETOL=1e-7, P_CG_MAXIT=300, pressure_solver=solver, p_rel_tol=1e-10, p_max_iter=3000,
        precision=precision)


def step_bytes_per_cell(st, family, mixed=False, mom_pc=0):
    """Algorithmic bytes per cell for one RK3 step from the solver counts
    (fixed 563 B per substep + 88 per biharmonic, 136 per momentum (140 in the
    mixed form, whose operator kernel reads five FP32 coefficients instead of
    rho and mu; 16 more with the polynomial preconditioner) and BYTES_PER_ITER
    per pressure iteration)."""
    total, mom = 0, (140 if mixed else 136) + (16 if mom_pc else 0)
    for k in range(3):
        total += 563 + 88 * st["n_ch"][k] + mom * (st["n_u"][k] + st["n_v"][k]) + BYTES_PER_ITER[family] * st["n_p"][k]
    return total


def family_of(solver):
    return "jacobi" if solver == "jacobi" else ("mg" if solver in ("mg", "learned_mg") else "fft")


SOLVERS = {"fft": pvof.PV_PSOLVE_FFT_PCG, "jacobi": pvof.PV_PSOLVE_JACOBI_PCG, "mg": pvof.PV_PSOLVE_MG_PCG,
           "learned": pvof.PV_PSOLVE_LEARNED, "learned_fcg": pvof.PV_PSOLVE_LEARNED_FCG,
           "learned_mg": pvof.PV_PSOLVE_LEARNED_MG}
LABEL = {"fft": "FFT-PCG", "jacobi": "Jacobi-PCG", "mg": "MG-PCG", "learned": "learned warm start + FFT-PCG",
         "learned_fcg": "learned warm start + learned-preconditioned FCG",
         "learned_mg": "learned warm start + MG-PCG"}


def cpu_baseline(params, F, min_seconds=10.0, max_steps=3):
    """Time the C restatement of the reference algorithm (the oracle built with
    OpenMP over rows, oracle/_build/libpvo_omp.so) on whole RK3 steps of the
    same workload from the same initial field, on the host cores this process
    may use (OMP_NUM_THREADS when set): steps until min_seconds have passed
    (at least one, at most max_steps)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    if "OMP_NUM_THREADS" not in os.environ:
        os.environ["OMP_NUM_THREADS"] = str(len(os.sched_getaffinity(0)))
    import oracle_lib

    lib = oracle_lib.load_oracle(os.path.join(ROOT, "oracle", "_build", "libpvo_omp.so"))
    lib.pvo_threads.restype = C.c_int
    threads = int(lib.pvo_threads())
    o = oracle_lib.Oracle(params, lib)
    o.init_from_F(F)
    steps, elapsed = 0, 0.0
    while steps < max_steps and (steps == 0 or elapsed < min_seconds):
        t0 = time.perf_counter()
        rc = o.call("Temporal_int_rk3", 1)
        elapsed += time.perf_counter() - t0
        if rc != 0:
            return None
        steps += 1
    cells = params.NXM * params.NYM
    return {"value": round(cells * steps / elapsed / 1e6, 5), "unit": "Mcell-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"{steps} whole RK3 step(s) ({params.NXM}x{params.NYM}, same initial field, FFT-PCG pressure "
                      f"to {params.p_rel_tol:g}, CG_ETOL {params.CG_ETOL:g}) of the C restatement of the reference "
                      f"algorithm with OpenMP over rows on {threads} host threads: {elapsed:.1f} s"}


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


def run(lib, args, solver, precision, steps, warmup, F, net_params=None, timers=True, opts=None):
    """Time `steps` RK3 steps of one configuration from the seeded field F
    (after `warmup` untimed steps); returns the measurements.  opts overrides
    the MG forms' settings (warm_start, ch_spectral, mg_cycle, mg_sweeps)."""
    prec = pvof.PV_PREC_MIXED if precision == "mixed" else pvof.PV_PREC_FP64
    p = workload_params(args.n, SOLVERS[solver], prec, args.cg_etol)
    s = pvof.Solver(p, lib, device=0)
    s.init_from_F(F)
    if net_params is not None:
        pvof.Network(s).set_params(net_params)
    o = dict(warm_start=args.warm_start, ch_spectral=args.ch_spectral, mg_cycle=args.mg_cycle,
             mg_sweeps=args.mg_sweeps, mom_warm=args.mom_warm, mom_exact=args.mom_exact,
             mom_pc=args.mom_pc if precision == "mixed" else 0)
    o.update(opts or {})
    if family_of(solver) == "mg":
        s.check(lib.pv_set_pressure_warm_start(s.h, o["warm_start"]))
        s.check(lib.pv_set_ch_solver(s.h, o["ch_spectral"]))
        s.check(lib.pv_set_mg_cycle(s.h, o["mg_cycle"]))
        s.check(lib.pv_set_mg_sweeps(s.h, o["mg_sweeps"]))
        s.check(lib.pv_set_momentum_warm_start(s.h, o["mom_warm"]))
        s.check(lib.pv_set_momentum_cg_steps(s.h, o["mom_exact"]))
        s.check(lib.pv_set_momentum_preconditioner(s.h, o["mom_pc"]))
    if warmup:
        s.call("Temporal_int_rk3", warmup)
    s.call("synchronize")
    if timers:
        s.call("timers_enable", max(1, args.timer_period))
        s.call("timers_reset")
    stats, starts = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        s.call("Temporal_int_rk3", 1)
        stats.append(s.stats())
        if net_params is not None:
            r = C.c_double(0.0)
            lib.pv_learn_start_residual(s.h, C.byref(r))
            starts.append(r.value)
    s.call("synchronize")
    elapsed = time.perf_counter() - t0
    ms = (C.c_double * len(TIMERS))()
    cnt = (C.c_longlong * len(TIMERS))()
    wms = (C.c_double * len(TIMERS))()
    wcnt = (C.c_longlong * len(TIMERS))()
    if timers:
        s.check(lib.pv_timers_read(s.h, ms, cnt))
        # launches that did their work (CG launches queued past convergence
        # exit on the device stop flag and move next to nothing)
        s.check(lib.pv_timers_read_working(s.h, wms, wcnt))
    fam = family_of(solver)
    if fam == "mg" and precision == "mixed":      # mixed handles run the FP32 V-cycle
        fam = "mg32"
    bpc = {"jacobi": BYTES_PER_CELL, "fft": BYTES_PER_CELL_FFT, "mg": BYTES_PER_CELL_MG,
           "mg32": BYTES_PER_CELL_MG32}[fam]
    cells = p.NXM * p.NYM
    if precision == "mixed":
        # the FP32 operator reads r, diag, d_old and five FP32 coefficients per
        # component (formed by the start kernel) instead of rho and mu
        bpc = dict(bpc, MOM_APPLY=60)
        if o["mom_pc"]:
            # polynomial preconditioner: the operator reads z instead of r and
            # diag; the update also reads diag and four coefficients and
            # writes r_{n+1} and z
            bpc = dict(bpc, MOM_APPLY=52, MOM_UPDATE=80)
    if not any(sum(st["n_ch"]) for st in stats):
        # spectral CH start: the CH CG launches only take the stop test (no cell work)
        bpc = {k: v for k, v in bpc.items() if k not in ("CH_APPLY", "CH_UPDATE")}
    # the momentum CG runs u and v in the same launches: a launch moves the
    # per-component bytes once per component still iterating, on average
    # (n_u + n_v) / max(n_u, n_v) components (one after the other has stopped)
    nu = sum(sum(st["n_u"]) for st in stats)
    nv = sum(sum(st["n_v"]) for st in stats)
    mlaunch = sum(max(a, b) for st in stats for a, b in zip(st["n_u"], st["n_v"]))
    if mlaunch:
        comps = (nu + nv) / mlaunch
        bpc = dict(bpc, MOM_APPLY=bpc["MOM_APPLY"] * comps, MOM_UPDATE=bpc["MOM_UPDATE"] * comps)
    kernels = {}
    for k, name in enumerate(TIMERS):
        if cnt[k]:
            # avg_us and GBs over the working launches (at least a quarter of
            # the class's longest sampled launch), the bytes per launch being
            # those of a launch that does the work; avg_us_all over all
            avg_ms = wms[k] / wcnt[k] if wcnt[k] else ms[k] / cnt[k]
            kernels[name] = {"timed_launches": int(cnt[k]), "working_launches": int(wcnt[k]),
                             "avg_us": round(1e3 * avg_ms, 3), "avg_us_all": round(1e3 * ms[k] / cnt[k], 3),
                             "est_total_ms": round(ms[k] * max(1, args.timer_period), 3),
                             "GBs": (round(bpc[name] * cells / (avg_ms * 1e-3) / 1e9, 1) if name in bpc else None)}
    bytes_step = sum(step_bytes_per_cell(st, fam, precision == "mixed", o["mom_pc"]) for st in stats) / len(stats)
    out = {"solver": solver, "precision": precision, "steps": steps, "warmup": warmup,
           "settings": o if family_of(solver) == "mg" else None,
           "value": round(cells * steps / elapsed / 1e6, 3), "ms_per_step": round(1e3 * elapsed / steps, 3),
           "iterations_per_step": {key: float(np.mean([sum(st[key]) for st in stats]))
                                   for key in ("n_ch", "n_u", "n_v", "n_p")},
           "max_p_res": float(max(max(st["p_res"]) for st in stats)),
           "path_hbm_fraction": round(bytes_step * cells * steps / elapsed / 1e9 / HBM_PEAK_GBS, 4),
           "path_bytes_per_cell_step": round(bytes_step, 1), "kernels": kernels, "_bpc": bpc, "_cells": cells,
           "_params": p, "_family": fam}
    if net_params is not None:
        out["start_energy_ratio_last_step"] = starts[-1] if starts else None
        out["fallbacks"] = int(sum(sum(st["p_fallback"]) for st in stats))
    s.close()
    return out


def mixed_error(lib, args):
    """Per-field error of the fast mixed form (FP32 stencils and multigrid
    cycle, W on --mg-cycle in-workgroup levels, FP64 fluxes / solves / dots,
    spectral CH start, temporal pressure and momentum starts, momentum CG
    steps without EPS) against the FP64 path (FP64 stencils and V-cycle, the
    reference CH and momentum CGs, cold starts), both MG-PCG, stepped side by
    side from the same seeded field; errors at checkpoints up to
    --mixed-error-steps."""
    n, steps = args.n, args.mixed_error_steps
    if args.ic == "droplets":          # 32 liquid droplets of radius 0.02-0.06 in gas
        kw = dict(nmin=32, nmax=32, rmin=0.02, rmax=0.06, gas_inside=0)
    else:
        kw = dict(nmin=16, nmax=64, rmin=0.01, rmax=0.08, gas_inside=1)
    runs, secs = {}, {}
    for name, prec in (("fp64", pvof.PV_PREC_FP64), ("mixed", pvof.PV_PREC_MIXED)):
        p = workload_params(n, pvof.PV_PSOLVE_MG_PCG, prec)
        F = pvof.ic(lib, "bubbles", p, seed=SEED, **kw)
        s = pvof.Solver(p, lib, device=0)
        s.init_from_F(F)
        if name == "mixed":
            s.check(lib.pv_set_ch_solver(s.h, 1))
            s.check(lib.pv_set_pressure_warm_start(s.h, 1))
            s.check(lib.pv_set_mg_cycle(s.h, args.mg_cycle))
            s.check(lib.pv_set_momentum_warm_start(s.h, args.mom_warm))
            s.check(lib.pv_set_momentum_cg_steps(s.h, args.mom_exact))
            s.check(lib.pv_set_momentum_preconditioner(s.h, args.mom_pc))
        runs[name], secs[name] = s, 0.0
    R = (slice(3, 3 + n), slice(3, 3 + n))
    marks = sorted({m for m in (10, 50, 100, 250, steps) if m <= steps})
    checkpoints, done = [], 0
    for m in marks:
        for name, s in runs.items():
            t0 = time.perf_counter()
            for _ in range(m - done):
                s.call("Temporal_int_rk3", 1)
            s.call("synchronize")
            secs[name] += time.perf_counter() - t0
        done = m
        fields = {}
        for f in ("C_L", "U", "V", "P"):
            a, b = runs["mixed"].get(f)[R], runs["fp64"].get(f)[R]
            d = a - b
            fields[f] = {"rel_L2": float(np.sqrt((d * d).sum() / max((b * b).sum(), 1e-300))),
                         "rel_Linf": float(np.abs(d).max() / max(np.abs(b).max(), 1e-300))}
        cl = [runs[k].get("C_L")[R].sum() for k in ("fp64", "mixed")]
        t = [runs[k].stats()["time"] for k in ("fp64", "mixed")]
        checkpoints.append({"steps": m, "fields": fields, "time": t,
                            "phase_volume_rel_diff": float(abs(cl[1] - cl[0]) / abs(cl[0]))})
        print(f"mixed_error: step {m}: " + ", ".join(f"{f} {v['rel_L2']:.2e}" for f, v in fields.items()),
              file=sys.stderr, flush=True)
    return {"n": n, "steps": steps, "ic": args.ic, "checkpoints": checkpoints,
            "fp64_ms_per_step": round(1e3 * secs["fp64"] / steps, 2),
            "mixed_ms_per_step": round(1e3 * secs["mixed"] / steps, 2)}


def batch_bench(lib, args):
    """Batched cases: --batch-cases cases of sizes drawn (seeded) from the
    batched configuration's {96^2, 192^2, 256x512, 384x768}, FP64, MG-PCG with
    the temporal start, the momentum CG options of --mom-exact / --mom-warm,
    seeded bubbles
    (the IBM cylinder and wetting of that configuration are not part of this
    path).  Aggregate Mcell-steps/s of pv_batch_rk3 (one stream and host
    thread per case) against the same cases stepped one after another.  Each
    case's cycle and sweep count come from pv_batch_autotune_mg after the
    warm-up (untimed; the one-after-another run uses the same choices) unless
    --no-batch-autotune is given."""
    rng = np.random.default_rng(SEED)
    shapes = [(96, 96), (192, 192), (256, 512), (384, 768)]
    picks = [shapes[k] for k in rng.integers(0, len(shapes), args.batch_cases)]
    params = []
    for nx, ny in picks:
        p = workload_params(nx, pvof.PV_PSOLVE_MG_PCG, pvof.PV_PREC_FP64)
        p.NYM, p.ymax = ny, ny / nx
        p.Cn, p.Pe_CH = 2.0 / nx, nx / 2.0          # interface two cells wide on every size
        params.append(p)
    fields = [pvof.ic(lib, "bubbles", p, seed=SEED + k, nmin=2, nmax=6, rmin=0.05, rmax=0.15, gas_inside=1)
              for k, p in enumerate(params)]

    def setup(s, F):
        s.init_from_F(F)
        s.check(lib.pv_set_pressure_warm_start(s.h, 1))
        s.check(lib.pv_set_mg_cycle(s.h, args.mg_cycle))
        s.check(lib.pv_set_momentum_cg_steps(s.h, args.mom_exact))
        s.check(lib.pv_set_momentum_warm_start(s.h, args.mom_warm))
        if args.batch_graphs > 0:
            s.check(lib.pv_set_graphs(s.h, 1, args.batch_graphs))

    cells = sum(p.NXM * p.NYM for p in params)
    bt = pvof.Batch(params, lib)
    bt.set_threads(args.batch_threads)
    for c, F in zip(bt.cases, fields):
        setup(c, F)
    note(f"batch of {len(params)} cases, {cells / 1e6:.2f} Mcells: {args.warmup} + {args.steps} steps")
    bt.rk3(args.warmup)
    choices = None
    if args.batch_autotune:
        note("autotuning every case's multigrid cycle")
        choices = bt.autotune_mg()
    for c in bt.cases:
        c.call("synchronize")
    t0 = time.perf_counter()
    bt.rk3(args.steps)
    for c in bt.cases:
        c.call("synchronize")
    tb = time.perf_counter() - t0
    bt.close()
    note("the same cases one after another")
    ts = 0.0
    for k, (p, F) in enumerate(zip(params, fields)):
        s = pvof.Solver(p, lib, device=0)
        setup(s, F)
        s.call("Temporal_int_rk3", args.warmup)
        if choices is not None:         # as in the batch: tuned after the warm-up
            s.check(lib.pv_set_mg_cycle(s.h, choices[k][0]))
            s.check(lib.pv_set_mg_sweeps(s.h, choices[k][1]))
        s.call("synchronize")
        t0 = time.perf_counter()
        s.call("Temporal_int_rk3", args.steps)
        s.call("synchronize")
        ts += time.perf_counter() - t0
        s.close()
    return {"cases": len(params), "sizes": {f"{nx}x{ny}": picks.count((nx, ny)) for nx, ny in shapes},
            "cells": cells, "steps": args.steps, "warmup": args.warmup,
            "batch_Mcell_steps_per_s": round(cells * args.steps / tb / 1e6, 3),
            "one_after_another_Mcell_steps_per_s": round(cells * args.steps / ts / 1e6, 3),
            "batch_gain": round(ts / tb, 3), "precision": "fp64", "pressure_solver": "mg",
            "momentum_exact_steps": args.mom_exact, "momentum_warm_start": args.mom_warm,
            "mg_cycle_w_levels": args.mg_cycle if choices is None else "autotuned",
            "graph_chunk": args.batch_graphs or None,
            "host_threads": min(args.batch_threads, len(params)) if args.batch_threads else len(params),
            "mg_choices": None if choices is None else
            {f"{nx}x{ny}": sorted({f"W{w}/nu{nu}": sum(1 for q, ch in zip(picks, choices) if q == (nx, ny) and
                                                      ch == (w, nu))
                                   for q, (w, nu) in zip(picks, choices) if q == (nx, ny)}.items())
             for nx, ny in shapes}}


def config2_run(lib, args):
    """The plain FP64 path at scale: 512 x 1024 cells, heavy fluid above a
    seeded 8-mode interface, Atwood 0.5, Re 3000, Cn 0.005, viscosity ratio 1,
    FFT-PCG pressure solve, --config2-steps RK3 steps; square cells (a 1 x 2
    domain, interface at y = 1), which the uniform-grid pressure right-hand
    side needs.  Throughput, iteration counts, the phase-volume change, the
    clamp corrections as the reference reports them (summed over its G range,
    which includes the duplicated periodic column and the ghost row above the
    top wall, VOF_DIFFUSE.c:1093-1127) and over the interior cells alone, the
    differences of each from the volume change, the largest divergence and
    finiteness."""
    p = workload_params(512, pvof.PV_PSOLVE_FFT_PCG, pvof.PV_PREC_FP64)
    for k, v in dict(NYM=1024, ymax=2.0, Re=3000.0, rho2=1.0 / 3.0, mu2=1.0, Cn=0.005, Pe_CH=200.0,
                     default_dt=1e-4).items():
        setattr(p, k, v)
    F = pvof.ic(lib, "multimode_rt", p, seed=SEED, modes=8, amp_max=0.05, y0=1.0)
    s = pvof.Solver(p, lib, device=0)
    s.init_from_F(F)
    R = (slice(3, 3 + p.NYM), slice(3, 3 + p.NXM))
    pp = s.params
    meas = (pp.xmax - pp.xmin) / p.NXM * (pp.ymax - pp.ymin) / p.NYM * pp.twod_slab_thickness
    vol0 = s.get("C_L")[R].sum() * meas
    clamp, clamp_int, iters, div = 0.0, 0.0, {"n_ch": 0, "n_u": 0, "n_v": 0, "n_p": 0}, 0.0
    steps, done = args.config2_steps, 0
    t = 0.0
    while done < steps:
        k = min(50, steps - done)
        t0 = time.perf_counter()
        for _ in range(k):
            s.call("Temporal_int_rk3", 1)
            st = s.stats()
            clamp += sum(st["clamp_delta"])
            clamp_int += sum(st["clamp_delta_interior"])
            div = max(div, max(st["div_max"]))
            for key in iters:
                iters[key] += sum(st[key])
        s.call("synchronize")
        t += time.perf_counter() - t0
        done += k
        note(f"config 2: {done} of {steps} steps")
    vol = s.get("C_L")[R].sum() * meas
    finite = all(bool(np.isfinite(s.get(f)).all()) for f in ("C_L", "U", "V", "P"))
    cells = p.NXM * p.NYM
    return {"NXM": p.NXM, "NYM": p.NYM, "steps": steps, "time": s.stats()["time"],
            "Mcell_steps_per_s": round(cells * steps / t / 1e6, 3), "ms_per_step": round(1e3 * t / steps, 3),
            "iterations_per_step": {k: round(v / steps, 2) for k, v in iters.items()},
            "phase_volume_change_rel": float((vol - vol0) / vol0), "clamp_total_rel": float(clamp / vol0),
            "phase_volume_change_less_clamp_rel": float(abs((vol - vol0) - clamp) / vol0),
            "clamp_interior_total_rel": float(clamp_int / vol0),
            "phase_volume_change_less_interior_clamp_rel": float(abs((vol - vol0) - clamp_int) / vol0),
            "max_divergence": div, "finite": finite, "precision": "fp64", "pressure_solver": "fft"}


def note(msg):
    print(f"bench.py: {msg}", file=sys.stderr, flush=True)


def public(r):
    return {k: v for k, v in r.items() if not k.startswith("_") and k != "kernels"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--solver", choices=sorted(SOLVERS), default="learned_mg",
                    help="pressure solver of the timed configuration (learned*: network trained first, "
                         "untimed, on a held-out seed)")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed",
                    help="mixed: FP32 WENO / chemical-potential / momentum stencils, FP64 fluxes, solves and dots")
    ap.add_argument("--warm-start", type=int, choices=[0, 1], default=1,
                    help="MG-PCG solvers: 1 = also try the same RK3 stage's last correction as the start")
    ap.add_argument("--mg-cycle", type=int, choices=[0, 1, 2, 3, 4], default=2,
                    help="MG solvers: the first k in-workgroup levels take two coarse corrections (0 = V-cycle)")
    ap.add_argument("--mg-sweeps", type=int, choices=[1, 2], default=2,
                    help="MG solvers: Gauss-Seidel sweeps before and after each coarse correction")
    ap.add_argument("--mom-warm", type=int, choices=[0, 1], default=1,
                    help="MG-PCG forms: 1 = start each momentum CG from the same RK3 stage's last increment")
    ap.add_argument("--mom-pc", type=int, choices=[0, 1], default=1,
                    help="mixed MG-PCG forms: 1 = momentum CG preconditioned by the first-order Neumann polynomial "
                         "of the Jacobi splitting (0 = Jacobi, as the reference)")
    ap.add_argument("--mom-exact", type=int, choices=[0, 1], default=1,
                    help="MG-PCG forms: 1 = momentum CG step lengths without the reference's EPS = 1e-14 "
                         "(same stop test; the EPS stalls the CG near CG_ETOL)")
    ap.add_argument("--ch-spectral", type=int, choices=[0, 1], default=1,
                    help="MG-PCG forms: 1 = start the biharmonic CG from its exact spectral solve")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the plain FP64 FFT-PCG run and the no-network ablation")
    ap.add_argument("--plain-steps", type=int, default=10)
    ap.add_argument("--no-fp64-mg", action="store_true",
                    help="skip the FP64 MG-PCG comparison runs")
    ap.add_argument("--cg-etol", type=float, default=CG_ETOL,
                    help="momentum CG stop (RMS residual), default the reference's 1e-9")
    ap.add_argument("--train-collect-steps", type=int, default=4)
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=1024)
    ap.add_argument("--kernel-steps", type=int, default=20,
                    help="steps of the separate run whose kernels are timed with HIP events")
    ap.add_argument("--timer-period", type=int, default=16,
                    help="time every k-th launch of each kernel class with HIP events")
    ap.add_argument("--no-timers", action="store_true")
    ap.add_argument("--mixed-error-steps", type=int, default=0,
                    help="only report the per-field error of the mixed form against the FP64 path after this "
                         "many steps at --n (with --ic)")
    ap.add_argument("--ic", choices=["bubbles", "droplets"], default="bubbles")
    ap.add_argument("--config2-steps", type=int, default=0,
                    help="only run the plain FP64 path at scale (512 x 1024 Rayleigh-Taylor) for this many steps")
    ap.add_argument("--batch-cases", type=int, default=0,
                    help="only report batched throughput: this many cases of the batched configuration's sizes "
                         "(--steps, --warmup per case)")
    ap.add_argument("--batch-autotune", action=argparse.BooleanOptionalAction, default=True,
                    help="with --batch-cases: pick each case's multigrid cycle and sweeps with "
                         "pv_batch_autotune_mg after the warm-up (--no-batch-autotune: --mg-cycle on every case)")
    ap.add_argument("--batch-threads", type=int, default=16,
                    help="with --batch-cases: host threads sharing the cases (0 = one per case); more "
                         "threads than the CPUs the process has slow the batch down")
    ap.add_argument("--batch-graphs", type=int, default=0,
                    help="with --batch-cases: replay each case's pressure iterations from hipGraphs of this "
                         "many iterations (0 = eager launches)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only copy this many MiB device to device three times and exit: a kernel of known "
                         "traffic, to calibrate rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE in a run of its own")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs one subdomain on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note, not a code path)", file=sys.stderr)
        sys.exit(2)

    lib = pvof.load_lib()
    if args.mixed_error_steps > 0:
        print(json.dumps({"mixed_error": mixed_error(lib, args)}))
        return
    if args.batch_cases > 0:
        print(json.dumps({"batch": batch_bench(lib, args)}))
        return
    if args.config2_steps > 0:
        print(json.dumps({"config2": config2_run(lib, args)}))
        return
    if args.copy_mib > 0:
        import torch
        src = torch.ones(args.copy_mib * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": src.numel() * 8, "copies": 3}))
        return
    prec = pvof.PV_PREC_MIXED if args.precision == "mixed" else pvof.PV_PREC_FP64
    learned = None
    if args.solver.startswith("learned"):
        note("training the preconditioner network")
        learned = train_network(lib, args, prec)
    p0 = workload_params(args.n, SOLVERS[args.solver], prec)
    F = pvof.ic(lib, "bubbles", p0, seed=SEED, nmin=16, nmax=64, rmin=0.01, rmax=0.08, gas_inside=1)
    net = learned.pop("params") if learned is not None else None
    # the timed run records no events; a separate short run times the kernels
    note(f"timed run: {args.solver}, {args.precision}, {args.steps} steps after {args.warmup}")
    main_run = run(lib, args, args.solver, args.precision, args.steps, args.warmup, F, net, timers=False)
    if not args.no_timers:
        note(f"kernel timing run: {args.kernel_steps} steps")
        prof = run(lib, args, args.solver, args.precision, args.kernel_steps, 2, F, net, timers=True)
        main_run["kernels"], main_run["_bpc"] = prof["kernels"], prof["_bpc"]
    p = main_run["_params"]
    compare = {}
    if not args.no_compare:
        note(f"plain FP64 FFT-PCG run: {args.plain_steps} steps")
        plain = run(lib, args, "fft", "fp64", args.plain_steps, 1, F, timers=False)
        compare["plain_fp64"] = public(plain)
        compare["speedup_vs_plain_fp64"] = round(main_run["value"] / plain["value"], 3)
        if args.solver.startswith("learned"):
            base = {"learned_mg": "mg", "learned": "fft", "learned_fcg": "fft"}[args.solver]
            note(f"ablation run without the network: {base}")
            abl = run(lib, args, base, args.precision, args.steps, args.warmup, F, timers=False)
            compare["ablation_no_network"] = public(abl)
            compare["network_gain"] = round(main_run["value"] / abl["value"], 3)
        if not args.no_fp64_mg and family_of(args.solver) == "mg":
            note(f"FP64 MG-PCG runs: {args.plain_steps} steps each")
            # the best FP64 path: multigrid V-cycle, the reference CH CG, cold starts, no network
            ref = run(lib, args, "mg", "fp64", args.plain_steps, 1, F, timers=False,
                      opts=dict(warm_start=0, ch_spectral=0, mg_cycle=0, mom_warm=0, mom_exact=0, mom_pc=0))
            compare["fp64_mg"] = public(ref)
            compare["speedup_vs_fp64_mg"] = round(main_run["value"] / ref["value"], 3)
            # the timed form's algorithm in FP64 throughout (isolates the FP32 stencils and cycle)
            same = run(lib, args, "mg", "fp64", args.plain_steps, 1, F, timers=False)
            compare["fp64_mg_same_settings"] = public(same)
            compare["speedup_vs_fp64_mg_same_settings"] = round(main_run["value"] / same["value"], 3)

    form = LABEL[args.solver] + (", mixed FP32/FP64" if args.precision == "mixed" else ", FP64")
    out = {
        "metric": (METRIC if p.NXM == p.NYM == 1024 and args.solver.startswith("learned") and
                   args.precision == "mixed" else
                   f"Mcell-steps/s at {p.NXM}x{p.NYM}, plain FP64 vs {form}; % of HBM roofline"),
        "value": main_run["value"],
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_run["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic (seeded bubble field generated in-library)",
        "config": {"workload": f"planar {p.NXM}x{p.NYM} bubbly two-phase RK3 step, rho2/rho1=1e-3, "
                               f"{LABEL[args.solver]} pressure to 1e-10 relative residual",
                   "form": f"{form} (value); plain FP64 FFT-PCG in compare.plain_fp64",
                   "NXM": p.NXM, "NYM": p.NYM, "pressure_solver": args.solver, "precision": args.precision,
                   "warm_start": args.warm_start if family_of(args.solver) == "mg" else None,
                   "ch_spectral": args.ch_spectral if family_of(args.solver) == "mg" else None,
                   "momentum_warm_start": args.mom_warm if family_of(args.solver) == "mg" else None,
                   "momentum_exact_steps": args.mom_exact if family_of(args.solver) == "mg" else None,
                   "momentum_preconditioner": ("neumann1" if args.mom_pc else "jacobi")
                   if family_of(args.solver) == "mg" and args.precision == "mixed" else "jacobi",
                   "mg_cycle_w_levels": args.mg_cycle if family_of(args.solver) == "mg" else None,
                   "p_rel_tol": p.p_rel_tol, "ch_tol": p.ch_tol, "CG_ETOL": p.CG_ETOL},
        "path_hbm_fraction": main_run["path_hbm_fraction"],
        "path_bytes_per_cell_step": main_run["path_bytes_per_cell_step"],
        "iterations_per_step": main_run["iterations_per_step"],
        "max_p_res": main_run["max_p_res"],
        "roofline": None,
        "kernels": main_run["kernels"],
        "compare": compare,
        "cpu_baseline": None,
    }
    if learned is not None:
        learned["start_energy_ratio_last_step"] = main_run.get("start_energy_ratio_last_step")
        learned["fallbacks"] = main_run.get("fallbacks")
        out["learned"] = learned
    kernels, bpc, cells = main_run["kernels"], main_run["_bpc"], main_run["_cells"]
    timed = {k: v for k, v in kernels.items() if v["GBs"] is not None}
    if timed:
        dom = max(timed, key=lambda n: timed[n]["est_total_ms"])
        achieved = timed[dom]["GBs"]
        traffic = None
        if main_run["_family"] == "mg32" and cells == 1024 * 1024 and dom in TRAFFIC_MB_MG32_1024:
            traffic = TRAFFIC_MB_MG32_1024[dom] * 1e6
        out["roofline"] = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                           "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated)",
                           "bytes_per_cell": bpc[dom], "cells_per_launch": cells,
                           "avg_us": timed[dom]["avg_us"], "timer_period": args.timer_period}
    if not args.no_cpu_baseline:
        note(f"CPU baseline at {args.cpu_n}^2")
        pc = workload_params(args.cpu_n, pvof.PV_PSOLVE_FFT_PCG, cg_etol=args.cg_etol)
        Fc = pvof.ic(lib, "bubbles", pc, seed=SEED, nmin=16, nmax=64, rmin=0.01, rmax=0.08, gas_inside=1)
        out["cpu_baseline"] = cpu_baseline(pc, Fc)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
