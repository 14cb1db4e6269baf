"""Launch graphs, autotuning and batches of uneven cases on the GPU.

The oracle for these is the plain path itself: a case replayed from
hipGraphs, run inside a batch, or autotuned between steps must produce bit
for bit the fields of the same case run alone with eager launches (all
kernels are deterministic: fixed-order reductions, no atomics).  Graph
replays happen many times per step with eager work and host synchronisation
in between, and every step is compared.
"""
import numpy as np
import pytest

import pvof

pytestmark = pytest.mark.gpu

FIELDS = ("C_L", "U", "V", "P")


def case_params(nx, ny):
    return pvof.default_params(
        NXM=nx, NYM=ny, xmax=1.0, ymax=ny / nx, Re=100.0, We=200.0, rho1=1.0, rho2=1e-3, mu1=1.0, mu2=1e-2,
        Cn=2.0 / nx, Pe_CH=nx / 2.0, cfl=0.2, default_dt=1e-4, max_dt=1e-3, ch_tol=1e-12, CG_ETOL=1e-7,
        CG_MAXIT=120, pressure_solver=pvof.PV_PSOLVE_FFT_PCG, p_rel_tol=1e-10, p_max_iter=3000)


def bubble_field(lib, p, seed):
    return pvof.ic(lib, "bubbles", p, seed=seed, nmin=3, nmax=8, rmin=0.05, rmax=0.15, gas_inside=1)


@pytest.fixture(scope="module")
def lib(gpu):
    return pvof.load_lib()


def test_graph_replay_matches_eager(lib):
    p = case_params(128, 128)
    F = bubble_field(lib, p, 7)
    a, b = pvof.Solver(p, lib), pvof.Solver(p, lib)
    a.init_from_F(F)
    b.init_from_F(F)
    b.check(lib.pv_set_graphs(b.h, 1, 8))
    for _ in range(3):
        a.call("Temporal_int_rk3", 1)
        b.call("Temporal_int_rk3", 1)
        assert a.stats()["n_p"] == b.stats()["n_p"]
        for f in FIELDS:
            assert np.array_equal(a.get(f), b.get(f)), f


@pytest.mark.parametrize("threads", [0, 2, 4])
def test_batch_cases_match_solo_runs(lib, threads):
    """threads: host threads sharing the 6 cases (0 = one per case)."""
    sizes = [(64, 64), (64, 128), (128, 128), (128, 64), (256, 128), (128, 256)]
    params = [case_params(nx, ny) for nx, ny in sizes]
    fields = [bubble_field(lib, p, 100 + k) for k, p in enumerate(params)]
    bt = pvof.Batch(params, lib)
    bt.set_threads(threads)
    for c, F in zip(bt.cases, fields):
        c.init_from_F(F)
    bt.rk3(3)
    for p, F, c in zip(params, fields, bt.cases):
        s = pvof.Solver(p, lib)
        s.init_from_F(F)
        s.call("Temporal_int_rk3", 3)
        for f in FIELDS:
            assert np.array_equal(c.get(f), s.get(f)), (p.NXM, p.NYM, f)
    bt.close()


@pytest.mark.parametrize("solver", [pvof.PV_PSOLVE_FFT_PCG, pvof.PV_PSOLVE_MG_PCG])
def test_batch_of_the_batched_configuration_sizes(lib, solver):
    """The batched configuration's sizes, 96^2, 192^2, 256 x 512 and
    384 x 768 (rows of 2^a 3^b for the FFT-x preconditioner), two cases each,
    three RK3 steps with either pressure solver: every case equals its solo
    run bitwise and its pressure solves meet the tolerance."""
    sizes = [(96, 96), (192, 192), (256, 512), (384, 768)] * 2
    params = []
    for nx, ny in sizes:
        p = case_params(nx, ny)
        p.pressure_solver = solver
        params.append(p)
    fields = [bubble_field(lib, p, 300 + k) for k, p in enumerate(params)]
    bt = pvof.Batch(params, lib)
    for c, F in zip(bt.cases, fields):
        c.init_from_F(F)
    bt.rk3(3)
    for p, F, c in zip(params, fields, bt.cases):
        assert max(c.stats()["p_res"]) <= 1e-10
        s = pvof.Solver(p, lib)
        s.init_from_F(F)
        s.call("Temporal_int_rk3", 3)
        for f in FIELDS:
            assert np.array_equal(c.get(f), s.get(f)), (p.NXM, p.NYM, f)
        s.close()
    bt.close()


def test_autotune_keeps_the_trajectory(lib):
    p = case_params(128, 128)
    F = bubble_field(lib, p, 9)
    a, b = pvof.Solver(p, lib), pvof.Solver(p, lib)
    a.init_from_F(F)
    b.init_from_F(F)
    a.call("Temporal_int_rk3", 1)
    b.call("Temporal_int_rk3", 1)
    ch = pvof.C.c_int(-1)
    b.check(lib.pv_autotune(b.h, pvof.C.byref(ch)))
    assert ch.value in (0, 8, 16, 32)
    a.call("Temporal_int_rk3", 2)
    b.call("Temporal_int_rk3", 2)
    for f in FIELDS:
        assert np.array_equal(a.get(f), b.get(f)), f


def test_mg_graph_replay_follows_cycle_change(lib):
    """MG-PCG replayed from hipGraphs stays bitwise equal to eager launches
    when the multigrid cycle changes between steps (graphs captured with the
    V-cycle are not replayed for the W)."""
    p = case_params(128, 128)
    p.pressure_solver = pvof.PV_PSOLVE_MG_PCG
    F = bubble_field(lib, p, 11)
    a, b = pvof.Solver(p, lib), pvof.Solver(p, lib)
    a.init_from_F(F)
    b.init_from_F(F)
    b.check(lib.pv_set_graphs(b.h, 1, 8))
    for step in range(4):
        if step == 2:
            for s in (a, b):
                s.check(lib.pv_set_mg_cycle(s.h, 2))
        a.call("Temporal_int_rk3", 1)
        b.call("Temporal_int_rk3", 1)
        assert a.stats()["n_p"] == b.stats()["n_p"], step
        for f in FIELDS:
            assert np.array_equal(a.get(f), b.get(f)), (step, f)


@pytest.mark.parametrize("warm", [0, 1])
def test_mg_autotune_keeps_the_trajectory(lib, warm):
    """pv_autotune_mg between steps: the tuned handle's next steps equal, bit
    for bit, those of a handle that never tuned and was set to the same cycle
    and sweep count (with and without the temporal start, whose stored
    starts and counters the tuning must leave alone)."""
    p = case_params(128, 128)
    p.pressure_solver = pvof.PV_PSOLVE_MG_PCG
    F = bubble_field(lib, p, 13)
    a, b = pvof.Solver(p, lib), pvof.Solver(p, lib)
    for s in (a, b):
        s.init_from_F(F)
        s.check(lib.pv_set_pressure_warm_start(s.h, warm))
        s.call("Temporal_int_rk3", 2)
    w, nu = pvof.C.c_int(-1), pvof.C.c_int(-1)
    b.check(lib.pv_autotune_mg(b.h, pvof.C.byref(w), pvof.C.byref(nu)))
    print(f"\nchosen: W on {w.value} levels, {nu.value} sweeps")
    assert 0 <= w.value <= 3 and nu.value in (1, 2)
    a.check(lib.pv_set_mg_cycle(a.h, w.value))
    a.check(lib.pv_set_mg_sweeps(a.h, nu.value))
    for step in range(2):
        a.call("Temporal_int_rk3", 1)
        b.call("Temporal_int_rk3", 1)
        assert a.stats()["n_p"] == b.stats()["n_p"], step
        for f in FIELDS:
            assert np.array_equal(a.get(f), b.get(f)), (step, f)


def test_batch_mg_autotune(lib):
    """Every MG-PCG case of a batch gets a valid choice, other cases (-1, -1);
    an FFT-PCG handle is refused."""
    pm = case_params(96, 96)
    pm.pressure_solver = pvof.PV_PSOLVE_MG_PCG
    pf = case_params(64, 64)
    bt = pvof.Batch([pm, pf], lib)
    for k, c in enumerate(bt.cases):
        c.init_from_F(bubble_field(lib, c.params, 40 + k))
    bt.rk3(1)
    (w0, nu0), (w1, nu1) = bt.autotune_mg()
    assert 0 <= w0 <= 3 and nu0 in (1, 2)
    assert (w1, nu1) == (-1, -1)
    bt.rk3(1)
    for c in bt.cases:
        assert max(c.stats()["p_res"]) <= 1e-10
    bt.close()
    s = pvof.Solver(pf, lib)
    assert lib.pv_autotune_mg(s.h, None, None) == -1
    assert "MG" in lib.pv_last_error().decode()
