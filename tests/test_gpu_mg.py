"""Multigrid-preconditioned pressure CG (csrc/pv_mg.hip) on the GPU.

* One V-cycle of the GPU (through the pv_mg_vcycle_apply hook) against the
  NumPy restatement tests/mg_ref.py on bubbly densities, on grids whose
  levels all run in one workgroup (8 x 8, 32 x 32, 64 x 64), on grids with
  both kinds of level (64 x 128, 96 x 96, 256 x 256), for 1 and 2 sweeps,
  with the FP64 V-cycle (to 1e-11) and the FP32 one (to 1e-5).
* Pressure solves on uneven sizes (where the FFT preconditioner does not
  exist): the GPU's PHI has the stated relative residual under the NumPy
  operator, zero mean, and took few iterations.
* Whole RK3 steps with MG-PCG against FFT-PCG from the same state: both
  solve to the same relative residual, so the fields agree to that level.
"""
import ctypes as C

import numpy as np
import pytest

import mg_ref
import pvof
from helpers import regions, relerr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib(gpu):
    return pvof.load_lib()


def params(nx, ny, solver=pvof.PV_PSOLVE_MG_PCG):
    return pvof.default_params(
        NXM=nx, NYM=ny, xmax=1.0, ymax=ny / nx, Re=100.0, We=200.0, rho1=1.0, rho2=1e-3, mu1=1.0, mu2=1e-2,
        Cn=2.0 / nx, Pe_CH=nx / 2.0, cfl=0.2, default_dt=1e-4, max_dt=1e-3, ch_tol=1e-12, CG_ETOL=1e-7,
        CG_MAXIT=120, pressure_solver=solver, p_rel_tol=1e-10, p_max_iter=3000)


def canonical(p, a):
    """Cell array [ny][nx] -> canonical layout, x-ghosts periodic, y-ghosts copied."""
    out = np.zeros(pvof.shape_of(p))
    ext = np.pad(a, ((3, 4), (0, 0)), mode="edge")
    out[:, :] = np.pad(ext, ((0, 0), (3, 4)), mode="wrap")
    return out


@pytest.mark.parametrize("nx,ny", [(8, 8), (32, 32), (64, 64), (64, 128), (96, 96), (256, 256)])
@pytest.mark.parametrize("nu", [1, 2])
@pytest.mark.parametrize("fp32", [0, 1])
def test_vcycle_matches_reference(lib, nx, ny, nu, fp32):
    p = params(nx, ny)
    s = pvof.Solver(p, lib)
    s.check(lib.pv_set_mg_sweeps(s.h, nu))
    s.check(lib.pv_set_mg_precision(s.h, fp32))
    rho = mg_ref.bubbles(nx, ny, seed=nx + ny, nb=6)
    f = np.random.default_rng(nx * ny).standard_normal((ny, nx))
    f -= f.mean()
    s.set("RHO", canonical(p, rho))
    s.set("P_RES", canonical(p, f))
    s.call("mg_vcycle_apply")
    z = s.get("SCRATCH0")[regions(p)["cell"]]
    bx, by = mg_ref.faces(rho, float(nx * nx), float(nx * nx))
    lv, G = mg_ref.build(bx, by)
    zr = mg_ref.vcycle(lv, G, f, nu)
    print(f"\n{nx}x{ny} nu={nu} fp32={fp32} levels={len(lv)} relerr={relerr(z, zr):.2e}")
    assert relerr(z, zr) < (1e-5 if fp32 else 1e-11)


@pytest.mark.parametrize("nx,ny", [(64, 64), (64, 128), (256, 256), (1024, 1024)])
@pytest.mark.parametrize("wlevels", [1, 2, 3])
@pytest.mark.parametrize("fp32", [0, 1])
def test_wcycle_matches_reference(lib, nx, ny, wlevels, fp32):
    """pv_set_mg_cycle: the first wlevels in-workgroup levels take two coarse
    corrections; the GPU cycle equals the restatement's with those levels."""
    p = params(nx, ny)
    s = pvof.Solver(p, lib)
    s.check(lib.pv_set_mg_precision(s.h, fp32))
    s.check(lib.pv_set_mg_cycle(s.h, wlevels))
    rho = mg_ref.bubbles(nx, ny, seed=nx + 2 * ny, nb=8)
    f = np.random.default_rng(nx + ny).standard_normal((ny, nx))
    f -= f.mean()
    s.set("RHO", canonical(p, rho))
    s.set("P_RES", canonical(p, f))
    s.call("mg_vcycle_apply")
    z = s.get("SCRATCH0")[regions(p)["cell"]]
    bx, by = mg_ref.faces(rho, float(nx * nx), float(nx * nx))
    lv, G = mg_ref.build(bx, by)
    sizes = mg_ref.level_sizes(nx, ny)
    c0 = next(i for i, sz in enumerate(sizes) if sz[0] * sz[1] <= mg_ref.LDS_CELLS)
    wl = tuple(range(c0, min(c0 + wlevels, len(lv) - 1)))
    zr = mg_ref.vcycle(lv, G, f, 2, wl=wl)
    print(f"\n{nx}x{ny} fp32={fp32} W levels {wl} relerr={relerr(z, zr):.2e}")
    assert relerr(z, zr) < (1e-5 if fp32 else 1e-11)


@pytest.mark.parametrize("nx,ny", [(96, 96), (192, 96), (256, 512), (384, 768)])
@pytest.mark.parametrize("fp32", [0, 1])
def test_pressure_solve_residual(lib, nx, ny, fp32):
    p = params(nx, ny)
    s = pvof.Solver(p, lib)
    s.check(lib.pv_set_mg_precision(s.h, fp32))
    rho = mg_ref.bubbles(nx, ny, seed=5, nb=10)
    rng = np.random.default_rng(6)
    b = rng.standard_normal((ny, nx))
    s.set("RHO", canonical(p, rho))
    s.set("P_RHS", canonical(p, b))
    it = C.c_int(0)
    s.check(lib.pv_Pressure_solve_cg(s.h, C.byref(it)))
    phi = s.get("PHI")[regions(p)["cell"]]
    bx, by = mg_ref.faces(rho, float(nx * nx), float(nx * nx))
    bm = b - b.mean()
    res = np.linalg.norm(mg_ref.apply_a(bx, by, phi) - bm) / np.linalg.norm(bm)
    print(f"\n{nx}x{ny} fp32={fp32}: {it.value} iterations, relative residual {res:.2e}")
    assert res <= 1.05e-10
    assert abs(phi.mean()) < 1e-12 * np.abs(phi).max()
    assert 0 < it.value <= 30


@pytest.mark.parametrize("nx,ny", [(2048, 2048), (1024, 1536)])
@pytest.mark.parametrize("fp32", [0, 1])
def test_vcycle_many_tiles(lib, nx, ny, fp32):
    """Levels with more tiles than the GPU runs at once (4096 and 1536 level-0
    tiles): the V-cycle still equals the restatement, gives bit-identical
    results when applied twice, and is symmetric, <V a, b> = <a, V b> (the
    property the CG needs of its preconditioner)."""
    p = params(nx, ny)
    s = pvof.Solver(p, lib)
    s.check(lib.pv_set_mg_precision(s.h, fp32))
    rho = mg_ref.bubbles(nx, ny, seed=3, nb=32, rmin=0.02, rmax=0.06)
    s.set("RHO", canonical(p, rho))
    rng = np.random.default_rng(7)
    fs = [rng.standard_normal((ny, nx)) for _ in range(2)]
    fs = [f - f.mean() for f in fs] + [fs[0] - fs[0].mean()]
    zs = []
    for f in fs:
        s.set("P_RES", canonical(p, f))
        s.call("mg_vcycle_apply")
        zs.append(s.get("SCRATCH0")[regions(p)["cell"]].copy())
    bx, by = mg_ref.faces(rho, float(nx * nx), float(nx * nx))
    lv, G = mg_ref.build(bx, by)
    zr = mg_ref.vcycle(lv, G, fs[0], 2)
    (a, va), (b, vb) = (fs[0], zs[0]), (fs[1], zs[1])
    sym = abs((va * b).sum() - (a * vb).sum()) / abs((va * b).sum())
    print(f"\n{nx}x{ny} fp32={fp32} relerr={relerr(zs[0], zr):.2e} sym={sym:.2e}")
    assert np.array_equal(zs[0], zs[2])
    assert relerr(zs[0], zr) < (1e-5 if fp32 else 1e-11)
    assert sym < (1e-5 if fp32 else 1e-12)


def test_rk3_steps_match_fft_pcg(lib):
    n = 128
    pm = params(n, n)
    F = pvof.ic(lib, "bubbles", pm, seed=222, nmin=8, nmax=24, rmin=0.02, rmax=0.1, gas_inside=1)
    out = {}
    for name, solver, fp32 in (("fft", pvof.PV_PSOLVE_FFT_PCG, 0), ("mg", pvof.PV_PSOLVE_MG_PCG, 0),
                               ("mg32", pvof.PV_PSOLVE_MG_PCG, 1)):
        s = pvof.Solver(params(n, n, solver), lib)
        s.check(lib.pv_set_mg_precision(s.h, fp32))
        s.init_from_F(F)
        iters = []
        for _ in range(3):
            s.call("Temporal_int_rk3", 1)
            st = s.stats()
            assert max(st["p_res"]) <= 1e-10
            iters += list(st["n_p"])
        out[name] = (s, iters)
    print(f"\npressure iterations per substep: FFT-PCG {out['fft'][1]}, MG-PCG {out['mg'][1]}, "
          f"MG-PCG with the FP32 V-cycle {out['mg32'][1]}")
    assert max(out["mg"][1]) * 5 < min(out["fft"][1])
    assert max(out["mg32"][1]) * 5 < min(out["fft"][1])
    R = regions(pm)
    for name in ("mg", "mg32"):
        for f in ("P", "U", "V", "C_L"):
            assert relerr(out[name][0].get(f)[R["cell"]], out["fft"][0].get(f)[R["cell"]]) < 1e-9, (name, f)


def test_temporal_warm_start(lib):
    """Starting each solve from the same RK3 stage's last correction (kept
    only when its residual energy is lower) changes the fields only below the
    solve tolerance and never costs iterations over the cold start."""
    n = 128
    pm = params(n, n)
    F = pvof.ic(lib, "bubbles", pm, seed=223, nmin=8, nmax=24, rmin=0.02, rmax=0.1, gas_inside=1)
    out = {}
    for warm in (0, 1):
        s = pvof.Solver(pm, lib)
        s.init_from_F(F)
        s.check(lib.pv_set_pressure_warm_start(s.h, warm))
        iters, ratio = 0, C.c_double(0.0)
        for _ in range(4):
            s.call("Temporal_int_rk3", 1)
            st = s.stats()
            assert max(st["p_res"]) <= 1e-10
            iters += sum(st["n_p"])
        lib.pv_learn_start_residual(s.h, C.byref(ratio))
        out[warm] = (s, iters, ratio.value)
    print(f"\nMG-PCG iterations over 4 steps: cold {out[0][1]}, warm {out[1][1]} "
          f"(last start energy ratio {out[1][2]:.3f})")
    assert out[1][1] <= out[0][1]
    assert 0.0 < out[1][2] < 1.0
    R = regions(pm)
    for f in ("P", "U", "V", "C_L"):
        assert relerr(out[1][0].get(f)[R["cell"]], out[0][0].get(f)[R["cell"]]) < 1e-9, f


def test_momentum_warm_start(lib):
    """Starting each momentum CG from u^{k-1} plus the same RK3 stage's last
    increment (scaled by the ratio of the time steps) leaves the stop test as
    it is: the fields agree with the cold start to the solve tolerance, the
    momentum solves take fewer iterations, and resetting the option restores
    the cold start exactly."""
    n = 128
    pm = params(n, n)      # CG_ETOL 1e-7: a few iterations per solve, above the EPS-guard stall of msolve_cg.c
    F = pvof.ic(lib, "bubbles", pm, seed=227, nmin=8, nmax=24, rmin=0.02, rmax=0.1, gas_inside=1)
    out = {}
    for warm in (0, 1):
        s = pvof.Solver(pm, lib)
        s.init_from_F(F)
        s.check(lib.pv_set_momentum_warm_start(s.h, warm))
        its = []
        for _ in range(6):
            s.call("Temporal_int_rk3", 1)
            st = s.stats()
            its.append(sum(st["n_u"]) + sum(st["n_v"]))
        out[warm] = (s, its)
    print(f"\nmomentum CG iterations per step: cold {out[0][1]}, warm {out[1][1]}")
    assert out[1][1][0] == out[0][1][0]             # no increment kept yet in the first step
    assert sum(out[1][1][2:]) < sum(out[0][1][2:])
    R = regions(pm)
    for f in ("U", "V", "P", "C_L"):
        assert relerr(out[1][0].get(f)[R["cell"]], out[0][0].get(f)[R["cell"]]) < 1e-6, f
    with pytest.raises(RuntimeError, match="momentum warm start"):
        out[1][0].check(lib.pv_set_momentum_warm_start(out[1][0].h, 2))
