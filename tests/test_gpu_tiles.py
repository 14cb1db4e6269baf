"""Tile shapes chosen at run time (pv_set_momentum_rows, pv_set_mg_tile0)
and the tile autotuner (pv_autotune_tiles) on the GPU.

* The multigrid's level-0 tile shape changes nothing the V-cycle computes:
  its tiles carry a halo as wide as the sweeps reach, so every shape gives
  the result of sweeping the whole level, bitwise; the V-cycle output is
  compared across the shapes and against the NumPy restatement.
* Rows per momentum block change only the order of the CG's partial sums:
  steps from the same state take iteration counts within one of each other
  and give fields equal to the solves' tolerance.
* The autotuner returns one of the candidates, leaves u, v and PHI as they
  were, and the handle then steps on.
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


def params(nx, ny, **kw):
    base = dict(NXM=nx, NYM=ny, xmax=1.0, ymax=ny / nx, Re=100.0, We=200.0, rho1=1.0, rho2=1e-3, mu1=1.0, mu2=1e-2,
                Cn=2.0 / nx, Pe_CH=nx / 2.0, cfl=0.2, default_dt=1e-4, max_dt=1e-3, ch_tol=1e-12, CG_ETOL=1e-7,
                CG_MAXIT=120, pressure_solver=pvof.PV_PSOLVE_MG_PCG, p_rel_tol=1e-10, p_max_iter=3000)
    base.update(kw)
    return pvof.default_params(**base)


def canonical(p, a):
    out = np.zeros(pvof.shape_of(p))
    ext = np.pad(a, ((3, 4), (0, 0)), mode="edge")
    out[:, :] = np.pad(ext, ((0, 0), (3, 4)), mode="wrap")
    return out


@pytest.mark.parametrize("fp32", [0, 1])
def test_mg_tile0_shapes_give_the_same_vcycle(lib, fp32):
    nx, ny = 256, 512                       # level 0 and three more levels tiled
    p = params(nx, ny)
    rho = mg_ref.bubbles(nx, ny, seed=31, nb=8)
    f = np.random.default_rng(37).standard_normal((ny, nx))
    f -= f.mean()
    out = {}
    for shape in (-1, 0, 1, 2):
        s = pvof.Solver(p, lib)
        s.check(lib.pv_set_mg_precision(s.h, fp32))
        s.check(lib.pv_set_mg_tile0(s.h, shape))
        s.set("RHO", canonical(p, rho))
        s.set("P_RES", canonical(p, f))
        s.call("mg_vcycle_apply")
        out[shape] = s.get("SCRATCH0")[regions(p)["cell"]]
        s.close()
    for shape in (0, 1, 2):
        assert np.array_equal(out[shape], out[-1]), shape
    bx, by = mg_ref.faces(rho, float(nx * nx), float(nx * nx))
    lv, G = mg_ref.build(bx, by)
    assert relerr(out[-1], mg_ref.vcycle(lv, G, f, 2)) < (1e-5 if fp32 else 1e-11)


def test_momentum_rows_and_mg_tiles_in_steps(lib):
    """Three RK3 steps with every momentum row count and level-0 tile shape:
    iteration counts within one of the defaults' (the stop tests see sums in
    another order), fields equal to the solves' tolerance, all solves
    converged."""
    n = 256
    p = params(n, n, precision=pvof.PV_PREC_MIXED, CG_ETOL=1e-9)
    F = pvof.ic(lib, "bubbles", p, seed=41, nmin=8, nmax=24, rmin=0.02, rmax=0.1, gas_inside=1)
    runs = {}
    for rows, shape in ((2, -1), (1, -1), (4, -1), (2, 1), (2, 2)):
        s = pvof.Solver(p, lib)
        s.init_from_F(F)
        s.check(lib.pv_set_momentum_cg_steps(s.h, 1))
        s.check(lib.pv_set_momentum_rows(s.h, rows))
        s.check(lib.pv_set_mg_tile0(s.h, shape))
        its = []
        for _ in range(3):
            s.call("Temporal_int_rk3", 1)
            st = s.stats()
            assert max(st["p_res"]) <= 1e-10
            assert max(st["n_u"]) < p.CG_MAXIT and max(st["n_v"]) < p.CG_MAXIT
            its.append((tuple(st["n_u"]), tuple(st["n_v"])))
        runs[rows, shape] = (s, its)
    ref, ref_its = runs[2, -1]
    R = regions(p)
    for key, (s, its) in runs.items():
        for (a, b), (ra, rb) in zip(its, ref_its):
            assert max(abs(x - y) for x, y in zip(a + b, ra + rb)) <= 1, (key, its, ref_its)
        for f in ("U", "V", "P", "C_L"):
            assert relerr(s.get(f)[R["cell"]], ref.get(f)[R["cell"]]) < 1e-6, (key, f)
    for s, _ in runs.values():
        s.close()


def test_autotune_tiles_keeps_the_state(lib):
    n = 256
    p = params(n, n, precision=pvof.PV_PREC_MIXED)
    F = pvof.ic(lib, "bubbles", p, seed=43, nmin=8, nmax=24, rmin=0.02, rmax=0.1, gas_inside=1)
    s = pvof.Solver(p, lib)
    s.init_from_F(F)
    s.call("Temporal_int_rk3", 1)
    s.call("synchronize")
    before = {f: s.get(f).copy() for f in ("U", "V", "PHI", "P", "C_L")}
    rows, shape = C.c_int(-9), C.c_int(-9)
    s.check(lib.pv_autotune_tiles(s.h, C.byref(rows), C.byref(shape)))
    print(f"\nautotune at {n}^2: momentum rows {rows.value}, MG level-0 tile {shape.value}")
    assert rows.value in (1, 2, 4)
    assert shape.value in (0, 1, 2)
    for f, a in before.items():
        assert np.array_equal(s.get(f), a), f
    s.call("Temporal_int_rk3", 2)
    st = s.stats()
    assert max(st["p_res"]) <= 1e-10
    for bad in ((lib.pv_set_momentum_rows, 3), (lib.pv_set_mg_tile0, 3), (lib.pv_set_mg_tile0, -2)):
        with pytest.raises(RuntimeError):
            s.check(bad[0](s.h, bad[1]))
    s.close()


def test_fft_tri_shapes_give_the_same_solve(lib):
    """The FFT preconditioner's y solves (partitioned Thomas) in several
    segment lengths and block widths: one RK3 step of the plain FP64 path
    each, pressure iteration counts within a few, fields equal to the
    solve's tolerance; shapes
    that do not fit are rejected and leave the handle as it was."""
    n = 256
    p = params(n, n, pressure_solver=pvof.PV_PSOLVE_FFT_PCG, CG_ETOL=1e-9)
    F = pvof.ic(lib, "bubbles", p, seed=47, nmin=8, nmax=24, rmin=0.02, rmax=0.1, gas_inside=1)
    R = regions(p)
    runs = {}
    for shape in ((0, 0), (16, 8), (8, 2), (4, 2), (3, 64), (16, 1)):
        s = pvof.Solver(p, lib)
        s.init_from_F(F)
        s.check(lib.pv_set_fft_tri_shape(s.h, *shape))
        s.call("Temporal_int_rk3", 1)
        st = s.stats()
        assert max(st["p_res"]) <= 1e-10
        runs[shape] = (st["n_p"], {f: s.get(f)[R["cell"]].copy() for f in ("U", "V", "P", "C_L")})
        if shape == (0, 0):
            for bad in ((17, 2), (-1, 2), (8, 3), (8, 128)):
                with pytest.raises(RuntimeError):
                    s.check(lib.pv_set_fft_tri_shape(s.h, *bad))
            s.call("Temporal_int_rk3", 1)          # the default shape is still in place
            assert max(s.stats()["p_res"]) <= 1e-10
        s.close()
    ref_its, ref = runs[0, 0]
    for shape, (its, fields) in runs.items():
        # ~300 iterations a solve: rounding moves the count by a few
        assert max(abs(a - b) for a, b in zip(its, ref_its)) <= 4, (shape, its, ref_its)
        for f, a in fields.items():
            assert relerr(a, ref[f]) < 1e-9, (shape, f, relerr(a, ref[f]))
