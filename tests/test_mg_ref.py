"""The NumPy restatement of the multigrid preconditioner (tests/mg_ref.py),
pinned by what the mathematics fixes rather than by retyping it:

* one V-cycle is a symmetric operator (pre- and post-smoothing are adjoint,
  restriction is the transpose of prolongation, the last level is exact), so
  CG preconditioned by it is plain CG;
* on a grid that is its own last level the V-cycle is the exact inverse of A
  on mean-free right-hand sides;
* MG-PCG solves the bubbly variable-density problem in a handful of
  iterations, far fewer than Jacobi-PCG, and to the right answer.
Each check fails on one of the planted mistakes of mg_ref (mutant=...).
"""
import numpy as np
import pytest

import mg_ref as M


def problem(n, seed=3):
    rho = M.bubbles(n, n, seed)
    bx, by = M.faces(rho, float(n * n), float(n * n))
    return rho, bx, by


def asymmetry(bx, by, mutant=None, seed=1, wl=()):
    n = bx.shape[0]
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n))
    b = rng.standard_normal((n, n))
    a -= a.mean()
    b -= b.mean()
    lv, G = M.build(bx, by, mutant)
    va, vb = M.vcycle(lv, G, a, mutant=mutant, wl=wl), M.vcycle(lv, G, b, mutant=mutant, wl=wl)
    return abs((va * b).sum() - (a * vb).sum()) / abs((va * b).sum())


def test_vcycle_is_symmetric():
    _, bx, by = problem(64)
    assert asymmetry(bx, by) < 1e-12
    assert asymmetry(bx, by, "post_red_first") > 1e-6


def w_without_middle_smoothing(levels, G, f, k, wl):
    """Two coarse corrections with no smoothing between them: the
    textbook W, which with these over-correcting coarse
    levels converges worse than the V-cycle (the smoothing between the
    corrections in mg_ref is what makes the W pay)."""
    bx, by = levels[k]
    if k == len(levels) - 1:
        fm = f - f.mean()
        x = -(G @ fm.ravel()).reshape(f.shape)
        return x - x.mean()
    x = np.zeros_like(f)
    for _ in range(2):
        x = M.sweep(bx, by, x, f, 0)
        x = M.sweep(bx, by, x, f, 1)
    for _ in range(2 if k in wl else 1):
        r = f - M.apply_a(bx, by, x)
        fc = r[0::2, 0::2] + r[1::2, 0::2] + r[0::2, 1::2] + r[1::2, 1::2]
        x = x + np.repeat(np.repeat(w_without_middle_smoothing(levels, G, fc, k + 1, wl), 2, 0), 2, 1)
    for _ in range(2):
        x = M.sweep(bx, by, x, f, 1)
        x = M.sweep(bx, by, x, f, 0)
    return x


def test_wcycle_symmetric_and_fewer_iterations():
    """The W on the in-workgroup levels (pre, correct, post, pre, correct,
    post) is symmetric, equals the V-cycle when no level takes it, and cuts
    the MG-PCG iterations of a bubbly problem."""
    n = 256
    _, bx, by = problem(n, seed=5)
    lv, G = M.build(bx, by)
    sizes = M.level_sizes(n, n)
    c0 = next(i for i, s in enumerate(sizes) if s[0] * s[1] <= M.LDS_CELLS)
    wl = (c0, c0 + 1)
    assert asymmetry(bx, by, wl=wl) < 1e-12
    f = np.random.default_rng(6).standard_normal((n, n))
    f -= f.mean()
    assert np.array_equal(M.vcycle(lv, G, f, wl=()), M.vcycle(lv, G, f))
    b = np.random.default_rng(7).standard_normal((n, n))
    _, itv = M.pcg(bx, by, b, lambda r: M.vcycle(lv, G, r), tol=1e-10)
    _, itw = M.pcg(bx, by, b, lambda r: M.vcycle(lv, G, r, wl=wl), tol=1e-10)
    _, itt = M.pcg(bx, by, b, lambda r: w_without_middle_smoothing(lv, G, r, 0, wl), tol=1e-10)
    assert itw < itv < itt


def test_last_level_alone_is_the_exact_inverse():
    _, bx, by = problem(8)
    lv, G = M.build(bx, by)
    assert len(lv) == 1
    f = np.random.default_rng(2).standard_normal((8, 8))
    x = M.vcycle(lv, G, f)
    assert abs(x.mean()) < 1e-14 * np.abs(x).max()
    r = M.apply_a(bx, by, x) - (f - f.mean())
    assert np.abs(r).max() < 1e-12 * np.abs(f).max()


def test_level_sizes():
    assert M.level_sizes(1024, 1024)[-1] == (16, 16)
    assert M.level_sizes(96, 96)[-1] == (12, 12)
    assert M.level_sizes(384, 768)[-1] == (6, 12)
    assert M.level_sizes(32, 64)[-1] == (8, 16)
    assert M.level_sizes(2048, 32)[-1] == (64, 1)         # 128 x 2 has 256 cells but is too wide
    with pytest.raises(ValueError):
        M.level_sizes(100, 100)
    with pytest.raises(ValueError):
        M.level_sizes(256, 2)                              # ends at 128 x 1: too wide


@pytest.mark.parametrize("mutant", [None, "restrict_mean", "coarse_sum"])
def test_mg_pcg_iterations_and_answer(mutant):
    n = 128                 # two levels above the 16 x 16 bottom: enough coarsening for the mutants to show
    _, bx, by = problem(n)
    rng = np.random.default_rng(4)
    xs = rng.standard_normal((n, n))
    xs -= xs.mean()
    b = M.apply_a(bx, by, xs)
    lv, G = M.build(bx, by, mutant)
    x, it = M.pcg(bx, by, b, lambda r: M.vcycle(lv, G, r, mutant=mutant), tol=1e-12)
    diag = bx + np.roll(bx, 1, 1)
    diag[:-1] += by[:-1]
    diag[1:] += by[:-1]
    _, itj = M.pcg(bx, by, b, lambda r: -r / diag, tol=1e-12)
    if mutant is None:
        assert it <= 12 and itj > 10 * it
        assert np.abs(x - x.mean() - xs).max() < 1e-8 * np.abs(xs).max()
    else:
        assert it > 2 * 12
