"""NumPy restatement of the multigrid preconditioner of the MG-PCG pressure
solve (test infrastructure; the product is csrc/pv_mg.hip).

The operator is the variable-coefficient pressure operator of the PCG
(psolve_cg.c:302-344 with the face coefficients 2 / (rho_a + rho_b) of
psolve_cg.c:63-101), x periodic, zero-gradient walls in y:
    (A x)_ij = bx_ij (x_E - x) - bx_(i-1)j (x - x_W) + by_ij (x_N - x) - by_i(j-1) (x - x_S).
One V-cycle V(f), arrays indexed [j][i]:
  * levels halve both sizes while a level has more than 4096 cells, then
    while both sizes are even and the level has more than 256 cells or more
    than 64 columns; the last level must have at most 256 cells and 64
    columns; coarse faces are half the sum of the two fine faces across the
    coarse face;
  * nu red-black Gauss-Seidel sweeps (red = (i + j) even first) from x = 0,
    restriction by the sum of the four fine residuals, the coarse V-cycle,
    piecewise-constant prolongation, nu sweeps black first; on the levels in
    `wl` that is done twice (the second pre-smoothing from the current x);
  * the last level: x = -P G P f, G the inverse of -A with its (0, 0) entry
    doubled, P the projection to zero mean.
`mutant` plants one plausible mistake (for the tests that show the checks
catch it): "post_red_first", "restrict_mean", "coarse_sum".
"""
import numpy as np

LDS_CELLS = 4096
BOTTOM_MIN = 256
BOTTOM = 256
BOTTOM_NX = 64


def faces(rho, idx2, idy2):
    """Face coefficients of the pressure operator from a cell density [ny][nx]."""
    bx = 2.0 / (rho + np.roll(rho, -1, axis=1)) * idx2
    by = np.zeros_like(rho)
    by[:-1] = 2.0 / (rho[:-1] + rho[1:]) * idy2
    return bx, by


def apply_a(bx, by, x):
    out = bx * (np.roll(x, -1, 1) - x) - np.roll(bx, 1, 1) * (x - np.roll(x, 1, 1))
    dy = by[:-1] * (x[1:] - x[:-1])
    out[:-1] += dy
    out[1:] -= dy
    return out


def level_sizes(nx, ny):
    sizes = [(nx, ny)]
    while ((sizes[-1][0] * sizes[-1][1] > BOTTOM_MIN or sizes[-1][0] > BOTTOM_NX)
           and sizes[-1][0] % 2 == 0 and sizes[-1][1] % 2 == 0):
        sizes.append((sizes[-1][0] // 2, sizes[-1][1] // 2))
    if sizes[-1][0] * sizes[-1][1] > BOTTOM or sizes[-1][0] > BOTTOM_NX:
        raise ValueError(f"{nx} x {ny} does not halve to a last level of <= {BOTTOM} cells and <= {BOTTOM_NX} columns")
    return sizes


def build(bx, by, mutant=None):
    ny, nx = bx.shape
    levels = [(bx, by)]
    scale = 1.0 if mutant == "coarse_sum" else 0.5
    for _ in level_sizes(nx, ny)[1:]:
        fx, fy = levels[-1]
        levels.append((scale * (fx[0::2, 1::2] + fx[1::2, 1::2]), scale * (fy[1::2, 0::2] + fy[1::2, 1::2])))
    bxb, byb = levels[-1]
    M = -dense_operator(bxb, byb)
    M[0, 0] += M[0, 0]
    return levels, np.linalg.inv(M)


def dense_operator(bx, by):
    ny, nx = bx.shape
    n = nx * ny
    A = np.zeros((n, n))
    for c in range(n):
        e = np.zeros(n)
        e[c] = 1.0
        A[:, c] = apply_a(bx, by, e.reshape(ny, nx)).ravel()
    return A


def sweep(bx, by, x, f, colour):
    ny, nx = x.shape
    cw = np.roll(bx, 1, 1)
    cn = np.zeros_like(by)
    cn[:-1] = by[:-1]
    cs = np.zeros_like(by)
    cs[1:] = by[:-1]
    xn = np.zeros_like(x)
    xn[:-1] = x[1:]
    xs = np.zeros_like(x)
    xs[1:] = x[:-1]
    v = (bx * np.roll(x, -1, 1) + cw * np.roll(x, 1, 1) + cn * xn + cs * xs - f) / (bx + cw + cn + cs)
    j, i = np.indices(x.shape)
    m = (i + j) % 2 == colour
    x = x.copy()
    x[m] = v[m]
    return x


def vcycle(levels, G, f, nu=2, k=0, mutant=None, wl=()):
    """One cycle; the levels in wl take two coarse corrections with smoothing
    between them (pre, correct, post, pre, correct, post)."""
    bx, by = levels[k]
    if k == len(levels) - 1:
        fm = f - f.mean()
        x = -(G @ fm.ravel()).reshape(f.shape)
        return x - x.mean()
    x = np.zeros_like(f)
    first, second = (0, 1) if mutant == "post_red_first" else (1, 0)
    for rep in range(2 if k in wl else 1):
        for _ in range(nu):
            x = sweep(bx, by, x, f, 0)
            x = sweep(bx, by, x, f, 1)
        r = f - apply_a(bx, by, x)
        fc = r[0::2, 0::2] + r[1::2, 0::2] + r[0::2, 1::2] + r[1::2, 1::2]
        if mutant == "restrict_mean":
            fc = 0.25 * fc
        xc = vcycle(levels, G, fc, nu, k + 1, mutant, wl)
        x = x + np.repeat(np.repeat(xc, 2, 0), 2, 1)
        for _ in range(nu):
            x = sweep(bx, by, x, f, first)
            x = sweep(bx, by, x, f, second)
    return x


def pcg(bx, by, b, precond, tol=1e-10, maxit=3000):
    """CG on A x = b - mean(b) with the preconditioner z = precond(r);
    returns (x, iterations) at ||r|| <= tol ||b - mean||."""
    r = b - b.mean()
    x = np.zeros_like(b)
    bn = np.linalg.norm(r)
    z = precond(r)
    d = z.copy()
    rz = (r * z).sum()
    for k in range(maxit):
        ad = apply_a(bx, by, d)
        a = rz / (d * ad).sum()
        x += a * d
        r -= a * ad
        if np.linalg.norm(r) <= tol * bn:
            return x, k + 1
        z = precond(r)
        rz2 = (r * z).sum()
        d = z + rz2 / rz * d
        rz = rz2
    return x, maxit


def bubbles(nx, ny, seed, nb=12, rmin=0.05, rmax=0.15, ratio=1e-3):
    """Cell density of a field of gas bubbles (tanh profile) in liquid."""
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / nx
    X, Y = np.meshgrid(x, y)
    C = np.ones((ny, nx))
    for _ in range(nb):
        cx, cy, r = rng.uniform(0, 1), rng.uniform(0.1, 0.9) * ny / nx, rng.uniform(rmin, rmax)
        dx = np.minimum(abs(X - cx), 1 - abs(X - cx))
        C = np.minimum(C, 0.5 * (1 + np.tanh((np.sqrt(dx ** 2 + (Y - cy) ** 2) - r) * nx / 4)))
    return C + (1 - C) * ratio
