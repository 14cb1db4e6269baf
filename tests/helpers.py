"""Shared helpers for the parity tests (regions, comparisons, state copies)."""
import numpy as np

import oracle_lib
import pvof


def regions(p):
    """Index slices of the canonical layout for each storage kind."""
    nx, ny = p.NXM, p.NYM
    return {
        "cell": (slice(3, 3 + ny), slice(3, 3 + nx)),            # i < NXM, j < NYM
        "G": (slice(3, 4 + ny), slice(3, 4 + nx)),               # the reference G range
        "frame": (slice(0, ny + 6), slice(0, nx + 6)),            # ghosts, minus the stale last row/col
        "u": (slice(3, 3 + ny), slice(4, 4 + nx)),               # u unknowns i = 1..NXM, j < NYM
        "v": (slice(4, 3 + ny), slice(3, 3 + nx)),               # v unknowns j = 1..NYM-1
    }


def relerr(a, b):
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)


def oracle_state(p, F, nsteps, lib=None):
    o = oracle_lib.Oracle(p, lib)
    o.init_from_F(F)
    if nsteps:
        rc = o.call("Temporal_int_rk3", nsteps)
        assert rc == 0
    return o


def copy_state(o, s, stage=None):
    """Copy every field and the time state of oracle o into solver s."""
    for name in pvof.FIELDS:
        s.set(name, o.get(name))
    st = o.stats()
    s.call("set_time_state", st["time"], st["dt"], st["dt_old"],
           st["which_stage"] if stage is None else stage)
    if stage is not None:
        o.call("set_time_state", st["time"], st["dt"], st["dt_old"], stage)


class RefArray:
    """A double*** laid out as the reference allocates it (Memory.c:36-96 for
    the L range, :270-330 for the no-ghost G range): one block of values
    [k][j][i], a table of row pointers and a table of slab pointers, both
    offset so that a[k0][j0][i0] is the first value.  `ptr` is the double***
    to hand to the reference-signature routines; `data` views the values as
    a numpy array [k][j][i] (index 0 = k0, j0, i0)."""

    def __init__(self, nxm, nym, ghost=True):
        import ctypes as C
        if ghost:
            self.k0, self.j0, self.i0 = -3, -3, -3
            kc, jc, ic = 8, nym + 7, nxm + 7
        else:
            self.k0, self.j0, self.i0 = 0, 0, 0
            kc, jc, ic = 2, nym + 1, nxm + 1
        self.data = np.zeros((kc, jc, ic))
        base = self.data.ctypes.data
        self.rows = (C.c_size_t * (kc * jc))()
        for k in range(kc):
            for j in range(jc):
                self.rows[k * jc + j] = base + 8 * ((k * jc + j) * ic - self.i0)
        rows_addr = C.addressof(self.rows)
        self.slabs = (C.c_size_t * kc)()
        for k in range(kc):
            self.slabs[k] = rows_addr + 8 * (k * jc - self.j0)
        self.ptr = C.c_void_p(C.addressof(self.slabs) - 8 * self.k0)

    def canonical(self, k=0):
        """Slab k as the canonical [NYM+7][NXM+7] frame (G-range arrays are
        placed at their (i, j) with zeros around)."""
        kk = k - self.k0
        if self.k0 == -3:
            return self.data[kk].copy()
        jc, ic = self.data.shape[1:]
        out = np.zeros((jc + 6, ic + 6))
        out[3:3 + jc, 3:3 + ic] = self.data[kk]
        return out

    def set_canonical(self, a):
        """Every k slab from a canonical frame."""
        for kk in range(self.data.shape[0]):
            if self.k0 == -3:
                self.data[kk] = a
            else:
                jc, ic = self.data.shape[1:]
                self.data[kk] = a[3:3 + jc, 3:3 + ic]
