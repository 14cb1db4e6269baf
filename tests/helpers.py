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
