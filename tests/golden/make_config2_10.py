"""Writes config2_10_steps.npz: the oracle's (oracle/pvo.c) state after 10
RK3 steps of the plain configuration at scale (512 x 1024 cells, heavy fluid
above a seeded 8-mode interface, Atwood 0.5, Re 3000,
Cn 0.005, viscosity ratio 1, FP64, FFT-PCG pressure solve to 1e-13 relative,
CH solve to 1e-13), for the GPU test of the plain path at that size
(tests/test_gpu_parity.py::test_config2_10_steps).  The oracle needs about
4 minutes for it on one core; the fixture keeps every 16th cell of C_L, U, V
and P in each direction, the maximum magnitude and the L2 norm of each full
field and the time.  The domain is 1 x 2 with the interface at y = 1, not
1 x 4 at y = 2: the reference's pressure right-hand side for a uniform grid
scales every direction by the x spacing (Pressure.c:154-156), so it needs
square cells, and 1024 rows on a height of 4 would make them 1:2 (the
re-solve loop of Temporal_int.c:1051-1107 then never reaches its divergence
bound, in the oracle as in the reference).  Run from the repository root:
    python tests/golden/make_config2_10.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, ".."))
sys.path.insert(0, os.path.join(HERE, "..", "..", "parties-vof-ibm_amd"))
import oracle_lib  # noqa: E402
import pvof  # noqa: E402
from helpers import regions  # noqa: E402

STEPS = 10
STRIDE = 16
FIELDS = (("C_L", "cell"), ("U", "u"), ("V", "v"), ("P", "cell"))


def config2(**kw):
    base = dict(NXM=512, NYM=1024, xmax=1.0, ymax=2.0, Re=3000.0, We=200.0, rho1=1.0, rho2=1.0 / 3.0,
                mu1=1.0, mu2=1.0, Cn=0.005, Pe_CH=200.0, cfl=0.2, default_dt=1e-4, max_dt=1e-3,
                pressure_solver=pvof.PV_PSOLVE_FFT_PCG, p_rel_tol=1e-13, ch_tol=1e-13, ch_iter_max=1000)
    base.update(kw)
    return pvof.default_params(**base)


def initial_field(lib, p):
    return pvof.ic(lib, "multimode_rt", p, seed=20261202, modes=8, amp_max=0.05, y0=1.0)


def main():
    lib = pvof.load_lib()
    p = config2()
    o = oracle_lib.Oracle(p)
    o.init_from_F(initial_field(lib, p))
    R = regions(p)
    for k in range(STEPS):
        assert o.call("Temporal_int_rk3", 1) == 0
        print(f"oracle: {k + 1} steps", flush=True)
    out = {"steps": np.int64(STEPS), "time": np.float64(o.stats()["time"])}
    for f, r in FIELDS:
        a = o.get(f)[R[r]]
        out[f + "_sub"] = a[::STRIDE, ::STRIDE].copy()
        out[f + "_max"] = np.float64(np.abs(a).max())
        out[f + "_l2"] = np.float64(np.sqrt((a * a).sum()))
    np.savez_compressed(os.path.join(HERE, "config2_10_steps.npz"), **out)
    print("wrote config2_10_steps.npz", flush=True)


if __name__ == "__main__":
    main()
