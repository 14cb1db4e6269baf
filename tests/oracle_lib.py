"""ctypes binding of the CPU oracle (oracle/pvo.c) -- test infrastructure only.

The oracle is the checker: tests compare the HIP library against it, never the
other way round, and the product never imports this module.
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "parties-vof-ibm_amd"))
import pvof  # noqa: E402

ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "libpvo.so")


def build_oracle():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def load_oracle(path=None):
    """Load the oracle library (default: the one built from oracle/pvo.c)."""
    if path is None:
        path = ORACLE_SO
        if not os.path.exists(path):
            build_oracle()
    lib = C.CDLL(path)
    vp, dp = C.c_void_p, C.POINTER(C.c_double)
    lib.pvo_create.restype = vp
    lib.pvo_create.argtypes = [C.POINTER(pvof.Params)]
    lib.pvo_destroy.argtypes = [vp]
    lib.pvo_field.restype = dp
    lib.pvo_field.argtypes = [vp, C.c_int]
    lib.pvo_get_params.argtypes = [vp, C.POINTER(pvof.Params)]
    lib.pvo_init_from_F.argtypes = [vp, dp]
    lib.pvo_set_time_state.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_int]
    lib.pvo_set_momentum_cg_steps.argtypes = [vp, C.c_int]
    lib.pvo_get_stats.argtypes = [vp, C.POINTER(pvof.Stats)]
    lib.pvo_Temporal_int_rk3.argtypes = [vp, C.c_int]
    lib.pvo_Dtime_cfl.restype = C.c_double
    lib.pvo_Dtime_cfl.argtypes = [vp]
    lib.pvo_VOF_DIFFUSE_solve_implicit_biharmonic.argtypes = [vp, C.c_int, dp]
    lib.pvo_VOF_DIFFUSE_bound_liquid_fraction.restype = C.c_double
    lib.pvo_VOF_DIFFUSE_bound_liquid_fraction.argtypes = [vp]
    lib.pvo_clamp_interior.restype = C.c_double
    lib.pvo_clamp_interior.argtypes = [vp]
    lib.pvo_Pressure_compute_velocity_divergence.restype = C.c_double
    lib.pvo_Pressure_compute_velocity_divergence.argtypes = [vp]
    for name in ("Velocity_solve_cg", "Pressure_solve_cg", "Temporal_int_all_the_equations"):
        getattr(lib, "pvo_" + name).restype = C.c_int
    return lib


class Oracle:
    """Owner of one oracle state with the same field API as pvof.Solver."""

    def __init__(self, params, lib=None):
        self.lib = lib or load_oracle()
        self.h = C.c_void_p(self.lib.pvo_create(C.byref(params)))
        self.params = pvof.Params()
        self.lib.pvo_get_params(self.h, C.byref(self.params))
        self.shape = pvof.shape_of(self.params)

    def close(self):
        if self.h:
            self.lib.pvo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def view(self, name):
        ptr = self.lib.pvo_field(self.h, pvof.F[name])
        return np.ctypeslib.as_array(ptr, shape=self.shape)

    def get(self, name):
        return self.view(name).copy()

    def set(self, name, a):
        self.view(name)[...] = a

    def init_from_F(self, Fld):
        Fld = np.ascontiguousarray(Fld, dtype=np.float64)
        self.lib.pvo_init_from_F(self.h, Fld.ctypes.data_as(C.POINTER(C.c_double)))

    def stats(self):
        st = pvof.Stats()
        self.lib.pvo_get_stats(self.h, C.byref(st))
        return st.as_dict()

    def call(self, name, *args):
        return getattr(self.lib, "pvo_" + name)(self.h, *args)
