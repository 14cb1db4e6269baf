"""ctypes view of the C ABI in include/parties_vof.h.

Host-side mirror for Python callers (tests, bench.py).  The product is the C/HIP
library libparties_vof.so built in this directory; this module only loads it and
declares argument types.  Nothing here computes physics.

Reference parameter names follow PARTIES/src/IO/default.inp.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libparties_vof.so")

PV_PSOLVE_JACOBI_PCG = 0
PV_PSOLVE_FFT_PCG = 1
PV_PSOLVE_LEARNED = 2
PV_PSOLVE_LEARNED_FCG = 3
PV_PSOLVE_MG_PCG = 4
PV_PSOLVE_LEARNED_MG = 5
PV_PREC_FP64 = 0
PV_PREC_MIXED = 1
PV_NOALLOC_MAX_N = 16384
PV_VEL_TYPE_NORMAL = 0
PV_VEL_TYPE_CG = 1

FIELDS = [
    "C_L", "C_G", "RHO", "MU",
    "PSI_LG", "LAP_C", "BULK_S", "PSI",
    "CH_AUX1", "CH_AUX2", "CH_RHS_N", "CH_RHS_NM1",
    "CH_RES", "CH_DIR", "FLUX_X", "FLUX_Y",
    "FSX_NEW", "FSY_NEW", "FSX_OLD", "FSY_OLD",
    "U", "V", "U_BC", "V_BC",
    "U_EXPL", "U_EXPL_OLD", "U_IMPL", "U_VISC", "U_VISC_OLD", "U_RHS",
    "V_EXPL", "V_EXPL_OLD", "V_IMPL", "V_VISC", "V_VISC_OLD", "V_RHS",
    "U_R", "U_D", "U_AD", "V_R", "V_D", "V_AD",
    "P", "PHI", "P_RHS", "P_RES", "P_D", "P_AD", "M_INV",
    "SCRATCH0", "SCRATCH1", "SCRATCH2",
    "F",
]
F = {name: k for k, name in enumerate(FIELDS)}
PV_NFIELDS = len(FIELDS)


class Params(C.Structure):
    _fields_ = [
        ("NXM", C.c_int), ("NYM", C.c_int),
        ("xmin", C.c_double), ("xmax", C.c_double), ("ymin", C.c_double), ("ymax", C.c_double),
        ("twod_slab_thickness", C.c_double),
        ("Re", C.c_double), ("dp_dx", C.c_double),
        ("We", C.c_double), ("rho1", C.c_double), ("rho2", C.c_double),
        ("mu1", C.c_double), ("mu2", C.c_double),
        ("Cn", C.c_double), ("Pe_CH", C.c_double),
        ("weno_order", C.c_int), ("ch_iter_max", C.c_int), ("ch_tol", C.c_double),
        ("cfl", C.c_double), ("default_dt", C.c_double), ("max_dt", C.c_double),
        ("constant_dt", C.c_int),
        ("grav", C.c_double * 3), ("richardson", C.c_double),
        ("P_CG_MAXIT", C.c_int), ("P_CG_ETOL", C.c_double),
        ("CG_MAXIT", C.c_int), ("CG_ETOL", C.c_double),
        ("pressure_solver", C.c_int), ("p_rel_tol", C.c_double), ("p_max_iter", C.c_int),
        ("precision", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("time", C.c_double), ("dt", C.c_double), ("dt_old", C.c_double),
        ("ntime", C.c_int), ("which_stage", C.c_int),
        ("n_ch", C.c_int * 3), ("n_u", C.c_int * 3), ("n_v", C.c_int * 3),
        ("n_p", C.c_int * 3), ("n_pass", C.c_int * 3),
        ("ch_res", C.c_double * 3), ("p_res", C.c_double * 3),
        ("div_max", C.c_double * 3), ("clamp_delta", C.c_double * 3),
        ("p_fallback", C.c_int * 3),
        ("clamp_delta_interior", C.c_double * 3),
    ]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            val = getattr(self, name)
            out[name] = list(val) if hasattr(val, "__len__") else val
        return out


def default_params(**kw):
    """IO/default.inp defaults (2-D planar, no IBM); keyword overrides."""
    p = Params()
    p.NXM, p.NYM = 64, 64
    p.xmin, p.xmax, p.ymin, p.ymax = 0.0, 1.0, 0.0, 1.0
    p.twod_slab_thickness = -1.0
    p.Re, p.dp_dx = 100.0, 0.0
    p.We, p.rho1, p.rho2, p.mu1, p.mu2 = 1.0, 1.0, 1.0, 1.0, 1.0
    p.Cn, p.Pe_CH = -1.0, -1.0
    p.weno_order, p.ch_iter_max, p.ch_tol = 5, 300, 1e-9
    p.cfl, p.default_dt, p.max_dt, p.constant_dt = 0.5, 1e-3, 0.05, 0
    p.grav[0], p.grav[1], p.grav[2] = 0.0, -1.0, 0.0
    p.richardson = 1.0
    p.P_CG_MAXIT, p.P_CG_ETOL = 300, 1e-7
    p.CG_MAXIT, p.CG_ETOL = 120, 1e-9
    p.pressure_solver, p.p_rel_tol, p.p_max_iter = PV_PSOLVE_JACOBI_PCG, 1e-10, 2000
    p.precision = PV_PREC_FP64
    for k, v in kw.items():
        if k == "grav":
            for i in range(3):
                p.grav[i] = v[i]
        else:
            setattr(p, k, v)
    return p


def config1_params(NXM=128, NYM=256, **kw):
    """Planar rising bubble: 1x2 domain, x periodic, y walls, rho2/rho1 = 1e-3,
    mu2/mu1 = 1e-2, Re = 100, We = Bo = 200, Cn = 0.01, Pe = 1/Cn, CFL 0.2."""
    base = dict(NXM=NXM, NYM=NYM, xmax=1.0, ymax=2.0, Re=100.0, We=200.0,
                rho1=1.0, rho2=1e-3, mu1=1.0, mu2=1e-2, Cn=0.01, Pe_CH=100.0,
                cfl=0.2, default_dt=1e-4, max_dt=1e-3)
    base.update(kw)
    return default_params(**base)


def shape_of(p):
    return (p.NYM + 7, p.NXM + 7)


def load_lib(path=LIB_PATH):
    """Load the product library; raises if it has not been built."""
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} missing: run __graft_entry__.build() first")
    lib = C.CDLL(path)
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    lib.pv_last_error.restype = C.c_char_p
    lib.pv_params_default.argtypes = [C.POINTER(Params)]
    lib.pv_create.argtypes = [C.POINTER(Params), C.c_int, C.POINTER(vp)]
    lib.pv_destroy.argtypes = [vp]
    lib.pv_get_params.argtypes = [vp, C.POINTER(Params)]
    lib.pv_set_field.argtypes = [vp, C.c_int, dp]
    lib.pv_get_field.argtypes = [vp, C.c_int, dp]
    lib.pv_fill_field.argtypes = [vp, C.c_int, C.c_double]
    lib.pv_init_from_F.argtypes = [vp, dp]
    lib.pv_set_time_state.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_int]
    lib.pv_set_step_count.argtypes = [vp, C.c_int]
    lib.pv_Pressure_compute_velocity_divergence_noalloc.argtypes = [vp, vp, C.c_int, C.c_int, C.c_int, C.c_double,
                                                                     C.c_double, vp, vp]
    lib.pv_get_stats.argtypes = [vp, C.POINTER(Stats)]
    lib.pv_Temporal_int_rk3.argtypes = [vp, C.c_int]
    lib.pv_Dtime_cfl.argtypes = [vp, dp]
    lib.pv_VOF_DIFFUSE_solve_implicit_biharmonic.argtypes = [vp, C.c_int, C.POINTER(C.c_int), dp]
    lib.pv_VOF_DIFFUSE_bound_liquid_fraction.argtypes = [vp, dp]
    lib.pv_Velocity_solve_cg.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
    lib.pv_Pressure_solve_cg.argtypes = [vp, C.POINTER(C.c_int)]
    lib.pv_Pressure_compute_velocity_divergence.argtypes = [vp, dp]
    lib.pv_synchronize.argtypes = [vp]
    lib.pv_timers_enable.argtypes = [vp, C.c_int]
    lib.pv_timers_reset.argtypes = [vp]
    lib.pv_timers_read.argtypes = [vp, dp, C.POINTER(C.c_longlong)]
    lib.pv_timers_read_working.argtypes = [vp, dp, C.POINTER(C.c_longlong)]
    fp = C.POINTER(C.c_float)
    lib.pv_cnn_init.argtypes = [vp, C.c_uint64]
    lib.pv_cnn_set_params.argtypes = [vp, fp]
    lib.pv_cnn_get_params.argtypes = [vp, fp]
    lib.pv_cnn_eval.argtypes = [vp, fp, fp, fp, fp, dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    lib.pv_cnn_adam_step.argtypes = [vp, C.c_float, C.c_int]
    ip = C.POINTER(C.c_int)
    lib.pv_set_pressure_solver.argtypes = [vp, C.c_int]
    lib.pv_set_mg_sweeps.argtypes = [vp, C.c_int]
    lib.pv_set_mg_cycle.argtypes = [vp, C.c_int]
    lib.pv_mg_vcycle_apply.argtypes = [vp]
    lib.pv_set_pressure_warm_start.argtypes = [vp, C.c_int]
    lib.pv_set_mg_precision.argtypes = [vp, C.c_int]
    lib.pv_set_ch_solver.argtypes = [vp, C.c_int]
    lib.pv_set_momentum_warm_start.argtypes = [vp, C.c_int]
    lib.pv_set_momentum_cg_steps.argtypes = [vp, C.c_int]
    lib.pv_set_momentum_preconditioner.argtypes = [vp, C.c_int]
    lib.pv_set_momentum_rows.argtypes = [vp, C.c_int]
    lib.pv_set_mg_tile0.argtypes = [vp, C.c_int]
    lib.pv_set_fft_tri_shape.argtypes = [vp, C.c_int, C.c_int]
    lib.pv_autotune_tiles.argtypes = [vp, ip, ip]
    lib.pv_learn_collect_begin.argtypes = [vp, ip, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64]
    lib.pv_learn_collect_end.argtypes = [vp, ip]
    lib.pv_learn_train.argtypes = [vp, C.c_int, C.c_int, C.c_float, C.c_uint64, dp, dp]
    lib.pv_learn_start_residual.argtypes = [vp, dp]
    lib.pv_set_graphs.argtypes = [vp, C.c_int, C.c_int]
    lib.pv_autotune.argtypes = [vp, ip]
    lib.pv_batch_create.argtypes = [C.POINTER(Params), C.c_int, C.c_int, C.POINTER(vp)]
    lib.pv_batch_destroy.argtypes = [vp]
    lib.pv_batch_size.argtypes = [vp]
    lib.pv_batch_case.argtypes = [vp, C.c_int]
    lib.pv_batch_case.restype = vp
    lib.pv_batch_rk3.argtypes = [vp, C.c_int]
    lib.pv_batch_set_threads.argtypes = [vp, C.c_int]
    lib.pv_batch_autotune.argtypes = [vp, ip]
    lib.pv_autotune_mg.argtypes = [vp, ip, ip]
    lib.pv_batch_autotune_mg.argtypes = [vp, ip, ip]
    for name in ("pv_ic_rising_bubble", "pv_ic_rayleigh_taylor"):
        getattr(lib, name).argtypes = [C.POINTER(Params), dp]
    lib.pv_ic_multimode_rt.argtypes = [C.POINTER(Params), C.c_uint64, C.c_int, C.c_double, C.c_double, dp]
    lib.pv_ic_bubbles.argtypes = [C.POINTER(Params), C.c_uint64, C.c_int, C.c_int, C.c_double,
                                  C.c_double, C.c_int, dp, C.POINTER(C.c_int)]
    return lib


def _dptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Solver:
    """Thin owner of one pv_bag (one 2-D subdomain on one GPU)."""

    def __init__(self, params, lib=None, device=0):
        self.lib = lib or load_lib()
        self.h = C.c_void_p()
        self.check(self.lib.pv_create(C.byref(params), device, C.byref(self.h)))
        self.params = Params()
        self.check(self.lib.pv_get_params(self.h, C.byref(self.params)))
        self.shape = shape_of(self.params)

    def check(self, rc):
        if rc != 0:
            raise RuntimeError(f"parties_vof error {rc}: {self.lib.pv_last_error().decode()}")
        return rc

    def close(self):
        if self.h:
            self.lib.pv_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get(self, name):
        a = np.zeros(self.shape)
        self.check(self.lib.pv_get_field(self.h, F[name], _dptr(a)))
        return a

    def set(self, name, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == self.shape
        self.check(self.lib.pv_set_field(self.h, F[name], _dptr(a)))

    def init_from_F(self, Fld):
        Fld = np.ascontiguousarray(Fld, dtype=np.float64)
        self.check(self.lib.pv_init_from_F(self.h, _dptr(Fld)))

    def stats(self):
        st = Stats()
        self.check(self.lib.pv_get_stats(self.h, C.byref(st)))
        return st.as_dict()

    def call(self, name, *args):
        return self.check(getattr(self.lib, "pv_" + name)(self.h, *args))


PV_CNN_NPARAM = 38177


def _fptr(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Network:
    """Host view of the learned preconditioner network of one Solver."""

    def __init__(self, solver, seed=1):
        self.s = solver
        solver.check(solver.lib.pv_cnn_init(solver.h, seed))

    def params(self):
        a = np.zeros(PV_CNN_NPARAM, dtype=np.float32)
        self.s.check(self.s.lib.pv_cnn_get_params(self.s.h, _fptr(a)))
        return a

    def set_params(self, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        self.s.check(self.s.lib.pv_cnn_set_params(self.s.h, _fptr(a)))

    def eval(self, x, tgt=None, px=0, border=0, want_grad=False):
        """x [B][H][W][3] -> out [B][H][W]; with tgt also (loss, gradient)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        B, H, W, _ = x.shape
        out = np.zeros((B, H, W), dtype=np.float32)
        loss = C.c_double(0.0)
        grad = np.zeros(PV_CNN_NPARAM, dtype=np.float32) if want_grad else None
        t = None if tgt is None else np.ascontiguousarray(tgt, dtype=np.float32)
        null = C.POINTER(C.c_float)()
        self.s.check(self.s.lib.pv_cnn_eval(self.s.h, _fptr(x), _fptr(t) if t is not None else null, _fptr(out),
                                            _fptr(grad) if grad is not None else null, C.byref(loss), B, H, W,
                                            px, border))
        if tgt is None:
            return out
        return out, loss.value, grad

    def adam_step(self, lr, t):
        self.s.check(self.s.lib.pv_cnn_adam_step(self.s.h, lr, t))

    def collect(self, snap_iters, crops_per_snap=32, crop_size=64, max_samples=20000, seed=1):
        """Start collecting training crops from the handle's plain FFT-PCG solves."""
        it = (C.c_int * len(snap_iters))(*snap_iters)
        self.s.check(self.s.lib.pv_learn_collect_begin(self.s.h, it, len(snap_iters), crops_per_snap, crop_size,
                                                       max_samples, seed))

    def collect_end(self):
        n = C.c_int(0)
        self.s.check(self.s.lib.pv_learn_collect_end(self.s.h, C.byref(n)))
        return n.value

    def train(self, steps, batch=32, lr=1e-3, seed=1):
        l0, l1 = C.c_double(0.0), C.c_double(0.0)
        self.s.check(self.s.lib.pv_learn_train(self.s.h, steps, batch, lr, seed, C.byref(l0), C.byref(l1)))
        return l0.value, l1.value


class Batch:
    """Independent cases of any size advanced together (one handle and stream each)."""

    def __init__(self, params_list, lib=None, device=0):
        self.lib = lib or load_lib()
        arr = (Params * len(params_list))(*params_list)
        self.h = C.c_void_p()
        rc = self.lib.pv_batch_create(arr, len(params_list), device, C.byref(self.h))
        if rc != 0:
            raise RuntimeError(f"parties_vof error {rc}: {self.lib.pv_last_error().decode()}")
        self.cases = [_CaseView(self.lib, self.lib.pv_batch_case(self.h, k), params_list[k])
                      for k in range(len(params_list))]

    def check(self, rc):
        if rc != 0:
            raise RuntimeError(f"parties_vof error {rc}: {self.lib.pv_last_error().decode()}")

    def rk3(self, nsteps):
        self.check(self.lib.pv_batch_rk3(self.h, nsteps))

    def set_threads(self, nthreads):
        self.check(self.lib.pv_batch_set_threads(self.h, nthreads))

    def autotune(self):
        ch = (C.c_int * len(self.cases))()
        self.check(self.lib.pv_batch_autotune(self.h, ch))
        return list(ch)

    def autotune_mg(self):
        """(W levels, sweeps) chosen per case; (-1, -1) for a case without MG-PCG."""
        w, nu = (C.c_int * len(self.cases))(), (C.c_int * len(self.cases))()
        self.check(self.lib.pv_batch_autotune_mg(self.h, w, nu))
        return list(zip(w, nu))

    def close(self):
        if self.h:
            self.lib.pv_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _CaseView(Solver):
    """A Solver view of one batch case (owned by the batch, not destroyed here)."""

    def __init__(self, lib, handle, params):
        self.lib = lib
        self.h = C.c_void_p(handle)
        self.params = Params()
        self.check(self.lib.pv_get_params(self.h, C.byref(self.params)))
        self.shape = shape_of(self.params)

    def close(self):
        self.h = C.c_void_p()


def ic(lib, kind, params, **kw):
    """Initial volume-fraction field F on the canonical layout."""
    Fld = np.zeros(shape_of(params))
    if kind == "rising_bubble":
        lib.pv_ic_rising_bubble(C.byref(params), _dptr(Fld))
    elif kind == "rayleigh_taylor":
        lib.pv_ic_rayleigh_taylor(C.byref(params), _dptr(Fld))
    elif kind == "multimode_rt":
        lib.pv_ic_multimode_rt(C.byref(params), kw.get("seed", 20261202), kw.get("modes", 8),
                               kw.get("amp_max", 0.05), kw.get("y0", 2.0), _dptr(Fld))
    elif kind == "bubbles":
        n = C.c_int(0)
        lib.pv_ic_bubbles(C.byref(params), kw.get("seed", 20261202), kw.get("nmin", 16),
                          kw.get("nmax", 64), kw.get("rmin", 0.01), kw.get("rmax", 0.08),
                          kw.get("gas_inside", 1), _dptr(Fld), C.byref(n))
    else:
        raise ValueError(kind)
    return Fld


SHIM_PATH = os.path.join(HERE, "_build", "libparties_vof_shim.so")
PVS_L_RANGE, PVS_G_RANGE = 0, 1
FATAL_FN = C.CFUNCTYPE(None, C.c_char_p, C.c_void_p)


def load_shim(path=SHIM_PATH):
    """Load the reference-signature layer (include/parties_vof_shim.h)."""
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} missing: run __graft_entry__.build() first")
    sh = C.CDLL(path)
    vp = C.c_void_p
    sh.pvs_last_error.restype = C.c_char_p
    sh.pvs_bind.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    sh.pvs_bind_field.argtypes = [vp, vp, C.c_int, C.c_int]
    sh.pvs_unbind.argtypes = [vp]
    sh.pvs_handle.argtypes = [vp]
    sh.pvs_handle.restype = vp
    sh.pvs_push.argtypes = [vp, vp]
    sh.pvs_pull.argtypes = [vp, vp]
    sh.pvs_set_time_max.argtypes = [vp, C.c_double]
    sh.pvs_set_fatal.argtypes = [FATAL_FN, vp]
    sh.Temporal_int_rk3.argtypes = [vp, vp]
    sh.Temporal_int_all_the_equations.argtypes = [vp, vp]
    sh.Dtime_cfl.argtypes = [vp]
    sh.Dtime_cfl.restype = C.c_double
    sh.Communication_update_ghost_nodes_flow_variable.argtypes = [vp, C.c_char, C.c_int, vp]
    for name in ("VOF_DIFFUSE_set_boundary_values",):
        getattr(sh, name).argtypes = [vp, vp]
    sh.VOF_DIFFUSE_compute_laplacian.argtypes = [vp, vp, vp]
    for name in ("VOF_DIFFUSE_advect_WENO5", "VOF_DIFFUSE_explicit_diffusion",
                 "VOF_DIFFUSE_solve_implicit_biharmonic"):
        getattr(sh, name).argtypes = [vp, vp]
    sh.Velocity_update_boundaries.argtypes = [vp, C.c_char, C.c_int, vp]
    sh.Velocity_solve_cg.argtypes = [vp, vp]
    sh.Pressure_solve_cg.argtypes = [vp]
    sh.Pressure_operator_variableCoeff.argtypes = [vp, vp, vp, vp, vp, vp]
    sh.Pressure_apply_BCs.argtypes = [vp, vp, vp]
    sh.Pressure_compute_velocity_divergence.argtypes = [vp]
    sh.Pressure_compute_velocity_divergence.restype = C.c_double
    sh.Pressure_solve_x_period.argtypes = [vp, vp, vp]
    for name in ("VOF_DIFFUSE_init", "VOF_DIFFUSE_update_phase_cache", "VOF_DIFFUSE_compute_bulk_S",
                 "VOF_DIFFUSE_compute_psi_LG", "VOF_DIFFUSE_step", "VOF_DIFFUSE_compute_f_sigma",
                 "VOF_DIFFUSE_apply_f_sigma_old", "VOF_DIFFUSE_update_density_viscosity",
                 "VOF_update_density_viscosity", "VOF_apply_f_sigma_old", "VOF_swap_f_sigma",
                 "Velocity_add_gravity_2_RHS", "Velocity_cell_center", "Velocity_u_set_implicit_explicit",
                 "Velocity_v_set_implicit_explicit", "Velocity_w_set_implicit_explicit", "Velocity_u_set_RHS",
                 "Velocity_v_set_RHS", "Velocity_w_set_RHS", "Pressure_set_RHS",
                 "Pressure_compute_preconditioner", "Pressure_project_velocity"):
        getattr(sh, name).argtypes = [vp]
    return sh


IO_PATH = os.path.join(HERE, "_build", "libparties_vof_io.so")
PV_IO_NTIMER = 34


class IOCounters(C.Structure):
    _fields_ = [("noutput", C.c_int), ("noutput_2d", C.c_int), ("output_time", C.c_double),
                ("output_time_2d", C.c_double), ("timer", C.c_double * PV_IO_NTIMER)]


def load_io(path=IO_PATH):
    """Load the HDF5 Data / Resume module (include/parties_vof_io.h)."""
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} missing: run __graft_entry__.build() first")
    io = C.CDLL(path)
    vp = C.c_void_p
    io.pv_io_last_error.restype = C.c_char_p
    io.pv_Output_h5_data.argtypes = [vp, C.c_char_p]
    io.pv_Output_h5_resume.argtypes = [vp, C.c_char_p, C.POINTER(IOCounters)]
    io.pv_Resume_h5.argtypes = [vp, C.c_char_p, C.c_char_p, C.POINTER(IOCounters)]
    return io
