"""The no-allocation divergence entry point against the handle's own
divergence routine, through the C ABI, on two calls in a row (the second on
a smaller grid).  The test's device arrays come from the HIP runtime the
library itself links (plumbing only)."""
import ctypes as C

import numpy as np
import pytest

import pvof

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2                        # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def hip_runtime():
    hip = C.CDLL("libamdhip64.so.7")   # already loaded by the library: the same runtime
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return hip


def to_device(hip, a):
    a = np.ascontiguousarray(a)
    ptr = C.c_void_p()
    assert hip.hipMalloc(C.byref(ptr), a.nbytes) == 0
    assert hip.hipMemcpy(ptr, a.ctypes.data, a.nbytes, H2D) == 0
    return ptr


def divergence_pair(lib, nxm, nym, seed):
    hip = hip_runtime()
    # dx = 1/nxm and dy = 2/nym are powers of two: the handle's metric arrays
    # hold exactly idx = nxm and idy = nym / 2
    p = pvof.default_params(NXM=nxm, NYM=nym, xmax=1.0, ymax=2.0)
    s = pvof.Solver(p, lib)
    rng = np.random.default_rng(seed)
    U = rng.standard_normal(s.shape)
    V = rng.standard_normal(s.shape)
    s.set("U", U)
    s.set("V", V)
    ref = C.c_double(0.0)
    s.call("Pressure_compute_velocity_divergence", C.byref(ref))
    s.close()
    W = nxm + 7                         # the canonical layout: (NYM+7) rows of NXM+7, cell (0,0) at [3][3]
    assert U.shape == (nym + 7, W)
    u, v = to_device(hip, U), to_device(hip, V)
    out = to_device(hip, np.array([-1.0]))
    off = (3 * W + 3) * 8
    try:
        rc = lib.pv_Pressure_compute_velocity_divergence_noalloc(u.value + off, v.value + off, nxm, nym, W,
                                                                 float(nxm), float(nym) / 2.0, out, None)
        assert rc == 0, lib.pv_last_error()
        got = np.zeros(1)
        assert hip.hipMemcpy(got.ctypes.data, out, 8, D2H) == 0      # null stream: after the kernels
    finally:
        for ptr in (u, v, out):
            hip.hipFree(ptr)
    return ref.value, float(got[0])


def test_noalloc_divergence_matches_the_handle_routine(gpu):
    lib = pvof.load_lib()
    first = divergence_pair(lib, 256, 512, 1)
    second = divergence_pair(lib, 64, 32, 2)       # smaller: stale partials must not leak in
    for ref, got in (first, second):
        assert ref > 0.0
        assert got == ref, (got, ref)


def test_noalloc_rejects_sizes_beyond_its_scratch(gpu):
    lib = pvof.load_lib()
    out = C.c_double()
    big = pvof.PV_NOALLOC_MAX_N + 1
    dummy = C.c_void_p(16)
    assert lib.pv_Pressure_compute_velocity_divergence_noalloc(dummy, dummy, big, 8, big + 8, 1.0, 1.0,
                                                               dummy, None) == -1
    assert b"outside" in lib.pv_last_error()
    assert lib.pv_Pressure_compute_velocity_divergence_noalloc(dummy, dummy, 8, 8, 4, 1.0, 1.0, dummy, None) == -1
