"""od_conservation_sums_noalloc (the allocation-free entry point) against
od_conservation_sums on the same states: bitwise equal, on consecutive calls
that share the library's fixed partial-sum array, the later ones on smaller
meshes; meshes above the fixed maximum are refused before any launch.

The test plays the caller that owns the device memory: it allocates the
state and the output with the HIP runtime the library itself is linked
against (a second HIP runtime in the process, such as torch's, would not
share its device pointers)."""
import ctypes

import numpy as np
import pytest

import opendynamo_amd as od
from helpers import smooth_state

pytestmark = pytest.mark.gpu

TWO_PI = 2 * np.pi
H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    """libamdhip64 as loaded by the library (same runtime, same pointers)."""
    od.lib()
    with open("/proc/self/maps") as f:
        paths = {line.split()[-1] for line in f if "libamdhip64.so" in line}
    assert paths, "the library did not load libamdhip64"
    hip = ctypes.CDLL(sorted(paths)[0])
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


class DeviceBuffer:
    def __init__(self, hip, nbytes):
        self.hip, self.ptr = hip, ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), nbytes) == 0

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.hip.hipFree(self.ptr)


def _noalloc(hip, p, state, real):
    with DeviceBuffer(hip, state.nbytes) as d_state, DeviceBuffer(hip, 5 * 8) as d_out:
        nan = np.full(5, np.nan)
        assert hip.hipMemcpy(d_out.ptr, nan.ctypes.data, 40, H2D) == 0
        assert hip.hipMemcpy(d_state.ptr, state.ctypes.data, state.nbytes, H2D) == 0
        rc = od.lib().od_conservation_sums_noalloc(p.nx, p.ny, p.nz, p.lx, p.ly, p.lz, d_state.ptr, real,
                                                   d_out.ptr, None)
        od._check(rc, "od_conservation_sums_noalloc")
        assert hip.hipDeviceSynchronize() == 0
        out = np.empty(5)
        assert hip.hipMemcpy(out.ctypes.data, d_out.ptr, 40, D2H) == 0
    return out


def _with_context(p, state):
    with od.Solver(p) as s:
        s.set_state(state)
        return s.sums()


def test_noalloc_sums_equal_context_sums_on_consecutive_calls():
    hip = _hip()
    cases = [((12, 10, 9), od.OD_F64, 1), ((3, 2, 5), od.OD_F64, 2), ((4, 3, 2), od.OD_F32, 3)]
    for (nx, ny, nz), real, seed in cases:
        print(f"CASE noalloc sums {nx}x{ny}x{nz} u_real={real}", flush=True)
        p = od.params(nx, ny, nz, TWO_PI, 2.0, np.pi, u_real=real, flux="rusanov", viscous=False)
        dt = np.float32 if real == od.OD_F32 else np.float64
        u = smooth_state(nx, ny, nz, p.lx, p.ly, p.lz, seed=seed, dtype=dt)
        got = _noalloc(hip, p, u, real)
        ref = _with_context(p, u)
        assert np.all(np.isfinite(got))
        assert got.tobytes() == np.asarray(ref, np.float64).tobytes(), (got, ref)


def test_noalloc_sums_refuse_meshes_above_the_maximum():
    hip = _hip()
    with DeviceBuffer(hip, 40) as d_out:
        zero = np.zeros(5)
        assert hip.hipMemcpy(d_out.ptr, zero.ctypes.data, 40, H2D) == 0
        L = od.lib()
        rc = L.od_conservation_sums_noalloc(1024, 1024, 2, 1.0, 1.0, 1.0, d_out.ptr, od.OD_F32, d_out.ptr, None)
        assert rc == 1
        assert b"exceed the maximum" in L.od_last_error()
        out = np.ones(5)
        assert hip.hipMemcpy(out.ctypes.data, d_out.ptr, 40, D2H) == 0
        assert np.all(out == 0.0)  # nothing was launched
