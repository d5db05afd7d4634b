"""ctypes front end of libodhost.so (include/odhost.h): config.toml and the
solver-native HDF5 state files."""
import ctypes
import json
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
HOSTLIB_PATH = os.path.join(PKG, "libodhost.so")
EXPORTS = ["odh_save_state", "odh_load_state", "odh_read_header", "odh_config_json", "odh_last_error"]

_lib = None


class HostError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(HOSTLIB_PATH):
            raise HostError(f"{HOSTLIB_PATH} is not built: run `python -m opendynamo_amd.build`")
        L = ctypes.CDLL(HOSTLIB_PATH)
        u64, dbl, cp, vp, sz = ctypes.c_uint64, ctypes.c_double, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_size_t
        L.odh_save_state.argtypes = [u64, dbl, cp, cp, cp, vp, u64, u64, u64, dbl, dbl, dbl, cp, ctypes.c_char_p, sz]
        L.odh_load_state.argtypes = [cp, u64, u64, u64, cp, vp, sz, ctypes.POINTER(u64), ctypes.POINTER(dbl)]
        L.odh_read_header.argtypes = [cp, ctypes.POINTER(u64), ctypes.POINTER(dbl), ctypes.c_char_p]
        L.odh_config_json.argtypes = [cp, ctypes.c_char_p, sz]
        L.odh_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        raise HostError(f"{what}: {lib().odh_last_error().decode(errors='replace')}")


def save_state(label, t, state, nx, ny, nz, lx, ly, lz, out_dir, prefix, compression="none"):
    """Write <out_dir>/<prefix>_<f32|f64>_<label:06>.h5; returns the path.
    lx/ly/lz are stored as the f32 config values widened (solver.rs:247-249)."""
    a = np.ascontiguousarray(state)
    kind = {np.dtype(np.float32): b"f32", np.dtype(np.float64): b"f64"}[a.dtype]
    buf = ctypes.create_string_buffer(4096)
    L = [float(np.float32(x)) for x in (lx, ly, lz)]
    _check(lib().odh_save_state(int(label), float(t), kind, out_dir.encode(), prefix.encode(), a.ctypes.data,
                                nx, ny, nz, *L, compression.encode(), buf, 4096), "odh_save_state")
    return buf.value.decode()


def read_header(path):
    u = (ctypes.c_uint64 * 5)()
    d = (ctypes.c_double * 4)()
    k = ctypes.create_string_buffer(8)
    _check(lib().odh_read_header(path.encode(), u, d, k), "odh_read_header")
    return {"nx": u[0], "ny": u[1], "nz": u[2], "fr_order": u[3], "step": u[4],
            "t": d[0], "lx": d[1], "ly": d[2], "lz": d[3], "u_real": k.value.decode()}


def load_state(path, nx, ny, nz, kind):
    """Validating loader (io.rs:163-234) in the file's native type."""
    dt = np.float32 if kind == "f32" else np.float64
    out = np.empty(5 * nx * ny * nz * 64, dtype=dt)
    step, t = ctypes.c_uint64(), ctypes.c_double()
    _check(lib().odh_load_state(path.encode(), nx, ny, nz, kind.encode(), out.ctypes.data, out.nbytes,
                                ctypes.byref(step), ctypes.byref(t)), "odh_load_state")
    return out, step.value, t.value


def config(path):
    """config.toml parsed by the C++ reader (reference keys and defaults)."""
    buf = ctypes.create_string_buffer(1 << 16)
    _check(lib().odh_config_json(path.encode(), buf, 1 << 16), "odh_config_json")
    return json.loads(buf.value.decode())
