"""ctypes front end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, never by the product package.  The C++ side
restates drunkseal/opendynamo's kernels op for op (see oracle.h).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("lz", ctypes.c_double),
        ("u_real", ctypes.c_int32), ("q_real", ctypes.c_int32),
        ("rhs_real", ctypes.c_int32), ("compute_real", ctypes.c_int32),
        ("flux_scheme", ctypes.c_int32), ("y_wall", ctypes.c_int32),
        ("viscous", ctypes.c_int32), ("viscosity_model", ctypes.c_int32),
        ("mu_ref", ctypes.c_double), ("t_ref", ctypes.c_double), ("s_suth", ctypes.c_double),
        ("body_force_x", ctypes.c_double),
        ("rk_accum", ctypes.c_int32), ("n_threads", ctypes.c_int32),
    ]


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_fr_operators.argtypes = [dp] * 7
        L.orc_fr_operators.restype = None
        L.orc_step.argtypes = [ctypes.POINTER(OrcParams), ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
        L.orc_step_mut.argtypes = [ctypes.POINTER(OrcParams), ctypes.c_void_p, dp, ctypes.c_double, ctypes.c_int]
        L.orc_rhs_total.argtypes = [ctypes.POINTER(OrcParams), ctypes.c_void_p, ctypes.c_void_p]
        L.orc_debug_array.argtypes = [ctypes.POINTER(OrcParams), ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        L.orc_point.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
        L.orc_sums.argtypes = [ctypes.POINTER(OrcParams), ctypes.c_void_p, dp]
        L.orc_cfl_dt.argtypes = [ctypes.POINTER(OrcParams), ctypes.c_void_p, ctypes.c_double]
        L.orc_cfl_dt.restype = ctypes.c_double
        _lib = L
    return _lib


_DT = {4: np.float32, 8: np.float64}
_SCHEME = {"slau2": 0, "ausm+": 1, "rusanov": 2}


def params(nx, ny, nz, lx, ly, lz, u_real=8, q_real=None, rhs_real=None, compute_real=None,
           flux="rusanov", y_wall=False, viscous=True, viscosity_model=0, mu_ref=0.00016,
           t_ref=273.15, s_suth=110.4, body_force_x=0.001, rk_accum=0, n_threads=0):
    """Build OrcParams.  lx/ly/lz are rounded to f32 like the config values."""
    q_real = u_real if q_real is None else q_real
    rhs_real = u_real if rhs_real is None else rhs_real
    compute_real = u_real if compute_real is None else compute_real
    p = OrcParams()
    p.nx, p.ny, p.nz = nx, ny, nz
    p.lx, p.ly, p.lz = (float(np.float32(x)) for x in (lx, ly, lz))
    p.u_real, p.q_real, p.rhs_real, p.compute_real = u_real, q_real, rhs_real, compute_real
    p.flux_scheme = _SCHEME[flux] if isinstance(flux, str) else int(flux)
    p.y_wall = int(bool(y_wall))
    p.viscous = int(bool(viscous))
    p.viscosity_model = int(viscosity_model)
    p.mu_ref, p.t_ref, p.s_suth = mu_ref, t_ref, s_suth
    p.body_force_x = body_force_x
    p.rk_accum = int(rk_accum)
    p.n_threads = int(n_threads)
    return p


def n_points(p):
    return p.nx * p.ny * p.nz * 64


def fr_operators():
    out = [np.zeros(n) for n in (4, 16, 4, 4, 4, 4, 4)]
    dp = ctypes.POINTER(ctypes.c_double)
    lib().orc_fr_operators(*[a.ctypes.data_as(dp) for a in out])
    names = ["points", "d", "extrap_left", "extrap_right", "g_left_prime", "g_right_prime", "weights"]
    return dict(zip(names, out))


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with code {rc}")


def step(p, state, dt, n_steps):
    """Return a copy of `state` advanced n_steps RK3 steps (dt rounded to f32)."""
    u = np.ascontiguousarray(state, dtype=_DT[p.u_real]).copy()
    assert u.size == 5 * n_points(p)
    _check(lib().orc_step(ctypes.byref(p), u.ctypes.data, float(np.float32(dt)), int(n_steps)), "step")
    return u


def step_mut(p, state, mu_t, dt, n_steps):
    """step() with a per-element eddy viscosity mu_t[E] added to the molecular one."""
    u = np.ascontiguousarray(state, dtype=_DT[p.u_real]).copy()
    m = np.ascontiguousarray(mu_t, dtype=np.float64)
    assert u.size == 5 * n_points(p) and m.size == p.nx * p.ny * p.nz
    _check(lib().orc_step_mut(ctypes.byref(p), u.ctypes.data, m.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                              float(np.float32(dt)), int(n_steps)), "step_mut")
    return u


def rhs_total(p, state):
    u = np.ascontiguousarray(state, dtype=_DT[p.u_real])
    out = np.zeros(5 * n_points(p), dtype=_DT[p.compute_real])
    _check(lib().orc_rhs_total(ctypes.byref(p), u.ctypes.data, out.ctypes.data), "rhs_total")
    return out


def debug_array(p, state, which):
    u = np.ascontiguousarray(state, dtype=_DT[p.u_real])
    E = p.nx * p.ny * p.nz
    if which == 0:
        out = np.zeros(5 * E * 64, dtype=_DT[p.rhs_real])
    elif which <= 6:
        out = np.zeros(E * 16 * 12, dtype=_DT[p.compute_real])
    else:
        out = np.zeros(5 * E * 16, dtype=_DT[p.compute_real])
    _check(lib().orc_debug_array(ctypes.byref(p), u.ctypes.data, int(which), out.ctypes.data), "debug")
    return out


def point(real, kind, values):
    inp = np.zeros(32)
    inp[: len(values)] = values
    out = np.zeros(5)
    dp = ctypes.POINTER(ctypes.c_double)
    _check(lib().orc_point(int(real), int(kind), inp.ctypes.data_as(dp), out.ctypes.data_as(dp)), "point")
    return out


def sums(p, state):
    u = np.ascontiguousarray(state, dtype=_DT[p.u_real])
    out = np.zeros(5)
    _check(lib().orc_sums(ctypes.byref(p), u.ctypes.data, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "sums")
    return out


def cfl_dt(p, state, cfl):
    u = np.ascontiguousarray(state, dtype=_DT[p.u_real])
    return lib().orc_cfl_dt(ctypes.byref(p), u.ctypes.data, float(cfl))
