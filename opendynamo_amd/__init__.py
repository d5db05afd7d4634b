"""opendynamo_amd — MI355X-native FR P=3 compressible Navier-Stokes RK3 step.

Python front end over the C ABI of libopendynamo_hip.so (include/opendynamo.h).
There is no CPU fallback: if the HIP library is not built, or no GPU is
present, the calls raise.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libopendynamo_hip.so")

OD_F32, OD_F64 = 4, 8
FLUX = {"slau2": 0, "ausm+": 1, "rusanov": 2}
_NP = {OD_F32: np.float32, OD_F64: np.float64}

# Every symbol include/opendynamo.h declares.
EXPORTS = [
    "od_default_params", "od_create", "od_destroy", "od_set_state", "od_get_state",
    "od_get_state_async", "od_wait", "od_step", "od_synchronize", "od_rhs",
    "od_conservation_sums", "od_cfl_dt", "od_time_steps", "od_time_kernels", "od_info",
    "od_last_error",
]


class Params(ctypes.Structure):
    """Mirror of od_params (include/opendynamo.h)."""
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("lz", ctypes.c_double),
        ("u_real", ctypes.c_int32), ("q_real", ctypes.c_int32),
        ("rhs_real", ctypes.c_int32), ("compute_real", ctypes.c_int32),
        ("flux_scheme", ctypes.c_int32), ("y_wall", ctypes.c_int32),
        ("viscous", ctypes.c_int32), ("viscosity_model", ctypes.c_int32),
        ("mu_ref", ctypes.c_double), ("t_ref", ctypes.c_double), ("s_suth", ctypes.c_double),
        ("body_force_x", ctypes.c_double),
        ("rk_accum", ctypes.c_int32), ("fast_math", ctypes.c_int32),
    ]


class ODError(RuntimeError):
    pass


_lib = None


def lib():
    """Load the HIP library; raise if it is not built (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ODError(f"{LIB_PATH} is not built: run `python -m opendynamo_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        vp, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)
        L.od_default_params.argtypes = [ctypes.POINTER(Params)]
        L.od_default_params.restype = None
        L.od_create.argtypes = [ctypes.POINTER(Params), ctypes.c_int, ctypes.POINTER(vp)]
        L.od_destroy.argtypes = [vp]
        L.od_destroy.restype = None
        L.od_set_state.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        L.od_get_state.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        L.od_get_state_async.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64)]
        L.od_wait.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(vp)]
        L.od_step.argtypes = [vp, ctypes.c_double, ctypes.c_uint32]
        L.od_synchronize.argtypes = [vp]
        L.od_rhs.argtypes = [vp, vp, ctypes.c_size_t]
        L.od_conservation_sums.argtypes = [vp, dp]
        L.od_cfl_dt.argtypes = [vp, ctypes.c_double, dp]
        L.od_time_steps.argtypes = [vp, ctypes.c_double, ctypes.c_uint32, dp]
        L.od_time_kernels.argtypes = [vp, ctypes.c_double, ctypes.c_uint32, dp,
                                      ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_int)]
        L.od_info.argtypes = [vp, dp]
        L.od_last_error.argtypes = []
        L.od_last_error.restype = ctypes.c_char_p
        for name in EXPORTS:
            if name not in ("od_default_params", "od_destroy", "od_last_error"):
                getattr(L, name).restype = ctypes.c_int
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        msg = lib().od_last_error().decode(errors="replace")
        raise ODError(f"{what} failed (status {rc}): {msg}")


def params(nx, ny, nz, lx, ly, lz, u_real=OD_F32, q_real=None, rhs_real=None, compute_real=None,
           flux="slau2", y_wall=False, viscous=True, viscosity_model=0, mu_ref=0.00016, t_ref=273.15,
           s_suth=110.4, body_force_x=0.001, rk_accum=0, fast_math=0):
    """Build od_params; lx/ly/lz are rounded to f32 like the config values."""
    p = Params()
    lib().od_default_params(ctypes.byref(p))
    p.nx, p.ny, p.nz = nx, ny, nz
    p.lx, p.ly, p.lz = (float(np.float32(x)) for x in (lx, ly, lz))
    p.u_real = u_real
    p.q_real = u_real if q_real is None else q_real
    p.rhs_real = u_real if rhs_real is None else rhs_real
    p.compute_real = u_real if compute_real is None else compute_real
    p.flux_scheme = FLUX[flux] if isinstance(flux, str) else int(flux)
    p.y_wall = int(bool(y_wall))
    p.viscous = int(bool(viscous))
    p.viscosity_model = int(viscosity_model)
    p.mu_ref, p.t_ref, p.s_suth = mu_ref, t_ref, s_suth
    p.body_force_x = body_force_x
    p.rk_accum = int(rk_accum)
    p.fast_math = int(fast_math)
    return p


class Solver:
    """One FR P=3 box on one GPU: owns the device state and the step graph."""

    def __init__(self, p, device=0):
        self.p = p
        self.n_values = 5 * p.nx * p.ny * p.nz * 64
        self._h = ctypes.c_void_p()
        _check(lib().od_create(ctypes.byref(p), int(device), ctypes.byref(self._h)), "od_create")

    def close(self):
        if self._h:
            lib().od_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def dtype(self):
        return _NP[self.p.u_real]

    def set_state(self, state):
        a = np.ascontiguousarray(state)
        if a.dtype not in (np.float32, np.float64):
            a = a.astype(self.dtype)
        code = OD_F32 if a.dtype == np.float32 else OD_F64
        _check(lib().od_set_state(self._h, a.ctypes.data, a.size, code), "od_set_state")

    def get_state(self, dtype=None):
        dt = self.dtype if dtype is None else np.dtype(dtype).type
        out = np.empty(self.n_values, dtype=dt)
        code = OD_F32 if dt == np.float32 else OD_F64
        _check(lib().od_get_state(self._h, out.ctypes.data, out.size, code), "od_get_state")
        return out

    def step(self, dt, n_steps=1):
        """Enqueue n RK3 steps (dt rounded to f32 like the config value)."""
        _check(lib().od_step(self._h, float(np.float32(dt)), int(n_steps)), "od_step")

    def synchronize(self):
        _check(lib().od_synchronize(self._h), "od_synchronize")

    def rhs(self):
        out = np.empty(self.n_values, dtype=_NP[self.p.compute_real])
        _check(lib().od_rhs(self._h, out.ctypes.data, out.size), "od_rhs")
        return out

    def sums(self):
        out = np.zeros(5)
        _check(lib().od_conservation_sums(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
               "od_conservation_sums")
        return out

    def cfl_dt(self, cfl):
        v = ctypes.c_double()
        _check(lib().od_cfl_dt(self._h, float(cfl), ctypes.byref(v)), "od_cfl_dt")
        return v.value

    def time_steps(self, dt, n_steps):
        v = ctypes.c_double()
        _check(lib().od_time_steps(self._h, float(np.float32(dt)), int(n_steps), ctypes.byref(v)), "od_time_steps")
        return v.value

    def time_kernels(self, dt, n_steps):
        ms = (ctypes.c_double * 4)()
        names = (ctypes.c_char_p * 4)()
        n = ctypes.c_int()
        _check(lib().od_time_kernels(self._h, float(np.float32(dt)), int(n_steps), ms, names, ctypes.byref(n)),
               "od_time_kernels")
        return {names[q].decode(): ms[q] for q in range(n.value)}

    def get_state_async(self):
        t = ctypes.c_uint64()
        _check(lib().od_get_state_async(self._h, ctypes.byref(t)), "od_get_state_async")
        return t.value

    def wait(self, ticket):
        ptr = ctypes.c_void_p()
        _check(lib().od_wait(self._h, ctypes.c_uint64(ticket), ctypes.byref(ptr)), "od_wait")
        buf = (ctypes.c_char * (self.n_values * self.p.u_real)).from_address(ptr.value)
        return np.frombuffer(buf, dtype=self.dtype).copy()

    def info(self):
        out = np.zeros(4)
        _check(lib().od_info(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "od_info")
        return {"dof": int(out[0]), "device_bytes": int(out[1]), "launches_per_step": int(out[2])}
