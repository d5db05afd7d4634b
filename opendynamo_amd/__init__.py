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


def _state_dtype(default, dtype):
    """float32 or float64 (the only types the C side writes); anything else
    is refused instead of letting 8-byte values overrun a smaller buffer."""
    dt = default if dtype is None else np.dtype(dtype).type
    if dt not in (np.float32, np.float64):
        raise ValueError(f"state dtype must be float32 or float64, not {np.dtype(dt).name}")
    return dt

# Every symbol include/opendynamo.h declares.
EXPORTS = [
    "od_default_params", "od_create", "od_destroy", "od_set_state", "od_get_state",
    "od_get_state_async", "od_wait", "od_release", "od_step", "od_synchronize", "od_rhs",
    "od_conservation_sums", "od_conservation_sums_noalloc", "od_cfl_dt", "od_time_steps", "od_time_kernels",
    "od_info", "od_last_error",
    "od_batch_create", "od_batch_destroy", "od_batch_set_state", "od_batch_get_state",
    "od_batch_step", "od_batch_synchronize", "od_batch_cfl_dt", "od_batch_conservation_sums",
    "od_batch_set_schedule", "od_batch_get_schedule", "od_batch_autotune", "od_batch_time_steps",
    "od_batch_info",
    "od_mlp_create", "od_mlp_destroy", "od_mlp_set_params", "od_mlp_get_params", "od_mlp_set_scaling",
    "od_mlp_set_output_scale",
    "od_mlp_forward", "od_mlp_loss_grad", "od_mlp_train", "od_closure_features", "od_closure_attach",
    "od_set_eddy_viscosity", "od_get_eddy_viscosity", "od_closure_stats",
]

MLP_FEATURES, MLP_HIDDEN, FEATURE_STRIDE, MLP_PARAMS = 12, 64, 16, 9217


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


class Box(ctypes.Structure):
    """Mirror of od_box."""
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("lz", ctypes.c_double),
    ]


class Schedule(ctypes.Structure):
    """Mirror of od_schedule."""
    _fields_ = [
        ("orient", ctypes.c_int32), ("pairs", ctypes.c_int32),
        ("order", ctypes.c_int32), ("groups_per_cu", ctypes.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ClosureOpts(ctypes.Structure):
    """Mirror of od_closure_opts."""
    _fields_ = [("update_interval", ctypes.c_int32), ("budget_kappa", ctypes.c_double)]


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
        L.od_release.argtypes = [vp, ctypes.c_uint64]
        L.od_step.argtypes = [vp, ctypes.c_double, ctypes.c_uint32]
        L.od_synchronize.argtypes = [vp]
        L.od_rhs.argtypes = [vp, vp, ctypes.c_size_t]
        L.od_conservation_sums.argtypes = [vp, dp]
        L.od_conservation_sums_noalloc.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                                   ctypes.c_double, ctypes.c_double, vp, ctypes.c_int, vp, vp]
        L.od_cfl_dt.argtypes = [vp, ctypes.c_double, dp]
        L.od_time_steps.argtypes = [vp, ctypes.c_double, ctypes.c_uint32, dp]
        L.od_time_kernels.argtypes = [vp, ctypes.c_double, ctypes.c_uint32, dp,
                                      ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_int)]
        L.od_info.argtypes = [vp, dp]
        L.od_last_error.argtypes = []
        L.od_last_error.restype = ctypes.c_char_p
        sp = ctypes.POINTER(Schedule)
        L.od_batch_create.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(Box), ctypes.c_int32, ctypes.c_int,
                                      ctypes.POINTER(vp)]
        L.od_batch_destroy.argtypes = [vp]
        L.od_batch_destroy.restype = None
        L.od_batch_set_state.argtypes = [vp, ctypes.c_int32, vp, ctypes.c_size_t, ctypes.c_int]
        L.od_batch_get_state.argtypes = [vp, ctypes.c_int32, vp, ctypes.c_size_t, ctypes.c_int]
        L.od_batch_step.argtypes = [vp, dp, ctypes.c_uint32]
        L.od_batch_synchronize.argtypes = [vp]
        L.od_batch_cfl_dt.argtypes = [vp, ctypes.c_double, dp]
        L.od_batch_conservation_sums.argtypes = [vp, dp]
        L.od_batch_set_schedule.argtypes = [vp, sp]
        L.od_batch_get_schedule.argtypes = [vp, sp]
        L.od_batch_autotune.argtypes = [vp, dp, ctypes.c_uint32, sp, dp, sp, dp, ctypes.c_int32,
                                        ctypes.POINTER(ctypes.c_int32)]
        L.od_batch_time_steps.argtypes = [vp, dp, ctypes.c_uint32, ctypes.c_int32, dp]
        L.od_batch_info.argtypes = [vp, dp]
        fp = ctypes.POINTER(ctypes.c_float)
        L.od_mlp_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        L.od_mlp_destroy.argtypes = [vp]
        L.od_mlp_destroy.restype = None
        L.od_mlp_set_params.argtypes = [vp, fp, ctypes.c_size_t]
        L.od_mlp_get_params.argtypes = [vp, fp, ctypes.c_size_t]
        L.od_mlp_set_scaling.argtypes = [vp, fp, fp, fp, fp]
        L.od_mlp_set_output_scale.argtypes = [vp, ctypes.c_float]
        L.od_mlp_forward.argtypes = [vp, fp, ctypes.c_size_t, fp]
        L.od_mlp_loss_grad.argtypes = [vp, fp, fp, ctypes.c_size_t, dp, fp]
        L.od_mlp_train.argtypes = [vp, fp, fp, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                                   ctypes.c_uint64, fp]
        L.od_closure_features.argtypes = [vp, fp, ctypes.c_size_t]
        L.od_closure_attach.argtypes = [vp, vp, ctypes.POINTER(ClosureOpts)]
        L.od_set_eddy_viscosity.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        L.od_get_eddy_viscosity.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        L.od_closure_stats.argtypes = [vp, dp]
        for name in EXPORTS:
            if name not in ("od_default_params", "od_destroy", "od_last_error", "od_batch_destroy",
                            "od_mlp_destroy"):
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
        dt = _state_dtype(self.dtype, dtype)
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

    def wait(self, ticket, release=True):
        """Copy of the snapshot of `ticket`; the pinned slot goes back to the
        pool afterwards unless release=False (then call release(ticket))."""
        ptr = ctypes.c_void_p()
        _check(lib().od_wait(self._h, ctypes.c_uint64(ticket), ctypes.byref(ptr)), "od_wait")
        buf = (ctypes.c_char * (self.n_values * self.p.u_real)).from_address(ptr.value)
        out = np.frombuffer(buf, dtype=self.dtype).copy()
        if release:
            self.release(ticket)
        return out

    def release(self, ticket):
        _check(lib().od_release(self._h, ctypes.c_uint64(ticket)), "od_release")

    def info(self):
        out = np.zeros(4)
        _check(lib().od_info(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "od_info")
        return {"dof": int(out[0]), "device_bytes": int(out[1]), "launches_per_step": int(out[2])}

    # ---- learned closure ----
    @property
    def n_elements(self):
        return self.p.nx * self.p.ny * self.p.nz

    def features(self):
        """[E][16]: 12 features, rho_e, Delta^2 |S|_e, kinetic energy, integral of 2 S^d:S^d."""
        out = np.empty((self.n_elements, FEATURE_STRIDE), dtype=np.float32)
        _check(lib().od_closure_features(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), out.size),
               "od_closure_features")
        return out

    def set_eddy_viscosity(self, mu_t):
        """Per-element eddy viscosity [E] (None switches it off)."""
        if mu_t is None:
            _check(lib().od_set_eddy_viscosity(self._h, None, 0, OD_F64), "od_set_eddy_viscosity")
            return
        a = np.ascontiguousarray(mu_t, dtype=np.float64)
        _check(lib().od_set_eddy_viscosity(self._h, a.ctypes.data, a.size, OD_F64), "od_set_eddy_viscosity")

    def get_eddy_viscosity(self):
        out = np.empty(self.n_elements)
        _check(lib().od_get_eddy_viscosity(self._h, out.ctypes.data, out.size, OD_F64), "od_get_eddy_viscosity")
        return out

    def attach_closure(self, mlp, update_interval=1, budget_kappa=1.0):
        """Attach an MLP (None detaches); od_step refreshes mu_t every update_interval steps."""
        if mlp is None:
            _check(lib().od_closure_attach(self._h, None, None), "od_closure_attach")
            self._mlp = None
            return
        o = ClosureOpts(int(update_interval), float(budget_kappa))
        _check(lib().od_closure_attach(self._h, mlp._h, ctypes.byref(o)), "od_closure_attach")
        self._mlp = mlp  # keep the network alive while attached

    def closure_stats(self):
        out = np.zeros(6)
        _check(lib().od_closure_stats(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
               "od_closure_stats")
        return {"updates": int(out[0]), "budget_fallbacks": int(out[1]), "out_of_range": int(out[2]),
                "kinetic_energy": out[3], "model_dissipation": out[4], "active": bool(out[5])}


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Batch:
    """An ensemble of independent boxes advanced together (od_batch_*).

    `boxes` is a list of (nx, ny, nz, lx, ly, lz); `p` (from params()) gives
    the shared physics, scheme and precision; its mesh fields are ignored.
    """

    def __init__(self, p, boxes, device=0):
        self.p = p
        self.boxes = [tuple(b) for b in boxes]
        arr = (Box * len(self.boxes))()
        for q, (nx, ny, nz, lx, ly, lz) in enumerate(self.boxes):
            arr[q] = Box(nx, ny, nz, *(float(np.float32(x)) for x in (lx, ly, lz)))
        self._h = ctypes.c_void_p()
        _check(lib().od_batch_create(ctypes.byref(p), arr, len(self.boxes), int(device), ctypes.byref(self._h)),
               "od_batch_create")

    def close(self):
        if self._h:
            lib().od_batch_destroy(self._h)
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

    def n_values(self, box):
        nx, ny, nz = self.boxes[box][:3]
        return 5 * 64 * nx * ny * nz

    def _dt(self, dt):
        a = np.array([float(np.float32(x)) for x in np.broadcast_to(dt, (len(self.boxes),))], dtype=np.float64)
        return a

    def set_state(self, box, state):
        a = np.ascontiguousarray(state)
        if a.dtype not in (np.float32, np.float64):
            a = a.astype(self.dtype)
        code = OD_F32 if a.dtype == np.float32 else OD_F64
        _check(lib().od_batch_set_state(self._h, int(box), a.ctypes.data, a.size, code), "od_batch_set_state")

    def get_state(self, box, dtype=None):
        dt = _state_dtype(self.dtype, dtype)
        out = np.empty(self.n_values(box), dtype=dt)
        code = OD_F32 if dt == np.float32 else OD_F64
        _check(lib().od_batch_get_state(self._h, int(box), out.ctypes.data, out.size, code), "od_batch_get_state")
        return out

    def step(self, dt, n_steps=1):
        """n RK3 steps of every box; dt is a scalar or one value per box (rounded to f32)."""
        d = self._dt(dt)
        _check(lib().od_batch_step(self._h, _dp(d), int(n_steps)), "od_batch_step")

    def synchronize(self):
        _check(lib().od_batch_synchronize(self._h), "od_batch_synchronize")

    def cfl_dt(self, cfl):
        out = np.zeros(len(self.boxes))
        _check(lib().od_batch_cfl_dt(self._h, float(cfl), _dp(out)), "od_batch_cfl_dt")
        return out

    def sums(self):
        out = np.zeros((len(self.boxes), 5))
        _check(lib().od_batch_conservation_sums(self._h, _dp(out)), "od_batch_conservation_sums")
        return out

    def set_schedule(self, orient=0, pairs=1, order=0, groups_per_cu=0):
        s = Schedule(orient, pairs, order, groups_per_cu)
        _check(lib().od_batch_set_schedule(self._h, ctypes.byref(s)), "od_batch_set_schedule")

    def get_schedule(self):
        s = Schedule()
        _check(lib().od_batch_get_schedule(self._h, ctypes.byref(s)), "od_batch_get_schedule")
        return s.as_dict()

    def autotune(self, dt, steps_per_trial=3, cap=256):
        """Time every candidate schedule, keep the fastest; the state is unchanged."""
        d = self._dt(dt)
        best = Schedule()
        best_ms = ctypes.c_double()
        trials = (Schedule * cap)()
        ms = np.zeros(cap)
        n = ctypes.c_int32()
        _check(lib().od_batch_autotune(self._h, _dp(d), int(steps_per_trial), ctypes.byref(best), ctypes.byref(best_ms),
                                       trials, _dp(ms), cap, ctypes.byref(n)), "od_batch_autotune")
        table = [dict(trials[q].as_dict(), ms_per_step=float(ms[q])) for q in range(min(n.value, cap))]
        return best.as_dict(), best_ms.value, table

    def time_steps(self, dt, n_steps, per_box=False):
        d = self._dt(dt)
        v = ctypes.c_double()
        _check(lib().od_batch_time_steps(self._h, _dp(d), int(n_steps), int(bool(per_box)), ctypes.byref(v)),
               "od_batch_time_steps")
        return v.value

    def info(self):
        out = np.zeros(4)
        _check(lib().od_batch_info(self._h, _dp(out)), "od_batch_info")
        return {"dof": int(out[0]), "device_bytes": int(out[1]), "launches_per_step": int(out[2]),
                "boxes": int(out[3])}


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class MLP:
    """The closure network on the GPU (od_mlp_*): 12 -> 64 -> 64 -> 64 -> 1,
    GELU hidden layers, softplus output, BF16 weights on MFMA."""

    def __init__(self, device=0, params=None):
        self._h = ctypes.c_void_p()
        _check(lib().od_mlp_create(int(device), ctypes.byref(self._h)), "od_mlp_create")
        if params is not None:
            self.set_params(params)

    def close(self):
        if self._h:
            lib().od_mlp_destroy(self._h)
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

    def set_params(self, params):
        a = np.ascontiguousarray(params, dtype=np.float32)
        _check(lib().od_mlp_set_params(self._h, _fp(a), a.size), "od_mlp_set_params")

    def get_params(self):
        out = np.empty(MLP_PARAMS, dtype=np.float32)
        _check(lib().od_mlp_get_params(self._h, _fp(out), out.size), "od_mlp_get_params")
        return out

    def set_scaling(self, shift, scale, lo=None, hi=None):
        f = lambda v, d: np.ascontiguousarray(np.full(MLP_FEATURES, d) if v is None else v, dtype=np.float32)
        sh, sc, l, h = f(shift, 0.0), f(scale, 1.0), f(lo, -np.inf), f(hi, np.inf)
        self._scaling = (sh, sc, l, h)
        _check(lib().od_mlp_set_scaling(self._h, _fp(sh), _fp(sc), _fp(l), _fp(h)), "od_mlp_set_scaling")

    def set_output_scale(self, s):
        """C = s * softplus(z)."""
        _check(lib().od_mlp_set_output_scale(self._h, float(s)), "od_mlp_set_output_scale")

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, MLP_FEATURES)
        y = np.empty(x.shape[0], dtype=np.float32)
        _check(lib().od_mlp_forward(self._h, _fp(x), x.shape[0], _fp(y)), "od_mlp_forward")
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, MLP_FEATURES)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.empty(MLP_PARAMS, dtype=np.float32)
        loss = ctypes.c_double()
        _check(lib().od_mlp_loss_grad(self._h, _fp(x), _fp(t), x.shape[0], ctypes.byref(loss), _fp(g)),
               "od_mlp_loss_grad")
        return loss.value, g

    def train(self, x, t, epochs, batch=65536, lr=1e-3, seed=0):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, MLP_FEATURES)
        t = np.ascontiguousarray(t, dtype=np.float32)
        hist = np.zeros(epochs, dtype=np.float32)
        _check(lib().od_mlp_train(self._h, _fp(x), _fp(t), x.shape[0], int(epochs), int(batch), float(lr),
                                  int(seed), _fp(hist)), "od_mlp_train")
        return hist
