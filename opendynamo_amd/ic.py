"""Initial-condition generator (counterpart of scripts/make_initial_state.py).

Builds the var-major state planes (rho, rhou, rhov, rhow, E) on the GL
solution points of the FR P=3 layout: elem-major 64-point blocks,
pt = i + 4j + 16k, elem = ex + nx*ey + nx*ny*ez (make_initial_state.py:16-22).

  channel_state   laminar channel / uniform base flow + the wall-damped
                  divergence-free trip (make_initial_state.py:122-241),
                  evaluated in float32
  fourier_state   periodic divergence-free random-Fourier perturbation of a
                  uniform flow (no reference counterpart; configs with
                  periodic boxes)
  ensemble_boxes  sizes and Fourier parameters of an ensemble of periodic
                  boxes (no reference counterpart)
"""
import math

import numpy as np

F32 = np.float32
GAMMA = 1.4


def gl_points():
    """P=3 Gauss-Legendre points by Newton on P_4 (fr_math.rs:31-53)."""
    n = 4
    roots = []
    for i in range(n):
        x = math.cos(math.pi * (i + 0.75) / (n + 0.5))
        while True:
            p0, p1 = 1.0, x
            for kk in range(2, n + 1):
                p0, p1 = p1, ((2.0 * kk - 1.0) * x * p1 - (kk - 1.0) * p0) / kk
            dp = n / (x * x - 1.0) * (x * p1 - p0)
            s = p1 / dp
            x -= s
            if abs(s) < 1e-14:
                break
        roots.append(x)
    return np.array(sorted(roots))


def point_coords(nx, ny, nz, lx, ly, lz, dtype=np.float64, y_center=False):
    """Physical coordinates of every solution point, flattened in plane order.
    x, z in [0, L); y in [0, ly) or, with y_center, in [-ly/2, ly/2]."""
    xi = gl_points().astype(dtype)
    e = np.arange(nx * ny * nz)
    ex, ey, ez = e % nx, (e // nx) % ny, e // (nx * ny)
    pt = np.arange(64)
    i, j, k = pt % 4, (pt // 4) % 4, pt // 16
    half = dtype(0.5)
    X = (ex[:, None].astype(dtype) + half + half * xi[i][None, :]) * dtype(lx / nx)
    Y = (ey[:, None].astype(dtype) + half + half * xi[j][None, :]) * dtype(ly / ny)
    Z = (ez[:, None].astype(dtype) + half + half * xi[k][None, :]) * dtype(lz / nz)
    if y_center:
        Y = Y - dtype(0.5 * ly)
    return X.ravel().astype(dtype), Y.ravel().astype(dtype), Z.ravel().astype(dtype)


def fourier_modes(n_modes, seed, lx, ly, lz, kmax=2):
    """Wave vectors k, unit amplitude vectors a (a . k = 0 exactly in f64
    up to rounding) and phases of the divergence-free perturbation."""
    rng = np.random.default_rng(seed)
    ks, amps, phases = [], [], []
    while len(ks) < n_modes:
        n = rng.integers(-kmax, kmax + 1, size=3)
        if not n.any():
            continue
        k = 2.0 * np.pi * n / np.array([lx, ly, lz])
        r = rng.standard_normal(3)
        a = np.cross(r, k)
        norm = np.linalg.norm(a)
        if norm < 1e-8 * np.linalg.norm(k):
            continue
        ks.append(k)
        amps.append(a / norm)
        phases.append(rng.uniform(0.0, 2.0 * np.pi))
    return np.array(ks), np.array(amps), np.array(phases)


def fourier_state(nx, ny, nz, lx, ly, lz, u0=(0.1, 0.0, 0.0), amplitude=1e-2, n_modes=8, seed=0,
                  rho=1.0, p=1.0 / GAMMA, kmax=2, dtype=np.float64):
    """Uniform flow u0 plus sum_m A a_m cos(k_m . x + phi_m), A = amplitude *
    sqrt(2 / n_modes) so that the RMS of |u'| is `amplitude`.  Each mode is
    divergence free (a_m . k_m = 0) and periodic on the box (integer wave
    numbers).  rho and p are uniform; E = p/(gamma-1) + rho |u|^2 / 2."""
    X, Y, Z = point_coords(nx, ny, nz, lx, ly, lz)
    ks, amps, phases = fourier_modes(n_modes, seed, lx, ly, lz, kmax)
    A = amplitude * math.sqrt(2.0 / n_modes)
    vel = [np.full_like(X, u0[0]), np.full_like(X, u0[1]), np.full_like(X, u0[2])]
    for k, a, ph in zip(ks, amps, phases):
        c = np.cos(k[0] * X + k[1] * Y + k[2] * Z + ph)
        for d in range(3):
            vel[d] = vel[d] + A * a[d] * c
    r = np.full_like(X, rho)
    E = p / (GAMMA - 1.0) + 0.5 * r * (vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2)
    return np.concatenate([r, r * vel[0], r * vel[1], r * vel[2], E]).astype(dtype)


def ensemble_boxes(n_boxes=24, seed=7, n_lo=16, n_hi=40, modes=(12, 32), amps=(0.02, 0.1),
                   length=2.0 * math.pi):
    """An ensemble of independent periodic cubes [0, length)^3: per box an
    element count n ~ U{n_lo..n_hi} per side, a Fourier IC with n_modes ~
    U{modes}, amplitude ~ U[amps) and its own seed, all drawn from one
    generator seeded with `seed`.  Returns a list of dicts."""
    rng = np.random.default_rng(seed)
    n = rng.integers(n_lo, n_hi + 1, size=n_boxes)
    m = rng.integers(modes[0], modes[1] + 1, size=n_boxes)
    a = rng.uniform(amps[0], amps[1], size=n_boxes)
    s = rng.integers(0, 2**31 - 1, size=n_boxes)
    return [dict(nx=int(n[q]), ny=int(n[q]), nz=int(n[q]), lx=length, ly=length, lz=length, n_modes=int(m[q]),
                 amplitude=float(a[q]), seed=int(s[q])) for q in range(n_boxes)]


def ensemble_state(box, dtype=np.float32):
    """Initial state of one ensemble box: Mach 0.1 mean flow in x plus its
    divergence-free Fourier perturbation."""
    return fourier_state(box["nx"], box["ny"], box["nz"], box["lx"], box["ly"], box["lz"], u0=(0.1, 0.0, 0.0),
                         amplitude=box["amplitude"], n_modes=box["n_modes"], seed=box["seed"], dtype=dtype)


def _fma32(a, b, c):
    """f32 a*b + c rounded once through f64 (the product of two f32 is exact
    in f64); matches a hardware f32 FMA except in rare double-rounding cases."""
    return (np.asarray(a, np.float64) * np.asarray(b, np.float64) + np.asarray(c, np.float64)).astype(F32)


def _centered_coord(length, elems, elem_idx, xi_at_pt):
    """Coordinate in [-L/2, L/2]: fma(dx, e + 0.5 + 0.5 xi, -0.5 L) in f32
    (make_initial_state.py:135-145)."""
    mid = (elem_idx[:, None].astype(F32) + F32(0.5)) + F32(0.5) * xi_at_pt[None, :]
    d = F32(length) / F32(elems)
    return _fma32(d, mid, F32(F32(-0.5) * F32(length))).ravel()


def _hash01(s):
    """frac(sin(s) * 43758.5453) in f32 (make_initial_state.py:176-179)."""
    x = (np.sin(F32(s), dtype=F32) * F32(43758.5453)).astype(F32)
    return F32(x - np.floor(x))


def trip_modes(n_modes, seed, lx, lz):
    """Per-mode parameters of the wall-damped trip (make_initial_state.py:200-219):
    streamwise/spanwise integer wavenumbers, phase, and the three amplitudes."""
    seed = F32(float(seed))
    two_pi = F32(6.283185307179586)
    modes = []
    for m in range(n_modes):
        mix = F32(F32(m) * F32(0.113))
        r = [
            _hash01(F32(F32(mix * F32(1.113)) + seed) - F32(3.7183)),
            _hash01(F32(F32(mix * F32(2.229)) + seed) + F32(1.7429)),
            _hash01(F32(F32(mix * F32(3.717)) + seed) + F32(4.8527)),
            _hash01(F32(F32(mix * F32(5.381)) + seed) - F32(2.9956)),
            _hash01(F32(F32(mix * F32(7.993)) + seed) + F32(6.1141)),
            _hash01(F32(F32(mix * F32(9.677)) + seed) + F32(8.2437)),
        ]
        mx = 1 + int(F32(r[0] * F32(24.0)))
        mz = 1 + int(F32(r[1] * F32(12.0)))
        modes.append(dict(
            mx=mx, mz=mz,
            kx=F32(F32(two_pi * F32(mx)) / F32(lx)),
            kz=F32(F32(two_pi * F32(mz)) / F32(lz)),
            phi=F32(two_pi * r[2]),
            av=F32(F32(2.0) * r[3] - F32(1.0)),
            ac=F32(F32(2.0) * r[4] - F32(1.0)),
            wj=F32(F32(0.5) + r[5]),
        ))
    return modes


def channel_state(nx, ny, nz, lx, ly, lz, mach_bulk, initial_condition="channel",
                  perturbation_amplitude=0.0, perturbation_modes=16, perturbation_seed=7, dtype=np.float32):
    """Channel (u = 1.5 M (1 - eta^2)) or uniform (u = M) base flow with rho = 1,
    p = 1/gamma, plus the divergence-free wall-damped trip, all in float32
    (make_initial_state.py:122-241).  The trip superposes, per mode,
      u' = -(kz C / kh) cos(th) - (kx V' / kh^2) sin(th)
      v' =  V cos(th)
      w' =  (kx C / kh) cos(th) - (kz V' / kh^2) sin(th)
    with th = kx x + kz z + phi, V = s av tau(y), C = s ac tau(y),
    tau = 1 - eta^2, which is divergence free for every mode."""
    xi = gl_points().astype(F32)
    e = np.arange(nx * ny * nz)
    ex, ey, ez = e % nx, (e // nx) % ny, e // (nx * ny)
    pt = np.arange(64)
    x = _centered_coord(lx, nx, ex, xi[pt % 4])
    y = _centered_coord(ly, ny, ey, xi[(pt // 4) % 4])
    z = _centered_coord(lz, nz, ez, xi[pt // 16])
    mach = F32(mach_bulk)
    total = x.size
    if initial_condition == "uniform":
        u = np.full(total, mach, dtype=F32)
    else:
        eta = (y / F32(F32(0.5) * F32(ly))).astype(F32)
        u = (F32(F32(1.5) * mach) * _fma32(-eta, eta, F32(1.0))).astype(F32)
    rho = np.ones(total, dtype=F32)
    p = F32(F32(1.0) / F32(1.4))
    E = (p / F32(F32(1.4) - F32(1.0)) + (F32(0.5) * rho * (u * u)).astype(F32)).astype(F32)
    m1 = (rho * u).astype(F32)
    m2 = np.zeros(total, dtype=F32)
    m3 = np.zeros(total, dtype=F32)
    amp = F32(perturbation_amplitude)
    if amp != F32(0.0):
        n_modes = max(1, int(perturbation_modes))
        hw = F32(F32(0.5) * F32(ly))
        eta = (y / hw).astype(F32)
        tau = (F32(1.0) - eta * eta).astype(F32)
        tau_y = ((F32(-2.0) * eta) / hw).astype(F32)
        weight = F32(amp / np.sqrt(F32(n_modes), dtype=F32))
        du = np.zeros(total, dtype=F32)
        dv = np.zeros(total, dtype=F32)
        dw = np.zeros(total, dtype=F32)
        for md in trip_modes(n_modes, perturbation_seed, lx, lz):
            kx, kz = md["kx"], md["kz"]
            th = ((kx * x + kz * z) + md["phi"]).astype(F32)
            c, s = np.cos(th, dtype=F32), np.sin(th, dtype=F32)
            kh2 = F32(kx * kx + kz * kz)
            kh = F32(np.sqrt(kh2, dtype=F32))
            sw = F32(weight * md["wj"])
            V = (F32(sw * md["av"]) * tau).astype(F32)
            Vp = (F32(sw * md["av"]) * tau_y).astype(F32)
            C = (F32(sw * md["ac"]) * tau).astype(F32)
            du += ((F32(-1.0) * (kz * C / kh)) * c - (kx * Vp / kh2) * s).astype(F32)
            dv += (V * c).astype(F32)
            dw += ((kx * C / kh) * c - (kz * Vp / kh2) * s).astype(F32)
        n1 = (m1 + rho * du).astype(F32)
        n2 = (m2 + rho * dv).astype(F32)
        n3 = (m3 + rho * dw).astype(F32)
        E = (E + (F32(0.5) * ((n1 * n1 + n2 * n2 + n3 * n3) - (m1 * m1 + m2 * m2 + m3 * m3)) / rho)).astype(F32)
        m1, m2, m3 = n1, n2, n3
    return np.concatenate([rho, m1, m2, m3, E]).astype(dtype)
