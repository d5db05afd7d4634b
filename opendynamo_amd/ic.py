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
