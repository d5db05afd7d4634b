"""Shared test helpers: synthetic states on the FR P=3 point layout."""
import numpy as np

GL = np.array([-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526])


def coords(nx, ny, nz, lx, ly, lz, y0=0.0):
    """Point coordinates in the var-plane layout: elem-major 64-point blocks,
    pt = i + 4j + 16k, elem = ex + nx*ey + nx*ny*ez."""
    e = np.arange(nx * ny * nz)
    ex, ey, ez = e % nx, (e // nx) % ny, e // (nx * ny)
    pt = np.arange(64)
    i, j, k = pt % 4, (pt // 4) % 4, pt // 16
    dx, dy, dz = lx / nx, ly / ny, lz / nz
    X = (ex[:, None] + 0.5 + 0.5 * GL[i][None, :]) * dx
    Y = (ey[:, None] + 0.5 + 0.5 * GL[j][None, :]) * dy + y0
    Z = (ez[:, None] + 0.5 + 0.5 * GL[k][None, :]) * dz
    return X.ravel(), Y.ravel(), Z.ravel()


def _fourier(rng, X, Y, Z, lx, ly, lz, n_modes=3):
    f = np.zeros_like(X)
    for _ in range(n_modes):
        kx, ky, kz = rng.integers(0, 3, size=3)
        ph = rng.uniform(0, 2 * np.pi)
        f += rng.uniform(-1, 1) * np.cos(2 * np.pi * (kx * X / lx + ky * Y / ly + kz * Z / lz) + ph)
    return f / n_modes


def smooth_state(nx, ny, nz, lx, ly, lz, seed=0, amp=0.05, mean_u=0.1, dtype=np.float64):
    """A smooth, positive-pressure periodic state, var-major [5][N]."""
    rng = np.random.default_rng(seed)
    X, Y, Z = coords(nx, ny, nz, lx, ly, lz)
    rho = 1.0 + amp * _fourier(rng, X, Y, Z, lx, ly, lz)
    u = mean_u + amp * _fourier(rng, X, Y, Z, lx, ly, lz)
    v = amp * _fourier(rng, X, Y, Z, lx, ly, lz)
    w = amp * _fourier(rng, X, Y, Z, lx, ly, lz)
    p = (1.0 / 1.4) * (1.0 + amp * _fourier(rng, X, Y, Z, lx, ly, lz))
    E = p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)
    return np.concatenate([rho, rho * u, rho * v, rho * w, E]).astype(dtype)


def uniform_state(nx, ny, nz, rho=1.0, u=0.1, v=0.0, w=0.0, p=1.0 / 1.4, dtype=np.float64):
    n = nx * ny * nz * 64
    E = p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)
    vals = [rho, rho * u, rho * v, rho * w, E]
    return np.concatenate([np.full(n, x) for x in vals]).astype(dtype)


def rel_err_per_var(a, b):
    """max |a-b| / max |b| for each of the 5 variable planes."""
    a = np.asarray(a, dtype=np.float64).reshape(5, -1)
    b = np.asarray(b, dtype=np.float64).reshape(5, -1)
    out = []
    for v in range(5):
        scale = np.max(np.abs(b[v]))
        out.append(np.max(np.abs(a[v] - b[v])) / (scale if scale > 0 else 1.0))
    return np.array(out)
