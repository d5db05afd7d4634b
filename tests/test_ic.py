"""Initial-condition generator: closed forms and divergence-free fields."""
import numpy as np

from oracle import oracle as O
from opendynamo_amd import ic


def test_fourier_modes_are_divergence_free_and_periodic():
    lx, ly, lz = 2 * np.pi, 2.0, np.pi
    ks, amps, _ = ic.fourier_modes(32, 5, lx, ly, lz)
    assert np.max(np.abs(np.einsum("md,md->m", ks, amps))) < 1e-14
    n = ks * np.array([lx, ly, lz]) / (2 * np.pi)
    np.testing.assert_allclose(n, np.round(n), atol=1e-12)
    np.testing.assert_allclose(np.linalg.norm(amps, axis=1), 1.0, rtol=1e-14)


def test_fourier_state_rms_and_base_flow():
    L = 2 * np.pi
    u = ic.fourier_state(8, 8, 8, L, L, L, u0=(0.1, 0, 0), amplitude=1e-2, n_modes=8, seed=0).reshape(5, -1)
    rho = u[0]
    np.testing.assert_array_equal(rho, 1.0)
    vel = u[1:4] / rho
    pert = vel - np.array([[0.1], [0], [0]])
    rms = np.sqrt(np.mean(np.sum(pert**2, axis=0)))
    assert 0.5e-2 < rms < 1.5e-2
    ke = 0.5 * rho * np.sum(vel**2, axis=0)
    np.testing.assert_allclose(0.4 * (u[4] - ke), 1 / 1.4, rtol=1e-13)


def test_laminar_channel_closed_form():
    nx, ny, nz, lx, ly, lz, M = 2, 4, 2, 2 * np.pi, 2.0, np.pi, 0.1
    u = ic.channel_state(nx, ny, nz, lx, ly, lz, M).reshape(5, -1)
    _, y, _ = ic.point_coords(nx, ny, nz, lx, ly, lz, y_center=True)
    exact = 1.5 * M * (1 - y * y)
    np.testing.assert_allclose(u[1], exact, rtol=2e-6, atol=1e-7)
    assert np.all(u[2] == 0) and np.all(u[3] == 0)
    uni = ic.channel_state(nx, ny, nz, lx, ly, lz, M, initial_condition="uniform").reshape(5, -1)
    np.testing.assert_array_equal(uni[1], np.float32(M))


def _trip_divergence(nx, ny, nz, lx, ly, lz, modes):
    st = ic.channel_state(nx, ny, nz, lx, ly, lz, 0.1, perturbation_amplitude=0.05, perturbation_modes=modes,
                          perturbation_seed=1, dtype=np.float64).reshape(5, -1)
    base = ic.channel_state(nx, ny, nz, lx, ly, lz, 0.1, dtype=np.float64).reshape(5, -1)
    vel = (st[1:4] - base[1:4]) / st[0]
    D = O.fr_operators()["d"].reshape(4, 4)
    jac = [2 / (lx / nx), 2 / (ly / ny), 2 / (lz / nz)]
    v = vel.reshape(3, nz, ny, nx, 4, 4, 4)   # (comp, ez, ey, ex, k, j, i)
    div = 0.0
    for d in range(3):
        ax = 5 - d                              # i, j, k axis of v[d]
        div = div + np.moveaxis(np.moveaxis(v[d], ax, -1) @ D.T, -1, ax) * jac[d]
    return np.max(np.abs(div))


def test_trip_is_divergence_free_to_truncation():
    """The element-local FR divergence of the trip velocity converges to zero
    at high order under refinement (a field with a true divergence would
    plateau at that divergence)."""
    lx, ly, lz = 2 * np.pi, 2.0, np.pi
    coarse = _trip_divergence(36, 4, 18, lx, ly, lz, modes=2)
    fine = _trip_divergence(72, 4, 36, lx, ly, lz, modes=2)
    # element-local derivative error is O(h^P), P = 3: the ratio tends to 8
    assert coarse / fine > 5.0, (coarse, fine)


def test_trip_modes_are_deterministic():
    a = ic.trip_modes(16, 1, 2 * np.pi, np.pi)
    b = ic.trip_modes(16, 1, 2 * np.pi, np.pi)
    assert a == b
    assert all(1 <= m["mx"] <= 24 and 1 <= m["mz"] <= 12 for m in a)
