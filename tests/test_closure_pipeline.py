"""Host-side pieces of the closure pipeline (opendynamo_amd/closure.py)
checked against closed forms: the fine-to-coarse L2 projection, the uniform
-grid interpolation, the energy spectrum and the subgrid target."""
import math

import numpy as np

from opendynamo_amd import closure as C
from helpers import coords


def test_projection_reproduces_coarse_polynomials_and_means():
    nf, r = 6, 3
    nc = nf // r
    L = 2.0
    X, Y, Z = coords(nf, nf, nf, L, L, L)
    # a cubic in the coarse element coordinates, per coarse element
    hc = L / nc
    xi = lambda q: (q % hc) / hc * 2 - 1
    f = 1 + 0.3 * xi(X) ** 3 - 0.2 * xi(Y) ** 2 * xi(Z) + 0.1 * xi(Z) ** 3 * xi(X) ** 0
    g = C.filter_field(f, nf, r)
    Xc, Yc, Zc = coords(nc, nc, nc, L, L, L)
    want = 1 + 0.3 * xi(Xc) ** 3 - 0.2 * xi(Yc) ** 2 * xi(Zc) + 0.1 * xi(Zc) ** 3
    assert np.max(np.abs(g - want)) < 1e-12
    # a non-polynomial field keeps its integral
    h = np.sin(3 * X) * np.cos(2 * Y) + np.exp(0.1 * Z)
    _, w = np.polynomial.legendre.leggauss(4)
    W = np.einsum("k,j,i->kji", w, w, w).ravel()
    integ = lambda v, n: np.sum(v.reshape(-1, 64) * W) * (L / n) ** 3 / 8
    assert abs(integ(C.filter_field(h, nf, r), nc) - integ(h, nf)) < 1e-12


def test_uniform_interpolation_and_spectrum_of_a_single_mode():
    n, L, A = 4, 2 * math.pi, 0.3
    X, Y, Z = coords(n, n, n, L, L, L)
    u = A * np.cos(2 * X + 1 * Y)
    E0 = 0.5 * 0 + 0
    rho = np.ones_like(X)
    state = np.concatenate([rho, rho * u, 0 * X, 0 * X, np.full_like(X, 2.5)])
    vel = C.uniform_velocity(state, n)
    N = 4 * n
    g = (np.arange(N) + 0.5) * L / N
    Zg, Yg, Xg = np.meshgrid(g, g, g, indexing="ij")
    exact = A * np.cos(2 * Xg + Yg)
    # P3 interpolation of a smooth mode: small, not exact
    assert np.max(np.abs(vel[0] - exact)) < 2e-2
    E = C.energy_spectrum(np.stack([exact, 0 * exact, 0 * exact]), L, backend="numpy")
    k = math.sqrt(5)
    assert abs(E.sum() - A * A / 4) < 1e-12  # Parseval: 1/2 <u^2> = A^2 / 4
    assert abs(E[int(round(k)) - 1] - A * A / 4) < 1e-12


def test_sgs_target_vanishes_for_uniform_flow():
    nf, r, L = 6, 3, 2.0
    X, Y, Z = coords(nf, nf, nf, L, L, L)
    rho = np.full_like(X, 1.1)
    state = np.concatenate([rho, rho * 0.2, rho * -0.1, rho * 0.05, np.full_like(X, 2.0)])
    cons, prod = C.filter_state(state, nf, r)
    mu = C.sgs_target(cons, prod, nf // r, L)
    assert np.all(mu == 0.0) or np.max(np.abs(mu)) < 1e-12


def test_sgs_target_recovers_a_constructed_eddy_viscosity():
    """tau_ij = -2 mu S^d_ij exactly -> the least-squares fit returns mu."""
    nc, L = 2, 2.0
    E = nc ** 3
    X, Y, Z = coords(nc, nc, nc, L, L, L)
    ux, uy, uz = 0.1 * X + 0.05 * Y, -0.02 * Z + 0.03 * X, 0.04 * Y
    rho = np.ones_like(X)
    cons = np.stack([rho, ux, uy, uz, np.full_like(X, 2.5)])
    G = np.array([[0.1, 0.05, 0.0], [0.03, 0.0, -0.02], [0.0, 0.04, 0.0]])
    S = 0.5 * (G + G.T)
    S -= np.eye(3) * np.trace(G) / 3
    mu = 0.007
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    vel = [ux, uy, uz]
    prod = np.stack([vel[a] * vel[b] - 2 * mu * S[a, b] for a, b in pairs])
    got = C.sgs_target(cons, prod, nc, L)
    assert np.allclose(got, mu, rtol=1e-10)
