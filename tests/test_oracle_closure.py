"""Pins the closure oracle (oracle/mlp.py) without a GPU: finite differences
of the MLP gradient, hand values of GELU/softplus, BF16 rounding, and closed
forms of the element features."""
import math

import numpy as np
import pytest

from oracle import mlp as M
from helpers import coords


def random_params(seed=0, scale=0.3):
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, scale, M.NPARAM)


def test_bf16_rounding_known_values():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-3], dtype=np.float32)
    r = M.bf16(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0 + 2 ** -7 or r[1] == 1.0  # exactly half way below 2^-7 spacing? check tie rule below
    # 1 + 2^-9 is below half of the bf16 spacing (2^-7) -> rounds down; 1 + 3 * 2^-9 rounds up
    assert r[2] == 1.0 and r[3] == 1.0 + 2 ** -7
    assert r[4] == -2.5
    # ties to even: 1 + 2^-8 is exactly half way between 1 and 1 + 2^-7; the even neighbour is 1
    assert r[1] == 1.0
    assert abs(r[5] - 3.0e-3) / 3.0e-3 < 2 ** -8


def test_activation_values():
    assert M.gelu(np.array([0.0]))[0] == 0.0
    # tanh-GELU at 1: 0.5 * (1 + tanh(sqrt(2/pi) * 1.044715))
    want = 0.5 * (1 + math.tanh(math.sqrt(2 / math.pi) * 1.044715))
    assert abs(M.gelu(np.array([1.0]))[0] - want) < 1e-15
    # close to the erf form (within the known ~1e-3 of the approximation)
    for x in (-2.0, -0.5, 0.7, 1.5, 3.0):
        exact = 0.5 * x * (1 + math.erf(x / math.sqrt(2)))
        assert abs(M.gelu(np.array([x]))[0] - exact) < 2e-3
    assert abs(M.softplus(np.array([0.0]))[0] - math.log(2.0)) < 1e-15
    assert M.softplus(np.array([30.0]))[0] == 30.0
    xs = np.linspace(-3, 3, 13)
    h = 1e-6
    num = (M.gelu(xs + h) - M.gelu(xs - h)) / (2 * h)
    assert np.max(np.abs(num - M.gelu_d(xs))) < 1e-8


def test_loss_grad_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = random_params(1)
    x = rng.normal(0, 1, (37, M.NF))
    t = rng.uniform(0, 1, 37)
    loss, g = M.loss_grad(p, x, t, rounding=False)
    idx = rng.choice(M.NPARAM, 60, replace=False)
    idx = np.concatenate([idx, [M.OFF["b4"][0], M.OFF["w4"][0], M.OFF["b1"][0], M.OFF["W1"][0] + 5]])
    h = 1e-6
    for q in idx:
        pp, pm = p.copy(), p.copy()
        pp[q] += h
        pm[q] -= h
        num = (M.loss_grad(pp, x, t, rounding=False)[0] - M.loss_grad(pm, x, t, rounding=False)[0]) / (2 * h)
        assert abs(num - g[q]) <= 1e-7 + 1e-5 * abs(num), (q, num, g[q])


def test_rounded_gradient_close_to_exact():
    """The BF16 path is a perturbation of the exact gradient, not another function."""
    rng = np.random.default_rng(4)
    p = random_params(2)
    x = rng.normal(0, 1, (256, M.NF))
    t = rng.uniform(0, 1, 256)
    l0, g0 = M.loss_grad(p, x, t, rounding=False)
    l1, g1 = M.loss_grad(p, x, t, rounding=True)
    assert abs(l1 - l0) / l0 < 2e-2
    assert np.linalg.norm(g1 - g0) / np.linalg.norm(g0) < 5e-2


def _element_state(vel_fn, nx=2, ny=2, nz=2, L=(2.0, 2.0, 2.0), rho=1.0, p=1 / 1.4):
    X, Y, Z = coords(nx, ny, nz, *L)
    u, v, w = vel_fn(X, Y, Z)
    E = p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)
    r = np.full_like(X, rho)
    return np.concatenate([r, r * u, r * v, r * w, E])


def test_features_linear_field_closed_form():
    G = np.array([[0.01, 0.02, -0.005], [0.0, -0.01, 0.015], [0.003, 0.0, 0.004]])

    def vel(X, Y, Z):
        return (0.1 + G[0, 0] * X + G[0, 1] * Y + G[0, 2] * Z, G[1, 0] * X + G[1, 1] * Y + G[1, 2] * Z,
                G[2, 0] * X + G[2, 1] * Y + G[2, 2] * Z)

    nx = ny = nz = 2
    L = (2.0, 2.0, 2.0)
    f = M.features(_element_state(vel, nx, ny, nz, L), nx, ny, nz, *L, mu_mol=1e-3)
    div = np.trace(G)
    S = 0.5 * (G + G.T)
    Sd = S - np.eye(3) * div / 3
    Om = 0.5 * (G - G.T)
    smag = math.sqrt(2 * (Sd ** 2).sum())
    omag = math.sqrt(2 * (Om ** 2).sum())
    Q = 0.5 * ((Om ** 2).sum() - (S ** 2).sum())
    R = -np.trace(G @ G @ G) / 3
    dl = 1.0 / 4.0
    for e in range(8):
        assert abs(f[e, 1] - omag / smag) < 1e-9
        assert abs(f[e, 2] - Q / smag ** 2) < 1e-9
        assert abs(f[e, 3] - R / smag ** 3) < 1e-9
        assert abs(f[e, 4] - div / smag) < 1e-9
        # a linear field has only degree-1 modes
        assert abs(f[e, 5] - 1.0) < 1e-9 and abs(f[e, 6]) < 1e-12 and abs(f[e, 7]) < 1e-12
        assert abs(f[e, 0] - math.log(1.0 * dl * dl * smag / 1e-3)) < 1e-9
        assert abs(f[e, 8]) < 1e-12
        assert abs(f[e, 13] - dl * dl * smag) < 1e-12
        assert abs(f[e, 15] - smag ** 2 * 1.0) < 1e-9  # element volume 1, |S|^2 uniform


def test_features_single_mode_energy_fraction():
    # u_x = P2(xi) inside each element (xi along x), other components zero
    def vel(X, Y, Z):
        xi = (X % 1.0) * 2.0 - 1.0
        return 0.5 * (3 * xi * xi - 1) * 0.01, 0 * X, 0 * X

    f = M.features(_element_state(vel, 2, 2, 2, (2.0, 2.0, 2.0)), 2, 2, 2, 2.0, 2.0, 2.0, mu_mol=1e-3)
    assert np.allclose(f[:, 6], 1.0, atol=1e-9) and np.allclose(f[:, 5], 0.0, atol=1e-9)
    assert np.allclose(f[:, 7], 0.0, atol=1e-9)
    # kinetic energy: 0.5 * 1e-4 * int P2^2 = 0.5 * 1e-4 * (2/5) * (element volume 1 / 2) per unit reference
    ke = 0.5 * 1e-4 * (2.0 / 5.0) * 0.5
    assert np.allclose(f[:, 14], ke, rtol=1e-9)


def test_mu_t_range_check():
    p = random_params(5, 0.1)
    rng = np.random.default_rng(6)
    feat = rng.normal(0, 1, (20, 16))
    feat[:, 12] = 1.0
    feat[:, 13] = 0.01
    lo, hi = np.full(12, -1.5), np.full(12, 1.5)
    mu, inr = M.mu_t_from_features(p, feat, np.zeros(12), np.ones(12), lo, hi)
    assert np.all(mu[~inr] == 0.0) and np.all(mu[inr] > 0.0)
    assert inr.sum() == np.all(np.abs(feat[:, :12]) <= 1.5, axis=1).sum()
