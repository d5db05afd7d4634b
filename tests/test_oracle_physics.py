"""Pin the oracle's point physics (device_functions.cuh restated) with
known answers, consistency and symmetry, in f64 and f32."""
import numpy as np
import pytest

from oracle import oracle as O

GAMMA = 1.4


def cons(rho, u, v, w, p):
    return np.array([rho, rho * u, rho * v, rho * w, p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v + w * w)])


def euler_exact(s, d):
    rho, m = s[0], s[1:4]
    vel = m / rho
    p = (GAMMA - 1) * (s[4] - 0.5 * rho * vel @ vel)
    f = np.zeros(5)
    f[0] = m[d]
    f[1:4] = m * vel[d]
    f[1 + d] += p
    f[4] = vel[d] * (s[4] + p)
    return f


def test_pressure_known_answer():
    s = cons(1.2, 0.3, -0.2, 0.1, 0.9)
    assert abs(O.point(8, 0, s)[0] - 0.9) < 1e-15
    # clamp at 1e-7 (device_functions.cuh:4)
    s_bad = np.array([1.0, 1.0, 0.0, 0.0, 0.1])
    assert O.point(8, 0, s_bad)[0] == 1e-7


@pytest.mark.parametrize("d", [0, 1, 2])
def test_euler_flux_known_answer(d):
    s = cons(1.1, 0.3, -0.4, 0.25, 0.75)
    np.testing.assert_allclose(O.point(8, 1 + d, s), euler_exact(s, d), rtol=1e-15, atol=1e-16)


@pytest.mark.parametrize("scheme", [0, 1, 2])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_riemann_consistency(scheme, d):
    """F(U, U, n) = F(U) . n for SLAU2, AUSM+ and Rusanov."""
    rng = np.random.default_rng(10 * scheme + d)
    n = np.zeros(3)
    n[d] = 1.0
    for _ in range(20):
        s = cons(rng.uniform(0.5, 2), *rng.uniform(-0.5, 0.5, 3), rng.uniform(0.3, 2))
        got = O.point(8, 7 + scheme, np.concatenate([s, s, n]))
        ref = euler_exact(s, d)
        np.testing.assert_allclose(got, ref, rtol=2e-15, atol=2e-15 * np.abs(ref).max())


def mirror(s, d):
    t = s.copy()
    t[1 + d] = -t[1 + d]
    return t


@pytest.mark.parametrize("scheme", [0, 1, 2])
def test_riemann_mirror_symmetry(scheme):
    """Reflecting the problem across the face flips the normal-odd fluxes."""
    rng = np.random.default_rng(7 + scheme)
    d = 0
    n = np.array([1.0, 0.0, 0.0])
    for _ in range(20):
        sl = cons(rng.uniform(0.5, 2), *rng.uniform(-0.6, 0.6, 3), rng.uniform(0.3, 2))
        sr = cons(rng.uniform(0.5, 2), *rng.uniform(-0.6, 0.6, 3), rng.uniform(0.3, 2))
        f = O.point(8, 7 + scheme, np.concatenate([sl, sr, n]))
        g = O.point(8, 7 + scheme, np.concatenate([mirror(sr, d), mirror(sl, d), n]))
        sign = np.array([-1, 1, -1, -1, -1])
        np.testing.assert_allclose(g, sign * f, rtol=1e-13, atol=1e-14)


def test_rusanov_sod_known_answer():
    """Sod states, n = x: lambda = sqrt(1.4) from the left state."""
    sl, sr = cons(1.0, 0, 0, 0, 1.0), cons(0.125, 0, 0, 0, 0.1)
    f = O.point(8, 9, np.concatenate([sl, sr, [1, 0, 0]]))
    lam = np.sqrt(1.4)
    exact = np.array([0.5 * lam * 0.875, 0.5 * (1.0 + 0.1), 0, 0, 0.5 * lam * (1.0 / 0.4 - 0.1 / 0.4)])
    np.testing.assert_allclose(f, exact, rtol=1e-15, atol=1e-16)


def test_upwinding_supersonic():
    """Both states supersonic to the right: AUSM+ and Rusanov-free upwind flux = F(U_L)."""
    sl = cons(1.0, 2.5, 0.1, -0.1, 1 / 1.4)   # c = 1, M = 2.5
    sr = cons(0.9, 2.2, 0.0, 0.2, 0.8 / 1.4)
    f = O.point(8, 8, np.concatenate([sl, sr, [1, 0, 0]]))  # AUSM+
    np.testing.assert_allclose(f, euler_exact(sl, 0), rtol=1e-15, atol=1e-15)


def test_viscous_couette():
    """u = a y: tau_xy = mu a, no heat flux (uniform p, rho)."""
    mu, a, u0 = 0.01, 0.3, 0.2
    s = cons(1.0, u0, 0, 0, 1 / 1.4)
    g = np.zeros(15)
    g[1 * 3 + 1] = a            # d(rho u)/dy
    g[4 * 3 + 1] = 1.0 * u0 * a  # dE/dy = d(rho u^2/2)/dy
    fy = O.point(8, 5, np.concatenate([s, g, [mu]]))
    np.testing.assert_allclose(fy, [0, mu * a, 0, 0, u0 * mu * a], rtol=1e-15, atol=1e-18)
    fx = O.point(8, 4, np.concatenate([s, g, [mu]]))
    np.testing.assert_allclose(fx, [0, 0, mu * a, 0, 0], rtol=1e-15, atol=1e-18)


def test_viscous_heat_conduction():
    """p = p0 + b x at rest: F_v,x = (0,0,0,0, mu gamma/(Pr (gamma-1)) b / rho)."""
    mu, b = 0.02, 0.05
    s = cons(1.0, 0, 0, 0, 1 / 1.4)
    g = np.zeros(15)
    g[4 * 3 + 0] = b / (GAMMA - 1)
    f = O.point(8, 4, np.concatenate([s, g, [mu]]))
    coeff = mu * GAMMA / (0.72 * (GAMMA - 1))
    np.testing.assert_allclose(f, [0, 0, 0, 0, coeff * b], rtol=1e-15, atol=1e-18)


def test_viscous_dilatation():
    """u = a x: tau_xx = mu (2a - 2a/3) = 4 mu a / 3, tau_yy = -2 mu a / 3."""
    mu, a = 0.01, 0.4
    s = cons(1.0, 0.0, 0, 0, 1 / 1.4)
    g = np.zeros(15)
    g[1 * 3 + 0] = a
    fx = O.point(8, 4, np.concatenate([s, g, [mu]]))
    fy = O.point(8, 5, np.concatenate([s, g, [mu]]))
    np.testing.assert_allclose(fx[1], 4 * mu * a / 3, rtol=1e-15)
    np.testing.assert_allclose(fy[2], -2 * mu * a / 3, rtol=1e-15)


def test_sutherland():
    mu_ref, t_ref, s_suth = 3e-5, 273.15, 110.4
    s = cons(1.0, 0.1, 0, 0, 1 / 1.4)   # gamma p / rho = 1
    got = O.point(8, 10, np.concatenate([s, [mu_ref, t_ref, s_suth]]))[0]
    assert abs(got - mu_ref) <= 2e-16 * mu_ref
    s4 = cons(1.0, 0.1, 0, 0, 4 / 1.4)  # theta = 4
    got4 = O.point(8, 10, np.concatenate([s4, [mu_ref, t_ref, s_suth]]))[0]
    np.testing.assert_allclose(got4, mu_ref * 8 * (t_ref + s_suth) / (4 * t_ref + s_suth), rtol=4e-16)


def test_f32_matches_f64_to_single_precision():
    rng = np.random.default_rng(3)
    for kind in range(1, 10):
        sl = cons(rng.uniform(0.8, 1.2), *rng.uniform(-0.3, 0.3, 3), rng.uniform(0.6, 0.9))
        sr = cons(rng.uniform(0.8, 1.2), *rng.uniform(-0.3, 0.3, 3), rng.uniform(0.6, 0.9))
        if kind <= 3:
            inp = sl
        elif kind <= 6:
            inp = np.concatenate([sl, rng.uniform(-0.1, 0.1, 15), [1e-3]])
        else:
            inp = np.concatenate([sl, sr, [0, 1, 0]])
        a, b = O.point(4, kind, inp), O.point(8, kind, inp)
        np.testing.assert_allclose(a, b, rtol=2e-6, atol=2e-6 * np.abs(b).max())
