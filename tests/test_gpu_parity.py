"""GPU parity: the HIP library against the CPU oracle on the same inputs.

Strict IEEE build on both sides (no contraction, correctly rounded / and
sqrt), identical association order -> the bar is bit equality; the looser
bound of 1e-13 relative per variable after 100 f64 steps is asserted as
well.
"""
import itertools

import numpy as np
import pytest

import opendynamo_amd as od
from helpers import rel_err_per_var, smooth_state, uniform_state
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TWO_PI = 2 * np.pi


def both(nx, ny, nz, lx, ly, lz, **kw):
    """Matching od_params / orc_params."""
    okw = dict(kw)
    okw.pop("fast_math", None)
    po = O.params(nx, ny, nz, lx, ly, lz, **okw)
    pg = od.params(nx, ny, nz, lx, ly, lz, **kw)
    return pg, po


def announce(**case):
    print("CASE", " ".join(f"{k}={v}" for k, v in case.items()), flush=True)


CASES = [
    # (flux, wall, viscous, viscosity_model)
    ("rusanov", False, False, 0),
    ("rusanov", False, True, 0),
    ("slau2", True, True, 0),
    ("ausm+", True, True, 1),
    ("slau2", False, True, 1),
    ("ausm+", False, False, 0),
    ("slau2", True, False, 0),
]


@pytest.mark.parametrize("flux,wall,viscous,visc_model", CASES)
@pytest.mark.parametrize("real", [8, 4])
def test_rhs_bitwise(flux, wall, viscous, visc_model, real):
    nx, ny, nz = 3, 4, 2
    lx, ly, lz = TWO_PI, 2.0, np.pi
    announce(test="rhs", flux=flux, wall=wall, viscous=viscous, visc_model=visc_model, real=real)
    pg, po = both(nx, ny, nz, lx, ly, lz, u_real=real, flux=flux, y_wall=wall, viscous=viscous,
                  viscosity_model=visc_model, mu_ref=2e-3)
    u = smooth_state(nx, ny, nz, lx, ly, lz, seed=3, amp=0.08, dtype=np.float64 if real == 8 else np.float32)
    ref = O.rhs_total(po, u)
    with od.Solver(pg) as s:
        s.set_state(u)
        got = s.rhs()
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("flux,wall,viscous,visc_model", CASES)
@pytest.mark.parametrize("real", [8, 4])
def test_steps_bitwise(flux, wall, viscous, visc_model, real):
    nx, ny, nz = 2, 3, 2
    lx, ly, lz = TWO_PI, 2.0, np.pi
    announce(test="steps", flux=flux, wall=wall, viscous=viscous, visc_model=visc_model, real=real)
    pg, po = both(nx, ny, nz, lx, ly, lz, u_real=real, flux=flux, y_wall=wall, viscous=viscous,
                  viscosity_model=visc_model, mu_ref=2e-3)
    u = smooth_state(nx, ny, nz, lx, ly, lz, seed=5, amp=0.05, dtype=np.float64 if real == 8 else np.float32)
    dt, n = 2e-3, 7
    ref = O.step(po, u, dt, n)
    with od.Solver(pg) as s:
        s.set_state(u)
        s.step(dt, n)
        got = s.get_state()
    np.testing.assert_array_equal(got, ref)


PRECISIONS = [
    # u, q, rhs, compute, rk_accum
    (8, 4, 4, 8, 0),
    (8, 8, 8, 4, 0),
    (8, 8, 8, 4, 1),
    (4, 8, 8, 8, 0),
    (4, 4, 8, 8, 1),
]


@pytest.mark.parametrize("u_real,q_real,rhs_real,compute_real,rk_accum", PRECISIONS)
def test_mixed_precision_bitwise(u_real, q_real, rhs_real, compute_real, rk_accum):
    nx, ny, nz = 2, 2, 3
    lx, ly, lz = TWO_PI, 2.0, np.pi
    announce(test="mixed", u=u_real, q=q_real, rhs=rhs_real, c=compute_real, accum=rk_accum)
    pg, po = both(nx, ny, nz, lx, ly, lz, u_real=u_real, q_real=q_real, rhs_real=rhs_real,
                  compute_real=compute_real, rk_accum=rk_accum, flux="slau2", y_wall=True, viscous=True,
                  viscosity_model=1, mu_ref=1e-3)
    u = smooth_state(nx, ny, nz, lx, ly, lz, seed=9, amp=0.05, dtype=np.float64 if u_real == 8 else np.float32)
    ref = O.step(po, u, 1e-3, 5)
    with od.Solver(pg) as s:
        s.set_state(u)
        s.step(1e-3, 5)
        got = s.get_state()
    np.testing.assert_array_equal(got, ref)


def test_euler8_periodic_f64_100_steps():
    """8^3 periodic Euler, Rusanov, f64, CFL 0.3, 100 steps: <= 1e-13 relative per variable."""
    from opendynamo_amd import ic
    nx = ny = nz = 8
    L = float(np.float32(TWO_PI))
    announce(test="euler 8^3 periodic f64", steps=100)
    pg, po = both(nx, ny, nz, L, L, L, u_real=8, flux="rusanov", y_wall=False, viscous=False)
    u = ic.fourier_state(nx, ny, nz, L, L, L, u0=(0.1, 0.0, 0.0), amplitude=1e-2, n_modes=8, seed=0)
    dt = O.cfl_dt(po, u, 0.3)
    with od.Solver(pg) as s:
        s.set_state(u)
        assert s.cfl_dt(0.3) == dt
        s.step(dt, 100)
        got = s.get_state()
    ref = O.step(po, u, dt, 100)
    err = rel_err_per_var(got, ref)
    print("euler 8^3 f64 rel err per var", err, flush=True)
    assert np.all(err <= 1e-13)
    np.testing.assert_array_equal(got, ref)


def test_deterministic_and_resume():
    """Byte-identical run to run, and N + checkpoint + N == 2N."""
    nx, ny, nz = 4, 4, 4
    lx, ly, lz = TWO_PI, 2.0, np.pi
    announce(test="determinism")
    pg = od.params(nx, ny, nz, lx, ly, lz, u_real=4, flux="slau2", y_wall=True, viscosity_model=1, mu_ref=1e-3)
    u = smooth_state(nx, ny, nz, lx, ly, lz, seed=1, amp=0.05, dtype=np.float32)
    runs = []
    for _ in range(2):
        with od.Solver(pg) as s:
            s.set_state(u)
            s.step(1e-3, 20)
            runs.append(s.get_state())
    assert runs[0].tobytes() == runs[1].tobytes()
    with od.Solver(pg) as s:
        s.set_state(u)
        s.step(1e-3, 10)
        mid = s.get_state()
    with od.Solver(pg) as s:
        s.set_state(mid)
        s.step(1e-3, 10)
        resumed = s.get_state()
    assert resumed.tobytes() == runs[0].tobytes()


def test_conservation_sums_and_async_copy():
    nx, ny, nz = 3, 3, 3
    lx, ly, lz = TWO_PI, TWO_PI, TWO_PI
    announce(test="sums")
    for real in (8, 4):
        pg, po = both(nx, ny, nz, lx, ly, lz, u_real=real, flux="rusanov", viscous=False)
        u = smooth_state(nx, ny, nz, lx, ly, lz, seed=2, dtype=np.float64 if real == 8 else np.float32)
        with od.Solver(pg) as s:
            s.set_state(u)
            got = s.sums()
            s.step(1e-3, 3)
            t = s.get_state_async()
            snap = s.wait(t)
            sync = s.get_state()
        np.testing.assert_allclose(got, O.sums(po, u), rtol=1e-13, atol=1e-13)
        assert snap.tobytes() == sync.tobytes()


def test_uniform_flow_free_stream():
    """Uniform state: every face correction vanishes, RHS = body force up to D row sums."""
    nx, ny, nz = 2, 2, 2
    announce(test="freestream")
    pg, po = both(nx, ny, nz, TWO_PI, TWO_PI, TWO_PI, u_real=8, flux="slau2", viscous=True)
    u = uniform_state(nx, ny, nz)
    with od.Solver(pg) as s:
        s.set_state(u)
        r = s.rhs().reshape(5, -1)
    np.testing.assert_array_equal(r.ravel(), O.rhs_total(po, u))
    assert np.all(np.abs(r[1] - 0.001) < 1e-14)
    assert np.all(np.abs(r[[0, 2, 3, 4]]) < 1e-14)


# ---------------------------------------------------------------------------
# fast_math = true (the reference's default): f32 kernels with reciprocal
# multiplies, hardware sqrt and contraction.  Not bitwise: checked against
# the FP64 oracle with a stated tolerance, and for determinism / resume.
# ---------------------------------------------------------------------------
FAST_TOL = 2e-5  # max |gpu - oracle_f64| / max |oracle_f64| per variable


# odd nx: two-element tiles; all dimensions even: 2x2x2-element block kernels
FAST_DIMS = [(3, 4, 2), (4, 4, 2)]


@pytest.mark.parametrize("dims", FAST_DIMS, ids=["tiles", "blocks"])
@pytest.mark.parametrize("flux,wall,viscous,visc_model", CASES)
def test_fast_math_against_fp64_oracle(flux, wall, viscous, visc_model, dims):
    nx, ny, nz = dims
    lx, ly, lz = TWO_PI, 2.0, np.pi
    announce(test="fast", dims=dims, flux=flux, wall=wall, viscous=viscous, visc_model=visc_model)
    pg = od.params(nx, ny, nz, lx, ly, lz, u_real=4, flux=flux, y_wall=wall, viscous=viscous,
                   viscosity_model=visc_model, mu_ref=2e-3, fast_math=1)
    po = O.params(nx, ny, nz, lx, ly, lz, u_real=8, flux=flux, y_wall=wall, viscous=viscous,
                  viscosity_model=visc_model, mu_ref=2e-3)
    u = smooth_state(nx, ny, nz, lx, ly, lz, seed=5, amp=0.05, dtype=np.float32)
    ref = O.step(po, u.astype(np.float64), 1e-3, 5)
    with od.Solver(pg) as s:
        s.set_state(u)
        s.step(1e-3, 5)
        got = s.get_state()
    err = rel_err_per_var(got, ref)
    print("fast-math rel err per var", err, flush=True)
    assert np.all(err < FAST_TOL), err


FAST_RHS_TOL = 1e-4  # max |gpu - oracle_f64| / max |oracle_f64| per variable of the RHS


@pytest.mark.parametrize("dims", FAST_DIMS, ids=["tiles", "blocks"])
@pytest.mark.parametrize("flux,wall,viscous,visc_model", CASES)
def test_fast_rhs_against_fp64_oracle(flux, wall, viscous, visc_model, dims):
    """od_rhs of the fast build (the update kernels' RHS-only mode) against
    the FP64 oracle's corrected RHS, tiles and blocks."""
    nx, ny, nz = dims
    lx, ly, lz = TWO_PI, 2.0, np.pi
    announce(test="fast rhs", dims=dims, flux=flux, wall=wall, viscous=viscous, visc_model=visc_model)
    pg = od.params(nx, ny, nz, lx, ly, lz, u_real=4, flux=flux, y_wall=wall, viscous=viscous,
                   viscosity_model=visc_model, mu_ref=2e-3, fast_math=1)
    po = O.params(nx, ny, nz, lx, ly, lz, u_real=8, flux=flux, y_wall=wall, viscous=viscous,
                  viscosity_model=visc_model, mu_ref=2e-3)
    u = smooth_state(nx, ny, nz, lx, ly, lz, seed=3, amp=0.08, dtype=np.float32)
    ref = O.rhs_total(po, u.astype(np.float64))
    with od.Solver(pg) as s:
        s.set_state(u)
        got = s.rhs()
    err = rel_err_per_var(got, ref)
    print("fast-math rhs rel err per var", err, flush=True)
    assert np.all(err < FAST_RHS_TOL), err


def test_fast_math_deterministic_and_resume():
    nx, ny, nz = 4, 4, 4
    lx, ly, lz = TWO_PI, 2.0, np.pi
    announce(test="fast determinism")
    pg = od.params(nx, ny, nz, lx, ly, lz, u_real=4, flux="slau2", y_wall=True, viscosity_model=1, mu_ref=1e-3,
                   fast_math=1)
    u = smooth_state(nx, ny, nz, lx, ly, lz, seed=1, amp=0.05, dtype=np.float32)
    with od.Solver(pg) as s:
        s.set_state(u)
        s.step(1e-3, 20)
        full = s.get_state()
    with od.Solver(pg) as s:
        s.set_state(u)
        s.step(1e-3, 10)
        mid = s.get_state()
    with od.Solver(pg) as s:
        s.set_state(mid)
        s.step(1e-3, 10)
        resumed = s.get_state()
    assert resumed.tobytes() == full.tobytes()


@pytest.mark.parametrize("real,fast", [(4, 1), (8, 0)])
def test_multi_step_graph_equals_single_steps(real, fast):
    """od_step runs n steps as replays of an 8-step graph plus one-step
    replays; the state is byte-identical to n separate one-step calls and
    to any other split of the same n."""
    nx, ny, nz = 4, 3, 5
    lx, ly, lz = TWO_PI, 2.0, np.pi
    announce(test="multi_step_graph", real=real, fast=fast)
    pg = od.params(nx, ny, nz, lx, ly, lz, u_real=real, flux="slau2", y_wall=True, viscosity_model=1, mu_ref=1e-3,
                   fast_math=fast)
    u = smooth_state(nx, ny, nz, lx, ly, lz, seed=4, amp=0.05, dtype=np.float32 if real == 4 else np.float64)
    out = []
    for split in ([17], [1] * 17, [8, 9], [3, 8, 6]):
        with od.Solver(pg) as s:
            s.set_state(u)
            for n in split:
                s.step(1e-3, n)
            out.append(s.get_state())
    assert np.isfinite(out[0]).all()
    for o in out[1:]:
        assert o.tobytes() == out[0].tobytes()
