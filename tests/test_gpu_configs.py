"""GPU tests on two named workloads at their real sizes.

* The 8^3 precision study: periodic [0, 2pi)^3 Euler, Rusanov, CFL 0.3,
  uniform flow u = (0.1, 0, 0) plus a divergence-free 8-mode Fourier field
  (amplitude 1e-2, seed 0).  The mixed pipeline (f64 state and RK
  accumulation, f32 flux / Riemann / derivatives) must equal the oracle run
  with the same types bit for bit after 10 and 100 steps; its per-variable
  error against the FP64 oracle is printed and bounded.
* The 24-box ensemble: periodic cubes with 16..40 elements per side (seed 7),
  P=3 NS f32 fast math, SLAU2, each box with its own Fourier IC and CFL dt.
  Under the autotuned schedule every box must equal the same box advanced
  alone, byte for byte, and the smallest box must agree with the FP64
  oracle within the fast build's tolerance.
"""
import numpy as np
import pytest

import opendynamo_amd as od
from helpers import rel_err_per_var
from opendynamo_amd import ic
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TWO_PI = 2 * np.pi
NAMES = ["rho", "rhou", "rhov", "rhow", "E"]


def announce(**case):
    print("CASE", " ".join(f"{k}={v}" for k, v in case.items()), flush=True)


def l2_rel(a, b):
    a = np.asarray(a, np.float64).reshape(5, -1)
    b = np.asarray(b, np.float64).reshape(5, -1)
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)


def test_precision_study_mixed_pipeline_bitwise():
    n = 8
    base = dict(flux="rusanov", y_wall=False, viscous=False, body_force_x=0.0)
    mixed = dict(u_real=8, compute_real=4, q_real=4, rhs_real=4, rk_accum=1)
    u0 = ic.fourier_state(n, n, n, TWO_PI, TWO_PI, TWO_PI, u0=(0.1, 0.0, 0.0), amplitude=1e-2, n_modes=8, seed=0)
    p64 = O.params(n, n, n, TWO_PI, TWO_PI, TWO_PI, u_real=8, **base)
    dt = O.cfl_dt(p64, u0, 0.3)
    po_mixed = O.params(n, n, n, TWO_PI, TWO_PI, TWO_PI, **mixed, **base)
    pg_mixed = od.params(n, n, n, TWO_PI, TWO_PI, TWO_PI, **mixed, **base)
    announce(test="precision_study", mesh=n, dt=dt)
    ref64 = {10: O.step(p64, u0, dt, 10)}
    ref64[100] = O.step(p64, ref64[10], dt, 90)
    refmx = {10: O.step(po_mixed, u0, dt, 10)}
    refmx[100] = O.step(po_mixed, refmx[10], dt, 90)
    with od.Solver(pg_mixed) as s:
        s.set_state(u0)
        s.step(dt, 10)
        g10 = s.get_state()
        s.step(dt, 90)
        g100 = s.get_state()
    for k, g in ((10, g10), (100, g100)):
        np.testing.assert_array_equal(g, refmx[k])  # the same pipeline on the CPU: bit for bit
        mx, l2 = rel_err_per_var(g, ref64[k]), l2_rel(g, ref64[k])
        print(f"after {k} steps, mixed vs FP64 oracle: " +
              ", ".join(f"{nm} max {a:.2e} L2 {b:.2e}" for nm, a, b in zip(NAMES, mx, l2)), flush=True)
        # f32 fluxes into an f64 accumulation: errors of f32 rounding size,
        # far below an all-f32 run (~1e-5 after 100 steps)
        assert np.all(mx < 5e-5) and np.all(l2 < 5e-6)
    # the all-f64 GPU pipeline on the same workload is bitwise the FP64 oracle
    with od.Solver(od.params(n, n, n, TWO_PI, TWO_PI, TWO_PI, u_real=8, **base)) as s:
        s.set_state(u0)
        s.step(dt, 100)
        np.testing.assert_array_equal(s.get_state(), ref64[100])


@pytest.fixture(scope="module")
def ensemble():
    boxes = ic.ensemble_boxes()
    kw = dict(u_real=od.OD_F32, flux="slau2", y_wall=False, viscous=True, viscosity_model=0, mu_ref=0.00016,
              body_force_x=0.0, fast_math=1)
    return boxes, kw


def test_ensemble_24_boxes_match_single_runs(ensemble):
    boxes, kw = ensemble
    dims = [(b["nx"], b["ny"], b["nz"], b["lx"], b["ly"], b["lz"]) for b in boxes]
    n_steps = 3
    announce(test="ensemble24", boxes=[b["nx"] for b in boxes], steps=n_steps)
    with od.Batch(od.params(1, 1, 1, 1.0, 1.0, 1.0, **kw), dims) as B:
        u0 = [ic.ensemble_state(b, np.float32) for b in boxes]
        for q in range(len(boxes)):
            B.set_state(q, u0[q])
        dt = B.cfl_dt(0.3)
        best, best_ms, table = B.autotune(dt, steps_per_trial=2)
        print(f"tuned schedule {best} at {best_ms:.3f} ms/step over {len(table)} trials", flush=True)
        for q in range(len(boxes)):  # tuning leaves the states as they were
            assert B.get_state(q).tobytes() == u0[q].tobytes()
        B.step(dt, n_steps)
        got = [B.get_state(q) for q in range(len(boxes))]
    for q, b in enumerate(boxes):
        p = od.params(b["nx"], b["ny"], b["nz"], b["lx"], b["ly"], b["lz"], **kw)
        with od.Solver(p) as s:
            s.set_state(u0[q])
            s.step(dt[q], n_steps)
            assert s.get_state().tobytes() == got[q].tobytes(), f"box {q} ({b['nx']}^3) differs from its single run"
    # the smallest box against the FP64 oracle (fast math: not bitwise)
    q = int(np.argmin([b["nx"] for b in boxes]))
    b = boxes[q]
    okw = {k: v for k, v in kw.items() if k != "fast_math"}
    okw["u_real"] = 8
    po = O.params(b["nx"], b["ny"], b["nz"], b["lx"], b["ly"], b["lz"], **okw)
    ref = O.step(po, u0[q].astype(np.float64), float(np.float32(dt[q])), n_steps)
    err = rel_err_per_var(got[q], ref)
    print(f"box {q} ({b['nx']}^3) f32 fast vs FP64 oracle after {n_steps} steps: "
          + ", ".join(f"{nm} {e:.2e}" for nm, e in zip(NAMES, err)), flush=True)
    assert np.all(err < 2e-5)
