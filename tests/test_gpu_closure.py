"""Learned closure on the GPU against the CPU oracle: the eddy-viscosity
coupling (bitwise, strict builds), the feature kernel, the BF16 MFMA MLP
forward and backward (tolerances stated per test), training, and both
fallbacks (training range, kinetic-energy budget) reducing to the plain FR
step exactly."""
import numpy as np
import pytest

import opendynamo_amd as od
from opendynamo_amd import ic
from oracle import mlp as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TWO_PI = 2 * np.pi


def announce(**case):
    print("CASE", " ".join(f"{k}={v}" for k, v in case.items()), flush=True)


def box_state(n=4, amp=0.05, seed=3, dtype=np.float64):
    return ic.fourier_state(n, n, n, TWO_PI, TWO_PI, TWO_PI, u0=(0.05, 0.0, 0.0), amplitude=amp, n_modes=8,
                            seed=seed, kmax=3, dtype=dtype)


def ns_kw(real, fast=0):
    return dict(u_real=real, flux="slau2", y_wall=False, viscous=True, viscosity_model=0, mu_ref=1e-3,
                body_force_x=0.0, fast_math=fast)


@pytest.mark.parametrize("real", [od.OD_F64, od.OD_F32])
def test_eddy_viscosity_step_matches_oracle(real):
    n = 4
    announce(test="mu_t_step", real=real)
    kw = ns_kw(real)
    rng = np.random.default_rng(1)
    mu_t = rng.uniform(0.0, 3e-3, n ** 3)
    u0 = box_state(n, dtype=np.float64 if real == 8 else np.float32)
    p = od.params(n, n, n, TWO_PI, TWO_PI, TWO_PI, **kw)
    with od.Solver(p) as s:
        s.set_state(u0)
        s.set_eddy_viscosity(mu_t)
        s.step(0.01, 3)
        got = s.get_state()
    okw = dict(kw)
    okw.pop("fast_math")
    po = O.params(n, n, n, TWO_PI, TWO_PI, TWO_PI, **okw)
    want = O.step_mut(po, u0, mu_t, 0.01, 3)
    assert got.tobytes() == want.tobytes()
    # a uniform mu_t is a shift of the molecular viscosity (f64: exact)
    if real == od.OD_F64:
        po2 = O.params(n, n, n, TWO_PI, TWO_PI, TWO_PI, **dict(okw, mu_ref=1e-3 + 2e-3))
        a = O.step_mut(po, u0, np.full(n ** 3, 2e-3), 0.01, 2)
        b = O.step(po2, u0, 0.01, 2)
        assert a.tobytes() == b.tobytes()


@pytest.mark.parametrize("n", [4, 3], ids=["blocks_4cubed", "tiles_3cubed"])
def test_fast_eddy_viscosity_step_against_fp64_oracle(n):
    """The fast f32 build with a per-element eddy viscosity (the closure's
    coupling) and a body force: 2x2x2-element block kernels at 4^3, two-element
    tiles at 3^3, both within the fast build's tolerance of the FP64 oracle
    (2e-5 of each variable's largest value, as for the other fast-math cases)."""
    announce(test="mu_t_step_fast", n=n)
    kw = dict(ns_kw(od.OD_F32, fast=1), viscosity_model=1, body_force_x=1e-3)
    rng = np.random.default_rng(7)
    mu_t = rng.uniform(0.0, 3e-3, n ** 3)
    u0 = box_state(n, dtype=np.float32)
    p = od.params(n, n, n, TWO_PI, TWO_PI, TWO_PI, **kw)
    with od.Solver(p) as s:
        s.set_state(u0)
        s.set_eddy_viscosity(mu_t)
        s.step(0.01, 3)
        got = s.get_state().astype(np.float64).reshape(5, -1)
    okw = dict(kw, u_real=od.OD_F64)
    okw.pop("fast_math")
    po = O.params(n, n, n, TWO_PI, TWO_PI, TWO_PI, **okw)
    want = O.step_mut(po, u0.astype(np.float64), mu_t, float(np.float32(0.01)), 3).reshape(5, -1)
    err = np.abs(got - want).max(axis=1) / np.abs(want).max(axis=1)
    print("fast f32 + mu_t vs FP64 oracle, rel err per var", err, flush=True)
    assert np.all(err <= 2e-5), err
    # the eddy viscosity acts: the same step without it differs
    with od.Solver(p) as s:
        s.set_state(u0)
        s.step(0.01, 3)
        plain = s.get_state().astype(np.float64).reshape(5, -1)
    assert np.abs(plain - want).max() > 10 * np.abs(got - want).max()


def test_zero_eddy_viscosity_is_plain_fr():
    n = 4
    p = od.params(n, n, n, TWO_PI, TWO_PI, TWO_PI, **ns_kw(od.OD_F32, fast=1))
    u0 = box_state(n, dtype=np.float32)
    with od.Solver(p) as s:
        s.set_state(u0)
        s.step(0.01, 3)
        plain = s.get_state()
        s.set_state(u0)
        s.set_eddy_viscosity(np.zeros(n ** 3))
        s.step(0.01, 3)
        assert s.get_state().tobytes() == plain.tobytes()
        s.set_state(u0)
        s.set_eddy_viscosity(None)
        s.step(0.01, 3)
        assert s.get_state().tobytes() == plain.tobytes()


@pytest.mark.parametrize("real", [od.OD_F32, od.OD_F64])
def test_features_match_oracle(real):
    n = 4
    announce(test="features", real=real)
    u0 = box_state(n, amp=0.08, dtype=np.float64)
    # density and pressure variations as well (the Fourier field has uniform rho and p)
    X, Y, Z = ic.point_coords(n, n, n, TWO_PI, TWO_PI, TWO_PI)
    f = 1.0 + 0.02 * np.sin(X + 2 * Z) * np.cos(Y)
    u0 = u0.reshape(5, -1) * f
    u0 = u0.ravel().astype(np.float64 if real == 8 else np.float32)
    p = od.params(n, n, n, TWO_PI, TWO_PI, TWO_PI, **ns_kw(real))
    with od.Solver(p) as s:
        s.set_state(u0)
        got = s.features()
    want = M.features(u0.astype(np.float64), n, n, n, float(np.float32(TWO_PI)), float(np.float32(TWO_PI)),
                      float(np.float32(TWO_PI)), mu_mol=1e-3)
    # f32 arithmetic on the GPU: 1e-3 of each feature's scale (features are
    # O(1e-3..10); scales below 1e-4 count as 1e-4)
    assert np.all(np.abs(want[:, 8:10]) > 1e-4)
    scale = np.maximum(np.abs(want).max(0), 1e-4)
    err = np.abs(got - want) / scale
    assert err.max() < 1e-3, err.max(0)


def mlp_case(seed=0, n=1000):
    rng = np.random.default_rng(seed)
    params = rng.normal(0.0, 0.25, M.NPARAM).astype(np.float32)
    x = rng.normal(0.0, 1.0, (n, M.NF)).astype(np.float32)
    shift = rng.normal(0.0, 0.1, M.NF).astype(np.float32)
    scale = rng.uniform(0.8, 1.2, M.NF).astype(np.float32)
    t = rng.uniform(0.0, 1.0, n).astype(np.float32)
    return params, x, shift, scale, t


def test_mlp_forward_matches_oracle():
    params, x, shift, scale, _ = mlp_case(0, 1000)  # 1000 = 62 tiles of 16 + 8
    announce(test="mlp_forward")
    with od.MLP(params=params) as m:
        m.set_scaling(shift, scale)
        got = m.forward(x)
        assert m.get_params().tobytes() == params.tobytes()
    want, _ = M.forward(params, x, shift, scale, rounding=True)
    # BF16 activations: an FP32-vs-exact sum difference can flip one BF16
    # rounding; bound: 2 % of the output scale, median 1e-3
    err = np.abs(got - want) / np.abs(want).max()
    assert err.max() < 2e-2 and np.median(err) < 1e-3, (err.max(), np.median(err))
    exact, _ = M.forward(params, x, shift, scale, rounding=False)
    assert np.abs(got - exact).max() / np.abs(exact).max() < 5e-2


@pytest.mark.parametrize("out_scale", [1.0, 0.01])
def test_mlp_gradient_matches_oracle(out_scale):
    params, x, shift, scale, t = mlp_case(1, 777)
    t = (t * out_scale).astype(np.float32)
    announce(test="mlp_grad", out_scale=out_scale)
    with od.MLP(params=params) as m:
        m.set_scaling(shift, scale)
        m.set_output_scale(out_scale)
        loss, g = m.loss_grad(x, t)
        y = m.forward(x)
    wl, wg = M.loss_grad(params, x, t, shift, scale, rounding=True, out_scale=out_scale)
    wy, _ = M.forward(params, x, shift, scale, rounding=True, out_scale=out_scale)
    assert np.abs(y - wy).max() <= 2e-2 * np.abs(wy).max()
    assert abs(loss - wl) / wl < 1e-2
    rel = np.linalg.norm(g - wg) / np.linalg.norm(wg)
    cos = float(np.dot(g, wg) / (np.linalg.norm(g) * np.linalg.norm(wg)))
    assert rel < 3e-2 and cos > 0.999, (rel, cos)
    # every block of the parameter vector, separately
    for name, (o, sz) in M.OFF.items():
        a, b = g[o:o + sz], wg[o:o + sz]
        assert np.linalg.norm(a - b) <= 5e-2 * np.linalg.norm(b) + 1e-7, name


def test_mlp_training_reduces_loss_and_is_deterministic():
    rng = np.random.default_rng(5)
    n = 20000
    x = rng.normal(0.0, 1.0, (n, M.NF)).astype(np.float32)
    t = (np.log1p(np.exp(0.8 * x[:, 0] - 0.5 * x[:, 3] + 0.3 * x[:, 7] * x[:, 2])) * 0.5).astype(np.float32)
    announce(test="mlp_train")
    with od.MLP() as a, od.MLP() as b:
        p0 = a.get_params()
        l0, _ = a.loss_grad(x, t)
        ha = a.train(x, t, epochs=40, batch=4096, lr=3e-3, seed=11)
        hb = b.train(x, t, epochs=40, batch=4096, lr=3e-3, seed=11)
        la, _ = a.loss_grad(x, t)
        assert a.get_params().tobytes() == b.get_params().tobytes()
        assert ha.tobytes() == hb.tobytes()
        assert not np.array_equal(a.get_params(), p0)
    assert la < 0.1 * l0, (l0, la, ha[-5:])


def closure_setup(n=4, fast=1):
    p = od.params(n, n, n, TWO_PI, TWO_PI, TWO_PI, **ns_kw(od.OD_F32, fast=fast))
    u0 = box_state(n, amp=0.08, dtype=np.float32)
    params, _, _, _, _ = mlp_case(2, 16)
    params = (params * 0.3).astype(np.float32)
    return p, u0, params


def test_closure_mu_t_matches_oracle_and_range_fallback():
    p, u0, params = closure_setup()
    announce(test="closure_mu_t")
    with od.Solver(p) as s, od.MLP(params=params) as m:
        s.set_state(u0)
        feat = s.features()
        lo, hi = feat[:, :12].min(0), feat[:, :12].max(0)
        # a range that excludes the two elements with the largest feature 1
        cut = np.sort(feat[:, 1])[-3]
        hi2 = hi.copy()
        hi2[1] = cut
        m.set_scaling(np.zeros(12), np.ones(12), lo, hi2)
        s.attach_closure(m, update_interval=2, budget_kappa=0.0)
        s.step(0.01, 1)
        mu = s.get_eddy_viscosity()
        st = s.closure_stats()
    want, inr = M.mu_t_from_features(params, feat.astype(np.float64), np.zeros(12), np.ones(12), lo, hi2)
    assert st["updates"] == 1 and st["out_of_range"] == int((~inr).sum()) == 2
    assert np.all(mu[~inr] == 0.0) and np.all(mu[inr] > 0.0)
    assert np.abs(mu - want).max() <= 2e-2 * np.abs(want).max()


def test_closure_all_out_of_range_is_plain_fr():
    p, u0, params = closure_setup()
    with od.Solver(p) as s, od.MLP(params=params) as m:
        s.set_state(u0)
        s.step(0.01, 4)
        plain = s.get_state()
        m.set_scaling(np.zeros(12), np.ones(12), np.full(12, 1e30), np.full(12, 2e30))
        s.set_state(u0)
        s.attach_closure(m, update_interval=2, budget_kappa=1.0)
        s.step(0.01, 4)
        assert s.get_state().tobytes() == plain.tobytes()
        assert s.closure_stats()["out_of_range"] == 2 * 64


def test_closure_budget_fallback_reverts_to_plain_fr():
    """kappa tiny: the second update sees the model's dissipated energy exceed
    kappa x the observed decay (the viscous box decays) and switches the model
    off for the interval; the run then equals a run with the first mu_t for
    one interval and the plain step for the next, byte for byte."""
    p, u0, params = closure_setup()
    announce(test="closure_budget")
    with od.Solver(p) as s, od.MLP(params=params) as m, od.Solver(p) as r:
        m.set_scaling(np.zeros(12), np.ones(12))
        s.set_state(u0)
        s.attach_closure(m, update_interval=3, budget_kappa=1e-12)
        s.step(0.01, 3)
        mu1 = s.get_eddy_viscosity()
        ke1 = s.closure_stats()["kinetic_energy"]
        assert s.closure_stats()["active"] and mu1.max() > 0
        s.step(0.01, 3)
        st = s.closure_stats()
        assert st["kinetic_energy"] < ke1  # the interval decayed
        assert st["updates"] == 2 and st["budget_fallbacks"] == 1 and not st["active"]
        assert np.all(s.get_eddy_viscosity() == 0.0)
        got = s.get_state()
        r.set_state(u0)
        r.set_eddy_viscosity(mu1)
        r.step(0.01, 3)
        r.set_eddy_viscosity(None)
        r.step(0.01, 3)
        assert got.tobytes() == r.get_state().tobytes()


def test_committed_closure_model_runs_on_the_coarse_mesh():
    """The trained network shipped with the package (48^3 -> 16^3, Re 400),
    at its calibrated output scale, loads, attaches and produces a positive
    eddy viscosity on its own coarse workload; the first 50 steps stay
    finite."""
    import os
    from opendynamo_amd import closure as C
    w = np.load(os.path.join(os.path.dirname(od.__file__), "closure_model", "closure_weights.npz"))
    cfg = C.DEFAULT
    n = cfg["fine"] // cfg["ratio"]
    announce(test="committed_model", n=n)
    p = C.case_params(n, cfg["L"], C.viscosity(cfg))
    u0 = C.initial_state(n, cfg["L"], 100, cfg["n_modes"], cfg["amplitude"], cfg["kmax"])
    with od.MLP(params=w["params"]) as m, od.Solver(p) as s:
        m.set_scaling(w["shift"], w["scale"], w["lo"], w["hi"])
        m.set_output_scale(float(w["out_scale"]) * float(w["output_multiplier"]))
        s.set_state(u0)
        dt = s.cfl_dt(0.3)
        s.attach_closure(m, update_interval=5, budget_kappa=1.0)
        s.step(dt, 50)
        mu = s.get_eddy_viscosity()
        st = s.closure_stats()
        assert np.isfinite(s.get_state()).all()
    # elements outside the training range (snapshots t >= 5; this is t < 0.3) get mu_t = 0
    assert st["updates"] == 10 and np.all(mu >= 0.0) and (mu > 0).mean() > 0.1
    assert st["out_of_range"] > 0
    # the eddy viscosity is a fraction of the molecular one at this Reynolds number
    assert 0.0 < np.median(mu[mu > 0]) < C.viscosity(cfg)


def test_energy_spectrum_gpu_fft_matches_numpy():
    """The spectra's GPU FFT path (torch.fft, hipFFT, f32) agrees with numpy's
    f64 FFT shell by shell on a small multi-mode field.  It runs in a child
    process: torch ships its own HIP runtime, which finds no GPU in a process
    where the library's runtime has already opened it (the closure pipeline's
    processes, which therefore take the numpy path)."""
    import json
    import os
    import subprocess
    import sys
    from opendynamo_amd import closure as C
    announce(test="energy_spectrum", n=6)
    n, L = 6, TWO_PI
    u = box_state(n=n, amp=0.1, seed=11, dtype=np.float64)
    vel = C.uniform_velocity(u, n)  # 24^3 uniform grid
    E_np = C.energy_spectrum(vel, L, backend="numpy")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = ("import json, sys, numpy as np; sys.path.insert(0, %r); from opendynamo_amd import closure as C; "
            "v = np.frombuffer(sys.stdin.buffer.read(), dtype=np.float64).reshape(3, 24, 24, 24); "
            "print(json.dumps(C.energy_spectrum(v, %r, backend='gpu').tolist()))") % (root, L)
    r = subprocess.run([sys.executable, "-c", code], input=np.ascontiguousarray(vel).tobytes(), capture_output=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr.decode(errors="replace")[-2000:]
    E_gpu = np.array(json.loads(r.stdout.decode().strip().splitlines()[-1]))
    assert E_gpu.shape == E_np.shape == (12,)
    # f32 FFT vs f64: a shell's relative error grows as sqrt(E_max / E_k) x f32 rounding,
    # so shells above 1e-4 E_max are held to 1e-4 relative and every shell to 1e-6 E_max
    big = E_np > 1e-4 * E_np.max()
    assert big.sum() >= 5
    assert np.all(np.abs(E_gpu[big] - E_np[big]) <= 1e-4 * E_np[big])
    assert np.max(np.abs(E_gpu - E_np)) <= 1e-6 * E_np.max()
    # Parseval: the shells k = 1 .. N/2 hold at most 1/2 <|u|^2>
    half_mean_sq = 0.5 * np.mean(np.sum(vel ** 2, axis=0))
    assert E_np.sum() <= half_mean_sq * (1 + 1e-12)


def test_closure_gains_over_the_plain_coarse_run():
    """What the trained network gains over the plain FR step on its coarse
    mesh (16^3, Re 400), on held-out seed 101, measured against the filtered
    48^3 run: at t = 5 the closure run's energy spectrum is closer to the
    fine run's than the plain 16^3 run's, and at t = 10 the plain run has
    diverged while the closure run is still finite."""
    import os
    from opendynamo_amd import closure as C
    w = np.load(os.path.join(os.path.dirname(od.__file__), "closure_model", "closure_weights.npz"))
    cfg = dict(C.DEFAULT)
    n = cfg["fine"] // cfg["ratio"]
    k_max = (4 * n) // 3
    times = [5.0, 10.0]
    announce(test="closure_gain", seed=101, n=n, fine=cfg["fine"], times=times)
    E_refs, _ = C.fine_spectra(cfg, 101, times)
    with od.MLP(params=w["params"]) as m:
        m.set_scaling(w["shift"], w["scale"], w["lo"], w["hi"])
        m.set_output_scale(float(w["out_scale"]) * float(w["output_multiplier"]))
        clos = C.coarse_errors(cfg, n, 101, times, E_refs, k_max, m)
    plain = C.coarse_errors(cfg, n, 101, times, E_refs, k_max, None)
    print("closure", clos["errs"], "plain", plain["errs"], flush=True)
    assert clos["errs"][0] is not None and plain["errs"][0] is not None
    assert clos["errs"][0] < 0.9 * plain["errs"][0]
    assert plain["errs"][1] is None and clos["errs"][1] is not None


def test_constant_coefficient_baseline_is_constant():
    """The constant-coefficient baseline (closure.constant_model) gives the
    same C = s for every element, whatever its features."""
    from opendynamo_amd import closure as C
    announce(test="constant_model")
    rng = np.random.default_rng(5)
    x = rng.normal(size=(4096, 12)).astype(np.float32)
    shift, scale = np.zeros(12, np.float32), np.ones(12, np.float32)
    lo, hi = np.full(12, -1e9, np.float32), np.full(12, 1e9, np.float32)
    s = 0.0123
    with C.constant_model(s, shift, scale, lo, hi) as m:
        y = m.forward(x)
    assert np.all(y == y[0])
    assert abs(y[0] - s) <= 5e-3 * s  # softplus(b4) = 1 up to the BF16 rounding of b4


def test_closure_beats_the_constant_coefficient_baseline():
    """What the network's element-local coefficient adds over a constant one:
    the constant-coefficient eddy viscosity in the same form, calibrated by
    the same multiplier search (the smallest multiplier that keeps the
    calibration runs finite, x11), against the network at its calibrated x8
    on held-out seed 101 (16^3, Re 400): the network's spectrum is closer to
    the filtered 48^3 run's at t = 5 (by more than a factor 2) and at t = 10,
    and it is still finite at t = 15, where the constant-coefficient run has
    diverged (the weaker constants that get close at t = 5 diverge earlier
    still: their calibration runs end before t = 10)."""
    import json
    import os
    from opendynamo_amd import closure as C
    mdir = os.path.join(os.path.dirname(od.__file__), "closure_model")
    w = np.load(os.path.join(mdir, "closure_weights.npz"))
    report = json.load(open(os.path.join(mdir, "closure_report.json")))
    cmult = float(report["constant_baseline"]["multiplier"])
    cfg = dict(C.DEFAULT)
    n = cfg["fine"] // cfg["ratio"]
    k_max = (4 * n) // 3
    times = [5.0, 10.0, 15.0]
    announce(test="closure_vs_constant", seed=101, n=n, constant_multiplier=cmult)
    E_refs, _ = C.fine_spectra(cfg, 101, times)
    base = float(w["out_scale"])
    with od.MLP(params=w["params"]) as m:
        m.set_scaling(w["shift"], w["scale"], w["lo"], w["hi"])
        m.set_output_scale(base * float(w["output_multiplier"]))
        net = C.coarse_errors(cfg, n, 101, times, E_refs, k_max, m)
    with C.constant_model(base * cmult, w["shift"], w["scale"], w["lo"], w["hi"]) as m:
        const = C.coarse_errors(cfg, n, 101, times, E_refs, k_max, m)
    print("network", net["errs"], "constant", const["errs"], flush=True)
    assert all(e is not None for e in net["errs"])
    assert const["errs"][0] is not None and const["errs"][1] is not None
    assert net["errs"][0] < 0.5 * const["errs"][0]
    assert net["errs"][1] < const["errs"][1]
    assert const["errs"][2] is None
