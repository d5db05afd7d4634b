"""Learned closure: training data from filtered fine runs, GPU training, and
the spectral evaluation against the fine run (no reference counterpart).

Pipeline (every step on the GPU through the library except the filter and
the spectra, which are numpy post-processing):

  1. fine runs: plain FR P=3 NS on n_f^3 periodic boxes from a divergence-
     free Fourier field, snapshots at fixed times;
  2. filter: L2 projection of each fine snapshot onto the P=3 space of the
     coarse mesh (n_c = n_f / r elements per side, r integer), applied to
     rho, rho u_i and rho u_i u_j (Favre filtering);
  3. target: per coarse element, the least-squares eddy viscosity of the
     deviatoric subgrid stress tau_ij = bar(rho u_i u_j) - bar(rho u_i)
     bar(rho u_j) / bar(rho) against -2 mu_t S~^d_ij, clipped at 0, divided by
     rho_e Delta^2 |S|_e of the filtered state -> coefficient C >= 0;
  4. features: the library's feature kernel on the filtered state (the same
     kernel the closure evaluates at run time);
  5. training: od_mlp_train (BF16 MFMA forward/backward, Adam) on the
     standardised features, with the training range kept for the fallback;
  6. a-posteriori calibration: one multiplier of the network's output
     scale, chosen on calibration seeds (from the training seeds) as the
     one with the smallest spectral error, averaged over every check time
     (or, with calib_objective = "eval_time", at the evaluation time only),
     among those whose coarse runs all stay finite to the run time;
  7. evaluation: held-out seeds, energy spectra of coarse+closure, plain
     coarse and plain intermediate runs against the fine run, at every
     check time the runs stay finite.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

import opendynamo_amd as od
from opendynamo_amd import ic

TWO_PI = 2.0 * math.pi


# ---------------------------------------------------------------------------
# FR P=3 point set
# ---------------------------------------------------------------------------
def _gl():
    x, w = np.polynomial.legendre.leggauss(4)
    return x, w


def _legendre(a, x):
    p0, p1 = np.ones_like(x), np.asarray(x, dtype=np.float64)
    if a == 0:
        return p0
    for k in range(2, a + 1):
        p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
    return p1


def projection_matrix(r):
    """[4][4r]: fine nodal values of the r fine elements along one axis of a
    coarse element -> coarse nodal values of their L2 projection onto P3."""
    x, w = _gl()
    xc = np.array([-1.0 + (2 * s + 1 + x[i]) / r for s in range(r) for i in range(4)])
    wc = np.array([w[i] / r for s in range(r) for i in range(4)])
    P = np.zeros((4, 4 * r))
    for a in range(4):
        coef = (2 * a + 1) / 2.0 * wc * _legendre(a, xc)  # c_a = sum coef * f
        P += np.outer(_legendre(a, x), coef)
    return P


def filter_field(f, nf, r):
    """Project a fine nodal field [E_f * 64] (element-major, pt = i + 4j + 16k)
    onto the coarse P3 space, n_c = nf / r; returns [E_c * 64]."""
    nc = nf // r
    P = projection_matrix(r)
    a = np.asarray(f, dtype=np.float64).reshape(nc, r, nc, r, nc, r, 4, 4, 4)  # cz sz cy sy cx sx k j i
    a = a.transpose(0, 2, 4, 1, 6, 3, 7, 5, 8).reshape(nc, nc, nc, 4 * r, 4 * r, 4 * r)
    a = np.einsum("kz,abczyx->abckyx", P, a)
    a = np.einsum("jy,abckyx->abckjx", P, a)
    a = np.einsum("ix,abckjx->abckji", P, a)
    return a.reshape(-1)


def filter_state(state, nf, r):
    """Filtered conservative state [5][E_c * 64] and the filtered products
    bar(rho u_i u_j) [6][E_c * 64] (ij = xx, yy, zz, xy, xz, yz)."""
    E = nf ** 3
    u = np.asarray(state, dtype=np.float64).reshape(5, E * 64)
    cons = np.stack([filter_field(u[v], nf, r) for v in range(5)])
    pairs = [(1, 1), (2, 2), (3, 3), (1, 2), (1, 3), (2, 3)]
    prod = np.stack([filter_field(u[a] * u[b] / u[0], nf, r) for a, b in pairs])
    return cons, prod


def _d_matrix():
    x, _ = _gl()
    D = np.zeros((4, 4))
    for i in range(4):
        for m in range(4):
            if i == m:
                D[i, m] = sum(1.0 / (x[i] - x[q]) for q in range(4) if q != i)
            else:
                num = np.prod([x[i] - x[q] for q in range(4) if q not in (i, m)])
                den = np.prod([x[m] - x[q] for q in range(4) if q != m])
                D[i, m] = num / den
    return D


def sgs_target(cons, prod, nc, L):
    """mu_t per coarse element: argmin_mu sum_q w_q |tau^d + 2 mu S~^d|^2,
    clipped at 0 (Favre-filtered velocity, density-weighted stress)."""
    _, w = _gl()
    D = _d_matrix()
    E = nc ** 3
    c = cons.reshape(5, E, 4, 4, 4)
    rho = c[0]
    ut = c[1:4] / rho  # Favre-filtered velocity
    pr = prod.reshape(6, E, 4, 4, 4)
    idx = {(0, 0): 0, (1, 1): 1, (2, 2): 2, (0, 1): 3, (0, 2): 4, (1, 2): 5}
    tau = np.empty((3, 3, E, 4, 4, 4))
    for a in range(3):
        for b in range(3):
            q = idx[(min(a, b), max(a, b))]
            tau[a, b] = pr[q] - c[1 + a] * c[1 + b] / rho
    tr = (tau[0, 0] + tau[1, 1] + tau[2, 2]) / 3.0
    for a in range(3):
        tau[a, a] -= tr
    jac = 2.0 / (L / nc)
    A = np.empty((3, 3) + ut.shape[1:])
    for a in range(3):
        A[a, 0] = np.einsum("im,ekjm->ekji", D, ut[a]) * jac
        A[a, 1] = np.einsum("jm,ekmi->ekji", D, ut[a]) * jac
        A[a, 2] = np.einsum("km,emji->ekji", D, ut[a]) * jac
    div = A[0, 0] + A[1, 1] + A[2, 2]
    S = 0.5 * (A + A.transpose(1, 0, 2, 3, 4, 5))
    for a in range(3):
        S[a, a] -= div / 3.0
    W = np.einsum("k,j,i->kji", w, w, w)
    num = -np.einsum("abekji,abekji,kji->e", tau, S, W)
    den = 2.0 * np.einsum("abekji,abekji,kji->e", S, S, W)
    return np.maximum(num / np.maximum(den, 1e-300), 0.0)


# ---------------------------------------------------------------------------
# Runs
# ---------------------------------------------------------------------------
def case_params(n, L, mu, flux="slau2", fast_math=1):
    return od.params(n, n, n, L, L, L, u_real=od.OD_F32, flux=flux, y_wall=False, viscous=True,
                     viscosity_model=0, mu_ref=mu, body_force_x=0.0, fast_math=fast_math)


def initial_state(n, L, seed, n_modes, amplitude, kmax):
    return ic.fourier_state(n, n, n, L, L, L, u0=(0.0, 0.0, 0.0), amplitude=amplitude, n_modes=n_modes, seed=seed,
                            kmax=kmax, dtype=np.float32)


# progress lines of long runs (set by main): at most one per PROGRESS_S seconds
PROGRESS = {"log": None, "last": 0.0}
PROGRESS_S = 30.0


def _progress(msg):
    if PROGRESS["log"] is not None and time.perf_counter() - PROGRESS["last"] >= PROGRESS_S:
        PROGRESS["log"](msg)
        PROGRESS["last"] = time.perf_counter()


def run(n, L, mu, seed, cfg, times, mlp=None, update_interval=5, budget_kappa=1.0, cfl=0.3):
    """Advance from the IC, returning the state at each of `times` (ascending)
    and the GPU wall time of the stepping.  dt from the CFL of the IC."""
    p = case_params(n, L, mu, cfg.get("flux", "slau2"))
    u0 = initial_state(n, L, seed, cfg["n_modes"], cfg["amplitude"], cfg["kmax"])
    out, wall = [], 0.0
    with od.Solver(p) as s:
        s.set_state(u0)
        dt = s.cfl_dt(cfl)
        if mlp is not None:
            s.attach_closure(mlp, update_interval=update_interval, budget_kappa=budget_kappa)
        t, step = 0.0, 0
        for T in times:
            n_steps = int(round((T - t) / dt))
            s.synchronize()
            t0 = time.perf_counter()
            s.step(dt, n_steps)
            s.synchronize()
            wall += time.perf_counter() - t0
            t += n_steps * dt
            step += n_steps
            out.append(s.get_state())
            _progress(f"  {n}^3 seed {seed}: t = {t:.2f} ({step} steps, {wall:.1f} s)")
        stats = s.closure_stats() if mlp is not None else None
    return dict(states=out, wall=wall, dt=dt, steps=step, stats=stats)


# ---------------------------------------------------------------------------
# Spectra
# ---------------------------------------------------------------------------
def uniform_velocity(state, n, m=4):
    """Velocity on the uniform (m n)^3 grid of cell centres (m per element
    per axis), by Lagrange interpolation inside each element."""
    x, _ = _gl()
    xu = -1.0 + (2.0 * np.arange(m) + 1.0) / m
    I = np.ones((m, 4))
    for a in range(m):
        for b in range(4):
            for q in range(4):
                if q != b:
                    I[a, b] *= (xu[a] - x[q]) / (x[b] - x[q])
    u = np.asarray(state, dtype=np.float64).reshape(5, n, n, n, 4, 4, 4)  # ez ey ex k j i
    vel = u[1:4] / u[0]
    g = np.einsum("ak,bj,ci,vzyxkji->vzaybxc", I, I, I, vel)
    return g.reshape(3, n * m, n * m, n * m)  # [v][z][y][x]


def energy_spectrum(vel, L, backend="numpy"):
    """Shell-summed E(k), k = 1 .. N/2 in units of 2 pi / L (integer shells).
    backend "numpy": numpy's f64 FFT on the host (what the pipeline uses: its
    grids are at most 192^3); "gpu": torch.fft (hipFFT) in f32, usable only
    in a process where torch opens the GPU before this library does (torch
    carries its own HIP runtime)."""
    N = vel.shape[1]
    if backend == "gpu":
        import torch
        t = torch.from_numpy(np.ascontiguousarray(vel, dtype=np.float32)).cuda()
        uh = torch.fft.rfftn(t, dim=(1, 2, 3)) / (N ** 3)
        e = (uh.abs() ** 2).sum(0).cpu().numpy().astype(np.float64)
    else:
        uh = np.fft.rfftn(vel, axes=(1, 2, 3)) / (N ** 3)
        e = (np.abs(uh) ** 2).sum(0)
    k1 = np.fft.fftfreq(N, 1.0 / N)
    kz, ky = np.meshgrid(k1, k1, indexing="ij")
    kx = np.arange(N // 2 + 1)
    kk = np.sqrt(kz[:, :, None] ** 2 + ky[:, :, None] ** 2 + kx[None, None, :] ** 2)
    # the rfft half-space counts the kx > 0 modes once: double them (except kx = 0 and Nyquist)
    wgt = np.full(kx.shape, 2.0)
    wgt[0] = 1.0
    if N % 2 == 0:
        wgt[-1] = 1.0
    shell = np.rint(kk).astype(int)
    E = np.bincount(shell.ravel(), weights=(0.5 * e * wgt[None, None, :]).ravel(), minlength=N)
    return E[1:N // 2 + 1]


def spectral_error(E, E_ref, k_max):
    k = np.arange(1, k_max + 1)
    return float(np.mean(np.abs(E[k - 1] - E_ref[k - 1]) / E_ref[k - 1]))


# ---------------------------------------------------------------------------
# Training set, training, evaluation
# ---------------------------------------------------------------------------
# Re 400: the plain FR step (no de-aliasing, like the reference) stays finite
# to t = 40 at 32^3 and 48^3 for this IC, while Re 1600 diverges at every
# resolution up to 48^3 (t ~ 5..23); plain 16^3 diverges at Re 400 too.
DEFAULT = dict(L=TWO_PI, mach=0.1, re=400.0, n_modes=64, amplitude=0.1, kmax=4, flux="slau2", fine=48, ratio=3,
               mid=32, snap_times=[5.0, 10.0, 15.0, 20.0, 25.0, 30.0, 35.0, 40.0], eval_time=5.0, run_time=40.0,
               train_seeds=list(range(8)), eval_seeds=[100, 101], epochs=300, batch=65536, lr=2e-3,
               update_interval=5, budget_kappa=1.0, calib_seeds=[0, 1], multipliers=[1, 2, 4, 8, 16, 32],
               constant_multipliers=[1, 2, 4, 6, 8, 11, 16, 23, 32, 45, 64], calib_objective="all_times")


def viscosity(cfg):
    """Re = u_rms L_ref / nu with L_ref = 1, u_rms = Mach (c = 1), rho = 1."""
    return cfg["mach"] * 1.0 / cfg["re"]


def build_dataset(cfg, log=print):
    mu = viscosity(cfg)
    nf, r = cfg["fine"], cfg["ratio"]
    nc = nf // r
    X, T = [], []
    pc = case_params(nc, cfg["L"], mu, cfg["flux"])
    with od.Solver(pc) as coarse:
        for seed in cfg["train_seeds"]:
            res = run(nf, cfg["L"], mu, seed, cfg, cfg["snap_times"])
            finite = [bool(np.isfinite(st).all()) for st in res["states"]]
            log(f"fine seed {seed}: {res['steps']} steps, {res['wall']:.1f} s, finite snapshots {finite}")
            if tuple(cfg["snap_times"]) == tuple(check_times(cfg)):  # the same run serves the spectra
                _FINE_CACHE[_fine_key(cfg, seed, cfg["snap_times"])] = (
                    [energy_spectrum(uniform_velocity(st, nf), cfg["L"]) for st in res["states"]],
                    dict(wall=res["wall"], steps=res["steps"]))
            for st, ok in zip(res["states"], finite):
                if not ok:
                    continue
                cons, prod = filter_state(st, nf, r)
                coarse.set_state(cons.reshape(-1).astype(np.float32))
                f = coarse.features()
                mu_t = sgs_target(cons, prod, nc, cfg["L"])
                C = mu_t / (f[:, 12].astype(np.float64) * f[:, 13].astype(np.float64))
                X.append(f[:, :12])
                T.append(C.astype(np.float32))
    return np.concatenate(X), np.concatenate(T)


def train(cfg, X, T, device=0, log=print):
    """Train on all seeds but the last, validate on the last (a-priori R^2)."""
    per_seed = len(T) // max(1, len(cfg["train_seeds"]))
    val = np.zeros(len(T), dtype=bool)
    val[len(T) - per_seed:] = True
    good = np.all(np.isfinite(X), axis=1) & np.isfinite(T)
    if not good.any():
        raise RuntimeError("no finite training samples (did the fine runs stay finite?)")
    log(f"{int(good.sum())} of {len(T)} samples finite")
    cap = np.quantile(T[good], 0.999)  # clip the extreme tail of the least-squares targets
    T = np.minimum(T, cap)
    tr, va = good & ~val, good & val
    Xt, Tt = X[tr], T[tr]
    shift = Xt.mean(0)
    scale = 1.0 / np.maximum(Xt.std(0), 1e-6)
    lo, hi = X[good].min(0), X[good].max(0)
    out_scale = float(Tt.mean())
    m = od.MLP(device)
    m.set_scaling(shift, scale, lo, hi)
    m.set_output_scale(out_scale)
    t0 = time.perf_counter()
    hist = m.train(Xt, Tt, epochs=cfg["epochs"], batch=cfg["batch"], lr=cfg["lr"], seed=1)
    secs = time.perf_counter() - t0
    r2 = lambda x, t: float(1.0 - np.mean((m.forward(x) - t) ** 2) / np.var(t))
    r2_train, r2_val = r2(Xt, Tt), r2(X[va], T[va])
    # a constant-coefficient model (the mean target) has R^2 = 0 by construction
    log(f"trained {cfg['epochs']} epochs on {len(Tt)} samples in {secs:.1f} s; loss {hist[0]:.3e} -> "
        f"{hist[-1]:.3e}; target variance {Tt.var():.3e}; R^2 train {r2_train:.3f}, held-out seed {r2_val:.3f}")
    return m, dict(shift=shift, scale=scale, lo=lo, hi=hi, out_scale=out_scale, loss_history=hist,
                   n_samples=int(len(Tt)), target_mean=out_scale, target_var=float(Tt.var()), train_s=secs,
                   r2_train=r2_train, r2_heldout=r2_val)


def constant_model(base_scale, shift, scale, lo, hi, device=0):
    """The constant-coefficient eddy viscosity in the closure's own form: a
    network with zero weights whose output bias makes softplus(.) = 1, so
    C = s for every element and mu_t = rho_e Delta^2 |S|_e s (a Smagorinsky
    type model).  It runs through the same kernels, training-range fallback,
    budget check and multiplier calibration as the trained network, which
    makes it the baseline that shows what the network's element-local
    variation of C adds."""
    params = np.zeros(od.MLP_PARAMS, np.float32)
    params[-1] = np.log(np.expm1(1.0))  # softplus(b4) = 1
    m = od.MLP(device, params=params)
    m.set_scaling(shift, scale, lo, hi)
    m.set_output_scale(base_scale)
    return m


def check_times(cfg):
    return [5.0 * q for q in range(1, int(cfg["run_time"] / 5.0) + 1)]


_FINE_CACHE = {}


def _fine_key(cfg, seed, times):
    return (cfg["fine"], cfg["L"], viscosity(cfg), seed, cfg["n_modes"], cfg["amplitude"], cfg["kmax"], cfg["flux"],
            tuple(float(t) for t in times))


def fine_spectra(cfg, seed, times):
    """Spectra of the fine run at `times` (kept for the process: calibration
    and evaluation of several models share the fine runs of a seed)."""
    nf = cfg["fine"]
    key = _fine_key(cfg, seed, times)
    if key not in _FINE_CACHE:
        fine = run(nf, cfg["L"], viscosity(cfg), seed, cfg, times)
        specs = [energy_spectrum(uniform_velocity(st, nf), cfg["L"]) for st in fine["states"]]
        _FINE_CACHE[key] = (specs, dict(wall=fine["wall"], steps=fine["steps"]))
    return _FINE_CACHE[key]


def coarse_errors(cfg, n, seed, times, E_refs, k_max, mlp=None):
    """Run n^3 from the seed's IC to each of `times`; spectral error at each
    time the state is finite (None after it diverges) and the GPU wall time."""
    res = run(n, cfg["L"], viscosity(cfg), seed, cfg, times, mlp=mlp, update_interval=cfg["update_interval"],
              budget_kappa=cfg["budget_kappa"])
    errs, specs = [], []
    for st, E_ref in zip(res["states"], E_refs):
        if np.isfinite(st).all():
            E = energy_spectrum(uniform_velocity(st, n), cfg["L"])
            errs.append(spectral_error(E, E_ref, k_max))
            specs.append(E[:k_max].tolist())
        else:
            errs.append(None)
            specs.append(None)
    fin = [e is not None for e in errs]
    finite_until = max([t for t, f in zip(times, fin) if f], default=0.0)
    return dict(errs=errs, spectra=specs, finite_until=finite_until, wall_s=res["wall"], steps=res["steps"],
                stats=res["stats"])


def calibrate(cfg, mlp, base_scale, log=print, multipliers=None):
    """A-posteriori multiplier of the output scale (see the module docstring).
    Returns the chosen multiplier and the trial table."""
    nc = cfg["fine"] // cfg["ratio"]
    k_max = (nc * 4) // 3
    times = check_times(cfg)
    te = times.index(cfg["eval_time"])
    refs = {seed: fine_spectra(cfg, seed, times)[0] for seed in cfg["calib_seeds"]}
    table = []
    key = "err_all_times" if cfg.get("calib_objective", "eval_time") == "all_times" else "err_eval_time"
    for m in (multipliers or cfg["multipliers"]):
        mlp.set_output_scale(base_scale * m)
        rows = [coarse_errors(cfg, nc, seed, times, refs[seed], k_max, mlp) for seed in cfg["calib_seeds"]]
        stable = all(r["finite_until"] >= cfg["run_time"] for r in rows)
        e = [r["errs"][te] for r in rows]
        err = float(np.mean(e)) if all(x is not None for x in e) else None
        ea = [x for r in rows for x in r["errs"]]
        err_all = float(np.mean(ea)) if all(x is not None for x in ea) else None
        table.append(dict(multiplier=m, err_eval_time=err, err_all_times=err_all, stable=stable,
                          finite_until=[r["finite_until"] for r in rows]))
        log(f"calibration x{m}: mean spectral error at t={cfg['eval_time']} {err}, over all check times "
            f"{err_all}, finite until {[r['finite_until'] for r in rows]}")
    ok = [t for t in table if t["stable"] and t[key] is not None]
    if ok:
        best = min(ok, key=lambda t: t[key])
    else:  # none stays finite: the most stable, then the most accurate
        best = max(table, key=lambda t: (min(t["finite_until"]), -(t["err_eval_time"] or 1e30)))
    mlp.set_output_scale(base_scale * best["multiplier"])
    return best["multiplier"], table


def evaluate(cfg, mlp, log=print, constant=None):
    nf, r, nm = cfg["fine"], cfg["ratio"], cfg["mid"]
    nc = nf // r
    k_max = (nc * 4) // 3  # 2/3 of the coarse grid's Nyquist wavenumber (4 nc / 2)
    times = check_times(cfg)
    te = times.index(cfg["eval_time"])
    rows = []
    for seed in cfg["eval_seeds"]:
        E_refs, fine = fine_spectra(cfg, seed, times)
        out = {"seed": seed, "times": times, "fine_wall_s": fine["wall"], "fine_steps": fine["steps"]}
        models = [("coarse_closure", nc, mlp), ("coarse_plain", nc, None), ("mid_plain", nm, None)]
        if constant is not None:
            models.insert(1, ("coarse_constant", nc, constant))
        for name, n, net in models:
            row = coarse_errors(cfg, n, seed, times, E_refs, k_max, net)
            row["err"] = row["errs"][te]  # the evaluation time
            row["spectrum"] = row["spectra"][te]
            del row["spectra"]
            out[name] = row
            log(f"seed {seed} {name}: spectral error at t={times} {row['errs']}, {row['wall_s']:.2f} s, "
                f"finite until t={row['finite_until']}, stats {row['stats']}")
        out["fine_spectrum"] = E_refs[te][:k_max].tolist()
        rows.append(out)
    return dict(k_max=k_max, runs=rows)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", default="closure_run")
    ap.add_argument("--fine", type=int, default=DEFAULT["fine"])
    ap.add_argument("--ratio", type=int, default=DEFAULT["ratio"])
    ap.add_argument("--mid", type=int, default=DEFAULT["mid"])
    ap.add_argument("--seeds", type=int, default=len(DEFAULT["train_seeds"]))
    ap.add_argument("--epochs", type=int, default=DEFAULT["epochs"])
    ap.add_argument("--snap-times", type=float, nargs="+", default=DEFAULT["snap_times"])
    ap.add_argument("--eval-time", type=float, default=DEFAULT["eval_time"])
    ap.add_argument("--weights", default=None, help="skip training: load weights (.npz) and evaluate")
    ap.add_argument("--dataset", default=None, help="skip the fine runs: load features/targets (.npz)")
    ap.add_argument("--flux", default=DEFAULT["flux"])
    ap.add_argument("--kmax", type=int, default=DEFAULT["kmax"])
    ap.add_argument("--multipliers", type=float, nargs="+", default=DEFAULT["multipliers"],
                    help="candidate output-scale multipliers of the a-posteriori calibration")
    ap.add_argument("--calib-objective", choices=["all_times", "eval_time"], default=DEFAULT["calib_objective"],
                    help="calibrate on the mean spectral error over every check time, or at the evaluation time")
    ap.add_argument("--multiplier", type=float, default=None,
                    help="output-scale multiplier (default: calibrated a posteriori on the calibration seeds)")
    ap.add_argument("--constant-multipliers", type=float, nargs="+", default=DEFAULT["constant_multipliers"],
                    help="candidate multipliers of the constant-coefficient baseline (calibrated the same way)")
    ap.add_argument("--no-constant", action="store_true", help="skip the constant-coefficient baseline")
    ap.add_argument("--constant-multiplier", type=float, default=None,
                    help="multiplier of the constant-coefficient baseline (default: calibrated)")
    ap.add_argument("--no-eval", action="store_true",
                    help="stop after training and calibration (the evaluation runs can follow with --weights)")
    args = ap.parse_args(argv)
    cfg = dict(DEFAULT, fine=args.fine, ratio=args.ratio, mid=args.mid, train_seeds=list(range(args.seeds)),
               epochs=args.epochs, snap_times=args.snap_times, eval_time=args.eval_time, flux=args.flux,
               kmax=args.kmax, multipliers=list(args.multipliers), calib_objective=args.calib_objective)
    os.makedirs(args.out, exist_ok=True)
    log = lambda s: print(s, flush=True)
    PROGRESS["log"] = log
    report = {"config": cfg}
    if args.weights:
        w = np.load(args.weights)
        mlp = od.MLP(params=w["params"])
        mlp.set_scaling(w["shift"], w["scale"], w["lo"], w["hi"])
        base = float(w["out_scale"])
        wts = {k: w[k] for k in ("params", "shift", "scale", "lo", "hi", "out_scale")}
        prev = os.path.join(os.path.dirname(os.path.abspath(args.weights)), "closure_report.json")
        if os.path.exists(prev):  # keep the training record of the weights
            with open(prev) as f:
                report["training"] = json.load(f).get("training")
    else:
        if args.dataset:
            d = np.load(args.dataset)
            X, T = d["X"], d["T"]
        else:
            t0 = time.perf_counter()
            X, T = build_dataset(cfg, log)
            report["dataset_s"] = time.perf_counter() - t0
            np.savez(os.path.join(args.out, "closure_dataset.npz"), X=X, T=T)
        mlp, info = train(cfg, X, T, log=log)
        base = info["out_scale"]
        wts = dict(params=mlp.get_params(), shift=info["shift"], scale=info["scale"], lo=info["lo"], hi=info["hi"],
                   out_scale=base)
        report["training"] = {k: (v.tolist() if isinstance(v, np.ndarray) else v) for k, v in info.items()}
    if args.multiplier is None:
        mult, table = calibrate(cfg, mlp, base, log)
        report["calibration"] = dict(multiplier=mult, trials=table)
    else:
        mult = args.multiplier
        mlp.set_output_scale(base * mult)
        report["calibration"] = dict(multiplier=mult, trials=None)
    # out_scale is the trained scale; the closure runs with out_scale * output_multiplier
    np.savez(os.path.join(args.out, "closure_weights.npz"), output_multiplier=mult, **wts)
    const = None
    if not args.no_constant:
        const = constant_model(base, wts["shift"], wts["scale"], wts["lo"], wts["hi"])
        if args.constant_multiplier is None:
            cmult, ctable = calibrate(cfg, const, base, log, multipliers=list(args.constant_multipliers))
        else:
            cmult, ctable = args.constant_multiplier, None
            const.set_output_scale(base * cmult)
        report["constant_baseline"] = dict(multiplier=cmult, trials=ctable)
        log(f"constant-coefficient baseline: multiplier x{cmult}")
    if args.no_eval:
        with open(os.path.join(args.out, "closure_report.json"), "w") as f:
            json.dump(report, f)
        return
    report["evaluation"] = evaluate(cfg, mlp, log, constant=const)
    with open(os.path.join(args.out, "closure_report.json"), "w") as f:
        json.dump(report, f)
    names = ("coarse_closure", "coarse_constant", "coarse_plain", "mid_plain")
    log(json.dumps([{k: r[k]["err"] for k in names if k in r} for r in report["evaluation"]["runs"]]))


if __name__ == "__main__":
    main()
