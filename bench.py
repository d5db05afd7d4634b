"""Benchmark: FR P=3 compressible Navier-Stokes RK3 step on one MI355X.

Workload (the headline configuration): 48x48x48 hex elements (7,077,888
solution points), box 2pi x 2 x pi, slip walls in y, periodic x/z, f32
state/RHS/flux, SLAU2, Sutherland viscosity, Mach 0.1, Re_bulk 3000
(mu_ref = rho U_b h / Re = 0.1 / 3000, h = ly/2), laminar channel
u = 1.5 M (1 - eta^2) plus the wall-damped divergence-free trip (16 modes,
amplitude 0.05, seed 1), dt from CFL 0.3.

    python bench.py [--gpus 1] [--steps 300] [--warmup 10] [--cpu-steps 20] [--copy-calibration]

A step is one full SSP-RK3 step (3 stages, one hipGraph launch) over the
resident state.  Prints one JSON line.

    python bench.py --workload ensemble [--steps 500]

measures the ensemble instead: 24 independent periodic cubes [0, 2pi)^3 with
16..40 elements per side (seed 7), P=3 NS f32 fast math, SLAU2, fixed mu
1.6e-4, no body force, Mach 0.1 mean flow plus a divergence-free Fourier
field per box (12..32 modes, amplitude 0.02..0.1, own seed), per-box dt from
CFL 0.3; the autotuner picks between one launch per kernel and stage for
all boxes (several tile/grid schedules) and box-by-box launches inside the
step graph; both are reported.  Also one JSON line.

    python bench.py --workload precision

runs the precision study on the 8^3 periodic Euler case (one JSON line).
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GDOF-steps/s, 48³ P=3 f32; speedup of 32³+closure vs plain 64³ at equal error"
HBM_PEAK_GBS = 8000.0
FP32_VECTOR_PEAK_TFLOPS = 157.3
REFERENCE_GDOF = 0.114  # README.md:205-209, GTX 1650 Ti, 48^3 f32 (~62 ms/step)

NX = NY = NZ = 48
LX, LY, LZ = 2 * np.pi, 2.0, np.pi
MACH, RE_BULK = 0.1, 3000.0
MU_REF = 1.0 * MACH * (LY / 2) / RE_BULK
CFL = 0.3
# Compulsory state traffic of SSP-RK3 with two registers, f32: stage 1 reads
# u^n and writes u1 (2 x 20 B), stages 2 and 3 read u^n, u_k and write
# (3 x 20 B): 160 B per DOF-step, i.e. 160/3 B per DOF per update launch.
ALG_BYTES_PER_DOF_STEP = 160.0
FLOP_PER_DOF_STEP = 3.0e3


def workload_kwargs(real, fast_math=0):
    kw = dict(u_real=real, flux="slau2", y_wall=True, viscous=True, viscosity_model=1, mu_ref=MU_REF)
    if fast_math:
        kw["fast_math"] = 1
    return kw


def run_gpu(od, u0, p, steps, warmup):
    """Graph-timed steps + per-kernel event timing on the resident state."""
    with od.Solver(p) as s:
        s.set_state(u0)
        dt = s.cfl_dt(CFL)
        # per-kernel event timing first (KERNEL_TIMING_STEPS steps, each
        # launch bracketed by events), so that the GPU is busy right up to
        # the warm-up steps: the first ~40 ms of work after an idle spell run
        # ~4 % slower (clock ramp; 20-step windows measured back to back read
        # 1.12 ms, then 1.073-1.076 ms per step at 48^3)
        kern = s.time_kernels(dt, KERNEL_TIMING_STEPS)
        s.step(dt, warmup)
        s.synchronize()
        t0 = time.perf_counter()
        ms = s.time_steps(dt, steps)
        s.synchronize()
        wall = time.perf_counter() - t0
        info = s.info()
        ok = bool(np.isfinite(s.get_state()).all())
    return dict(dt=dt, ms=ms, wall=wall, kern=kern, info=info, ok=ok)


def error_vs(ref, od, u0, p, dt, steps):
    with od.Solver(p) as s:
        s.set_state(u0)
        s.step(dt, steps)
        g = s.get_state().astype(np.float64).reshape(5, -1)
    r = ref.reshape(5, -1)
    names = ["rho", "rhou", "rhov", "rhow", "E"]
    return {
        "steps": steps,
        "max_rel": {n: float(np.max(np.abs(g[v] - r[v])) / np.max(np.abs(r[v]))) for v, n in enumerate(names)},
        "l2_rel": {n: float(np.linalg.norm(g[v] - r[v]) / np.linalg.norm(r[v])) for v, n in enumerate(names)},
    }


PMC_PROFILE = "profiles/fr_step_kernels_48cubed_f32_pmc.txt"
KERNEL_TIMING_STEPS = 30


def traffic_from_profile(kernel):
    """HBM bytes per launch of `kernel` (its fast-math build) measured with
    rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel
    calibration), from the committed summary, if it holds that kernel."""
    path = os.path.join(ROOT, PMC_PROFILE)
    if not os.path.exists(path):
        return None
    pat = r"traffic_bytes_per_launch\[od::fastmath::" + re.escape(kernel) + r"[^\]]*\]:\s*([0-9.eE+]+)"
    m = re.search(pat, open(path).read())
    return float(m.group(1)) if m else None


def hip_copy_calibration(nbytes):
    """Three device-to-device copies of `nbytes` through the HIP runtime (its
    blit copy kernel): the known-size reference for FETCH_SIZE / WRITE_SIZE."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    n = ctypes.c_size_t(nbytes)
    if hip.hipMalloc(ctypes.byref(a), n) or hip.hipMalloc(ctypes.byref(b), n):
        raise RuntimeError("hipMalloc failed in the copy calibration")
    hip.hipMemset(a, 0, n)
    for _ in range(3):
        hip.hipMemcpy(b, a, n, 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    return {"kernel": "HIP runtime device-to-device copy", "bytes_read": nbytes, "bytes_written": nbytes,
            "launches": 3}


def device_copy_rate(nbytes=1 << 30, reps=3):
    """Bytes moved (read + written) per second by the HIP runtime's own
    device-to-device copy of `nbytes`, event-timed on the null stream after
    one untimed copy: what this part delivers to a plain streaming kernel,
    the practical ceiling for a bandwidth-bound kernel's traffic."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    vp = ctypes.c_void_p
    a, b, e0, e1 = vp(), vp(), vp(), vp()
    n = ctypes.c_size_t(nbytes)
    if hip.hipMalloc(ctypes.byref(a), n) or hip.hipMalloc(ctypes.byref(b), n):
        raise RuntimeError("hipMalloc failed in the copy-rate measurement")
    try:
        if hip.hipMemset(a, 0, n) or hip.hipEventCreate(ctypes.byref(e0)) or hip.hipEventCreate(ctypes.byref(e1)):
            raise RuntimeError("copy-rate set-up failed")
        hip.hipMemcpyAsync(b, a, n, 3, vp())  # hipMemcpyDeviceToDevice
        hip.hipEventRecord(e0, vp())
        for _ in range(reps):
            hip.hipMemcpyAsync(b, a, n, 3, vp())
        hip.hipEventRecord(e1, vp())
        hip.hipEventSynchronize(e1)
        ms = ctypes.c_float()
        hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    return 2.0 * nbytes * reps / (ms.value / 1e3) / 1e9


def host_cpu():
    """The host's CPUs as this process sees them: affinity count, logical CPU
    count, the CPU model string and the thread allotment of the environment
    (OMP_NUM_THREADS, which a shared host sets to each GPU's share of CPUs)."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    env = os.environ.get("OMP_NUM_THREADS")
    allot = int(env) if env and env.isdigit() and int(env) > 0 else None
    return {"affinity_cpus": affinity, "logical_cpus": os.cpu_count(), "model": model,
            "omp_num_threads": allot}


def cpu_threads():
    """Oracle worker threads: every CPU this process may run on, limited to the
    environment's allotment (OMP_NUM_THREADS) where one is set."""
    h = host_cpu()
    n = h["affinity_cpus"]
    if h["omp_num_threads"]:
        n = min(n, h["omp_num_threads"])
    return max(1, n)


def run_ensemble(args):
    """The ensemble workload (see the module docstring)."""
    import opendynamo_amd as od
    from opendynamo_amd import ic

    boxes = ic.ensemble_boxes()
    p = od.params(1, 1, 1, 1.0, 1.0, 1.0, u_real=od.OD_F32, flux="slau2", y_wall=False, viscous=True,
                  viscosity_model=0, mu_ref=0.00016, body_force_x=0.0, fast_math=1)
    dims = [(b["nx"], b["ny"], b["nz"], b["lx"], b["ly"], b["lz"]) for b in boxes]
    dof = sum(64 * b["nx"] * b["ny"] * b["nz"] for b in boxes)
    with od.Batch(p, dims) as B:
        for q, b in enumerate(boxes):
            B.set_state(q, ic.ensemble_state(b, np.float32))
        dt = B.cfl_dt(CFL)
        B.synchronize()
        best, best_ms, table = B.autotune(dt, steps_per_trial=3)
        default_ms = [t["ms_per_step"] for t in table
                      if t["orient"] == 0 and t["pairs"] == 1 and t["order"] == 0 and t["groups_per_cu"] == 0][0]
        batched = min((t for t in table if t["order"] != 2), key=lambda t: t["ms_per_step"])
        box_by_box = [t["ms_per_step"] for t in table if t["order"] == 2][0]
        B.step(dt, args.warmup)
        B.synchronize()
        t0 = time.perf_counter()
        ms = B.time_steps(dt, args.steps)
        wall = time.perf_counter() - t0
        per_box_steps = max(1, min(20, args.steps))
        ms_per_box = B.time_steps(dt, per_box_steps, per_box=True) / per_box_steps
        info = B.info()
        finite = all(bool(np.isfinite(B.get_state(q)).all()) for q in range(len(boxes)))
        sums = B.sums()
    value = dof * args.steps / (ms / 1e3) / 1e9
    out = {
        "metric": "GDOF-steps/s, ensemble of 24 periodic boxes 16..40^3, P=3 NS f32", "value": round(value, 4),
        "unit": "GDOF·steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {
            "workload": "24 independent periodic cubes [0,2pi)^3, n ~ U{16..40} elements per side (seed 7), "
                        "P=3 NS f32 fast math, SLAU2, mu 1.6e-4, Mach 0.1 + div-free Fourier (12-32 modes, "
                        "amp 0.02-0.1, per-box seed), per-box dt from CFL 0.3",
            "boxes": [b["nx"] for b in boxes], "dof": dof, "launches_per_step": info["launches_per_step"],
            "device_bytes": info["device_bytes"],
        },
        "schedule": best, "autotune_trials": len(table), "autotune_best_ms": round(best_ms, 5),
        "default_schedule_ms": round(default_ms, 5),
        "one_launch_per_stage_best": {k: batched[k] for k in ("orient", "pairs", "order", "groups_per_cu")},
        "one_launch_per_stage_best_ms": round(batched["ms_per_step"], 5),
        "box_by_box_in_graph_ms": round(box_by_box, 5),
        "per_box_launches_no_graph_ms_per_step": round(ms_per_box, 5),
        "wall_s_timed": round(wall, 4), "state_finite": finite,
        "mass_per_box_min_max": [float(sums[:, 0].min()), float(sums[:, 0].max())],
        "autotune_table": table,
    }
    print(json.dumps(out), flush=True)


def run_precision(args):
    """Precision study on the 8^3 periodic Euler case (Rusanov, CFL 0.3, Fourier
    IC 8 modes amp 1e-2 seed 0): per-variable error against the FP64 CPU
    oracle after 10 and 100 steps for the all-f64 GPU pipeline, f64 state with
    f32 flux/Riemann/derivatives and the RK accumulation in f64, and all-f32;
    mass/energy drift and the mean rho*u profile after 2000 steps."""
    import opendynamo_amd as od
    from opendynamo_amd import ic
    from oracle import oracle as O

    n, L = 8, 2 * np.pi
    base = dict(flux="rusanov", y_wall=False, viscous=False, body_force_x=0.0)
    u0 = ic.fourier_state(n, n, n, L, L, L, u0=(0.1, 0.0, 0.0), amplitude=1e-2, n_modes=8, seed=0)
    pipelines = {
        "f64": dict(u_real=od.OD_F64),
        "f64_state_f32_flux_f64_accum": dict(u_real=od.OD_F64, compute_real=od.OD_F32, q_real=od.OD_F32,
                                             rhs_real=od.OD_F32, rk_accum=1),
        "f32": dict(u_real=od.OD_F32),
    }
    po = O.params(n, n, n, L, L, L, n_threads=cpu_threads(), **base)
    dt = O.cfl_dt(po, u0, CFL)
    refs = {10: O.step(po, u0, dt, 10)}
    refs[100] = O.step(po, refs[10], dt, 90)
    names = ["rho", "rhou", "rhov", "rhow", "E"]
    out = {"metric": "per-variable error vs the FP64 oracle, 8^3 periodic Euler (precision study)", "dt": dt,
           "pipelines": {}}
    long_states = {}
    for name, kw in pipelines.items():
        p = od.params(n, n, n, L, L, L, **base, **kw)
        row = {}
        with od.Solver(p) as s:
            s.set_state(u0)
            s0 = s.sums()
            done = 0
            for k in (10, 100):
                s.step(dt, k - done)
                done = k
                g = s.get_state(np.float64).reshape(5, -1)
                r = refs[k].reshape(5, -1)
                row[f"after_{k}"] = {
                    "max_rel": {nm: float(np.max(np.abs(g[v] - r[v])) / np.max(np.abs(r[v]))) for v, nm in enumerate(names)},
                    "l2_rel": {nm: float(np.linalg.norm(g[v] - r[v]) / np.linalg.norm(r[v])) for v, nm in enumerate(names)},
                }
            s.step(dt, 2000 - done)
            s2 = s.sums()
            row["drift_2000"] = {"mass_rel": float((s2[0] - s0[0]) / s0[0]), "energy_rel": float((s2[4] - s0[4]) / s0[4])}
            long_states[name] = s.get_state(np.float64).reshape(5, n, n, n, 4, 4, 4)
        out["pipelines"][name] = row
    # mean rho*u profile over y (x, z and in-element x, z points averaged), against the all-f64 GPU run
    prof = {k: v[1].mean(axis=(0, 2, 3, 5)).ravel() for k, v in long_states.items()}  # [ey][j]
    for k in prof:
        out["pipelines"][k]["mean_rhou_profile_max_abs_diff_vs_f64_2000"] = float(np.max(np.abs(prof[k] - prof["f64"])))
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-steps", type=int, default=20, help="oracle steps for the CPU baseline (0 = skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 1 GiB device-to-device copy of known size, the reference "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes")
    ap.add_argument("--workload", choices=["channel48", "ensemble", "precision"], default="channel48")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU slab decomposition is a design, not a build).", file=sys.stderr)
        sys.exit(2)
    if args.workload == "ensemble":
        run_ensemble(args)
        return
    if args.workload == "precision":
        run_precision(args)
        return

    import opendynamo_amd as od
    from opendynamo_amd import ic

    u0 = ic.channel_state(NX, NY, NZ, LX, LY, LZ, MACH, perturbation_amplitude=0.05, perturbation_modes=16,
                          perturbation_seed=1, dtype=np.float32)
    dof = NX * NY * NZ * 64
    # headline: fast_math = true, the reference's default (config.rs:89-91);
    # the strict IEEE build (bit-identical to the oracle) is reported beside it
    p = od.params(NX, NY, NZ, LX, LY, LZ, **workload_kwargs(od.OD_F32, fast_math=1))
    p_strict = od.params(NX, NY, NZ, LX, LY, LZ, **workload_kwargs(od.OD_F32))
    r = run_gpu(od, u0, p, args.steps, args.warmup)
    rs = run_gpu(od, u0, p_strict, args.steps, args.warmup)
    dt, ms, wall, kern, info, state_ok = r["dt"], r["ms"], r["wall"], r["kern"], r["info"], r["ok"]
    value = dof * args.steps / (ms / 1e3) / 1e9
    upd_name = next(k for k in kern if k.startswith("od_update"))  # od_update_block_kernel on even meshes
    upd_ms = kern[upd_name]
    alg_bytes_launch = ALG_BYTES_PER_DOF_STEP / 3.0 * dof
    achieved = alg_bytes_launch / (upd_ms / 1e3) / 1e9
    traffic = traffic_from_profile(upd_name)
    copy_gbs = device_copy_rate()
    roofline = {
        "bound": "hbm", "kernel": upd_name, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
        # the roofline is priced on HBM bytes; what limits the kernels as
        # measured (rocprofv3 FETCH/WRITE_SIZE and SQ_INSTS_VALU in the
        # committed profile): the update kernel's traffic arrives at or
        # above the rate of a plain device copy (part of it from the
        # memory-side cache), the visc kernel is VALU-issue bound
        "limiter": "memory (update kernel, HBM + memory-side cache) + valu (visc kernel)",
        "traffic_source": f"{PMC_PROFILE}: rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE of {upd_name} in separate "
                          "profiling runs of this workload, copy-calibrated (not measured in this run)",
        "algorithmic_bytes_per_launch": alg_bytes_launch, "launch_ms": round(upd_ms, 5),
        "launch_ms_over_steps": KERNEL_TIMING_STEPS,
        # the profiled HBM bytes per launch over the live launch time: how much
        # of the peak the kernel actually moves (the design's face traces and
        # viscous face fluxes on top of the compulsory state traffic)
        "traffic_frac": round(traffic / (upd_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 5) if traffic else None,
        # the same bytes against what a plain device copy reaches on this part
        # (measured in this run, after the timed steps)
        "device_copy_GBps": round(copy_gbs, 1),
        "traffic_frac_of_copy": round(traffic / (upd_ms / 1e3) / 1e9 / copy_gbs, 4) if traffic else None,
        "step_frac_alg_hbm": round(value * 1e9 * ALG_BYTES_PER_DOF_STEP / 1e9 / HBM_PEAK_GBS, 5),
        "valu_frac_alg_flops": round(value * FLOP_PER_DOF_STEP / 1e3 / FP32_VECTOR_PEAK_TFLOPS, 5),
    }
    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "GDOF·steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms / args.steps, 5),
        "higher_is_better": True, "scaling": "weak",
        "vs_baseline": round(value / REFERENCE_GDOF, 2), "dtype": "f32", "data": "synthetic",
        "config": {
            "workload": "FR P=3 compressible NS, 48x48x48 hex elements (7,077,888 DOF), box 2pi x 2 x pi, "
                        "y slip walls, x/z periodic, SLAU2, Sutherland, Mach 0.1, Re_bulk 3000, "
                        "channel IC + 16-mode trip (amp 0.05, seed 1), dt from CFL 0.3",
            "elements": [NX, NY, NZ], "dof": dof, "dt": dt, "launches_per_step": info["launches_per_step"],
            "device_bytes": info["device_bytes"], "baseline_reference": "0.114 GDOF-steps/s, GTX 1650 Ti, README.md:205",
        },
        "roofline": roofline,
        "kernel_ms": {k: round(v, 5) for k, v in kern.items()},
        "wall_s_timed": round(wall, 4),
        "state_finite": state_ok,
        "fast_math": True,
        "strict_ieee": {
            "value": round(dof * args.steps / (rs["ms"] / 1e3) / 1e9, 4),
            "ms_per_step": round(rs["ms"] / args.steps, 5),
            "kernel_ms": {k: round(v, 5) for k, v in rs["kern"].items()},
            "note": "fast_math = false: IEEE division/sqrt, no contraction; bitwise equal to the CPU oracle",
        },
    }
    if args.cpu_steps > 0:
        from oracle import oracle as O
        threads = cpu_threads()
        po = O.params(NX, NY, NZ, LX, LY, LZ, n_threads=threads, **workload_kwargs(8))
        u64 = u0.astype(np.float64)
        t0 = time.perf_counter()
        ref = O.step(po, u64, dt, args.cpu_steps)
        cpu_s = time.perf_counter() - t0
        cpu_value = dof * args.cpu_steps / cpu_s / 1e9
        hc = host_cpu()
        out["cpu_baseline"] = {
            "value": round(cpu_value, 6), "unit": "GDOF·steps/s", "cores": threads, "kind": "port",
            "sample": f"{args.cpu_steps} RK3 steps of the same 48^3 workload from the same IC, FP64 op-for-op "
                      f"oracle (C++ std::thread over element ranges, -O3 -march=native, no FP contraction), "
                      f"{threads} threads, {cpu_s:.1f} s",
            "per_thread_value": round(cpu_value / threads, 7),
            "host": hc,
            "threads_note": "threads = the CPUs this process may use, capped by OMP_NUM_THREADS when the "
                            "environment sets it (a shared host's share of CPUs per GPU)",
        }
        out["speedup_vs_cpu"] = round(value / cpu_value, 1)
        # per-variable error of the f32 GPU pipelines against the FP64 oracle
        out["error_vs_fp64_oracle"] = error_vs(ref, od, u0, p, dt, args.cpu_steps)
        out["strict_ieee"]["error_vs_fp64_oracle"] = error_vs(ref, od, u0, p_strict, dt, args.cpu_steps)
    if args.copy_calibration:
        out["copy_calibration"] = hip_copy_calibration(1 << 30)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
