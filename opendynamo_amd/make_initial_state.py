"""Write the label-0 initial state file for a config (counterpart of
scripts/make_initial_state.py).

    python -m opendynamo_amd.make_initial_state [config.toml]

Reads the config with the C++ reader (same keys and defaults as the solver),
evaluates [physics].initial_condition:
  "channel"  laminar channel u = 1.5 M (1 - eta^2)    (reference)
  "uniform"  uniform flow u = M                        (reference)
  "fourier"  uniform flow u = M plus the periodic divergence-free
             random-Fourier perturbation (no reference counterpart)
with the perturbation_* keys, and writes
<output_dir>/<prefix>_<f32|f64>_000000.h5 (step 0, t 0) with the [io]
compression.  channel / uniform are evaluated in float32 as the reference
does; fourier in float64.
"""
import os
import sys

import numpy as np

from . import host, ic


def build_state(cfg):
    m, ph, pr = cfg["mesh"], cfg["physics"], cfg["precision"]
    nx, ny, nz, lx, ly, lz = m["nx"], m["ny"], m["nz"], m["lx"], m["ly"], m["lz"]
    dtype = np.float32 if pr["u_real"] == "f32" else np.float64
    kind = ph["initial_condition"]
    if kind == "fourier":
        return ic.fourier_state(nx, ny, nz, lx, ly, lz, u0=(ph["mach_bulk"], 0.0, 0.0),
                                amplitude=ph["perturbation_amplitude"], n_modes=max(1, ph["perturbation_modes"]),
                                seed=ph["perturbation_seed"], dtype=dtype)
    return ic.channel_state(nx, ny, nz, lx, ly, lz, ph["mach_bulk"], initial_condition=kind,
                            perturbation_amplitude=ph["perturbation_amplitude"],
                            perturbation_modes=ph["perturbation_modes"], perturbation_seed=ph["perturbation_seed"],
                            dtype=dtype)


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    path = argv[0] if argv else "./config.toml"
    cfg = host.config(path)
    state = build_state(cfg)
    m, io = cfg["mesh"], cfg["io"]
    os.makedirs(io["output_dir"], exist_ok=True)
    out = host.save_state(0, 0.0, state, m["nx"], m["ny"], m["nz"], m["lx"], m["ly"], m["lz"], io["output_dir"],
                          io["prefix"], io["compression"])
    s = state.reshape(5, -1)
    u = s[1] / s[0]
    print(f"Initial state written: {out}")
    print(f"  mesh {m['nx']}x{m['ny']}x{m['nz']}  IC {cfg['physics']['initial_condition']}  "
          f"mach_bulk {cfg['physics']['mach_bulk']}")
    print(f"  u range [{u.min():.6f}, {u.max():.6f}]  |v'| max {np.abs(s[2] / s[0]).max():.6f}  "
          f"|w'| max {np.abs(s[3] / s[0]).max():.6f}")
    return out


if __name__ == "__main__":
    main()
