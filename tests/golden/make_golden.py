"""Write the golden fixtures of the FR P=3 step from the CPU oracle.

    python tests/golden/make_golden.py

Small cases are stored as input/output vectors (golden_small.npz); the 8^3
periodic Euler trajectory (Rusanov, f64, CFL 0.3, 100 steps) is stored as
SHA-256 digests of the state bytes after 1, 10 and 100 steps
(golden_euler8.json) since its IC is regenerated deterministically.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from helpers import smooth_state  # noqa: E402
from opendynamo_amd import ic  # noqa: E402
from oracle import oracle as O  # noqa: E402

TWO_PI = 2 * np.pi

SMALL = {
    # name: (dims, lengths, params, dt, steps, seed)
    "euler_rusanov_f64": ((2, 2, 2), (TWO_PI, TWO_PI, TWO_PI), dict(u_real=8, flux="rusanov", viscous=False), 2e-3, 10, 0),
    "ns_slau2_wall_suth_f64": ((2, 3, 2), (TWO_PI, 2.0, np.pi),
                               dict(u_real=8, flux="slau2", y_wall=True, viscosity_model=1, mu_ref=1e-3), 1e-3, 5, 1),
    "ns_ausm_f32": ((3, 2, 2), (TWO_PI, TWO_PI, np.pi), dict(u_real=4, flux="ausm+", mu_ref=2e-3), 1e-3, 5, 2),
    "mixed_u64_c32_accum": ((2, 2, 2), (TWO_PI, 2.0, np.pi),
                            dict(u_real=8, compute_real=4, q_real=4, rhs_real=4, rk_accum=1, flux="slau2",
                                 y_wall=True, mu_ref=1e-3), 1e-3, 5, 3),
}


def euler8():
    L = float(np.float32(TWO_PI))
    p = O.params(8, 8, 8, L, L, L, u_real=8, flux="rusanov", viscous=False)
    u = ic.fourier_state(8, 8, 8, L, L, L, u0=(0.1, 0.0, 0.0), amplitude=1e-2, n_modes=8, seed=0)
    return p, u


def main():
    arrays = {}
    meta = {}
    for name, (dims, lens, kw, dt, steps, seed) in SMALL.items():
        p = O.params(*dims, *lens, **kw)
        real = kw.get("u_real", 8)
        u = smooth_state(*dims, *lens, seed=seed, amp=0.06, dtype=np.float64 if real == 8 else np.float32)
        arrays[name + "_in"] = u
        arrays[name + "_out"] = O.step(p, u, dt, steps)
        meta[name] = dict(dims=dims, lengths=lens, params=kw, dt=dt, steps=steps)
    np.savez_compressed(os.path.join(HERE, "golden_small.npz"), **arrays)
    with open(os.path.join(HERE, "golden_small.json"), "w") as f:
        json.dump(meta, f, indent=1)
    p, u = euler8()
    dt = O.cfl_dt(p, u, 0.3)
    digests = {"dt": dt, "sha256_in": hashlib.sha256(u.tobytes()).hexdigest()}
    cur, done = u, 0
    for n in (1, 10, 100):
        cur = O.step(p, cur, dt, n - done)
        done = n
        digests[f"sha256_after_{n}"] = hashlib.sha256(cur.tobytes()).hexdigest()
        digests[f"sums_after_{n}"] = O.sums(p, cur).tolist()
    with open(os.path.join(HERE, "golden_euler8.json"), "w") as f:
        json.dump(digests, f, indent=1)
    print("golden fixtures written")


if __name__ == "__main__":
    main()
