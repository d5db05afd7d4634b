"""The headline workload at its full size (48^3 elements, 7,077,888 DOF, f32
state; channel IC with a 16-mode trip, SLAU2, Sutherland, y slip walls):

* the strict build equals the f32 op-for-op oracle byte for byte after two
  RK3 steps (the small-mesh parity tests at the size that runs);
* the fast build stays within the fast-math tolerance of the FP64 oracle;
* mass is conserved to f32 round-off (slip walls carry no mass flux, x and z
  are periodic, the body force acts on x-momentum only) in both builds.
"""
import os
import sys

import numpy as np
import pytest

import opendynamo_amd as od
from opendynamo_amd import ic
from oracle import oracle as O

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench as B  # noqa: E402  (the workload definition)

pytestmark = pytest.mark.gpu

STEPS = 2
FAST_TOL = 2e-5  # max |gpu - oracle_f64| / max |oracle_f64| per variable, as for the small fast-math cases


@pytest.fixture(scope="module")
def workload():
    u0 = ic.channel_state(B.NX, B.NY, B.NZ, B.LX, B.LY, B.LZ, B.MACH, perturbation_amplitude=0.05,
                          perturbation_modes=16, perturbation_seed=1, dtype=np.float32)
    with od.Solver(od.params(B.NX, B.NY, B.NZ, B.LX, B.LY, B.LZ, **B.workload_kwargs(od.OD_F32))) as s:
        s.set_state(u0)
        dt = s.cfl_dt(B.CFL)
    return u0, dt


def gpu_run(u0, dt, fast):
    p = od.params(B.NX, B.NY, B.NZ, B.LX, B.LY, B.LZ, **B.workload_kwargs(od.OD_F32, fast_math=fast))
    with od.Solver(p) as s:
        s.set_state(u0)
        m0 = s.sums()[0]
        s.step(dt, STEPS)
        return s.get_state(), m0, s.sums()[0]


def test_strict_build_bitwise_at_full_size(workload):
    u0, dt = workload
    print("CASE fullsize strict 48^3 f32", flush=True)
    got, m0, m1 = gpu_run(u0, dt, 0)
    po = O.params(B.NX, B.NY, B.NZ, B.LX, B.LY, B.LZ, n_threads=B.cpu_threads(), **B.workload_kwargs(4))
    ref = O.step(po, u0, dt, STEPS)
    assert got.tobytes() == np.asarray(ref, dtype=np.float32).tobytes()
    assert abs(m1 - m0) <= 1e-7 * abs(m0)


def test_fast_build_against_fp64_oracle_at_full_size(workload):
    u0, dt = workload
    print("CASE fullsize fast 48^3 f32", flush=True)
    got, m0, m1 = gpu_run(u0, dt, 1)
    po = O.params(B.NX, B.NY, B.NZ, B.LX, B.LY, B.LZ, n_threads=B.cpu_threads(), **B.workload_kwargs(8))
    ref = O.step(po, u0.astype(np.float64), dt, STEPS).reshape(5, -1)
    g = got.astype(np.float64).reshape(5, -1)
    err = np.max(np.abs(g - ref), axis=1) / np.max(np.abs(ref), axis=1)
    print("fast-math rel err per var", err, flush=True)
    assert np.all(err < FAST_TOL), err
    assert abs(m1 - m0) <= 1e-7 * abs(m0)
