"""Golden fixtures: the oracle (CPU) and the HIP library (GPU) both
reproduce the committed vectors bit for bit."""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SMALL = json.load(open(os.path.join(HERE, "golden_small.json")))
DATA = np.load(os.path.join(HERE, "golden_small.npz"))
EULER8 = json.load(open(os.path.join(HERE, "golden_euler8.json")))


def _euler8():
    from golden.make_golden import euler8
    return euler8()


@pytest.mark.parametrize("name", sorted(SMALL))
def test_oracle_reproduces_small_golden(name):
    m = SMALL[name]
    p = O.params(*m["dims"], *m["lengths"], **m["params"])
    out = O.step(p, DATA[name + "_in"], m["dt"], m["steps"])
    np.testing.assert_array_equal(out, DATA[name + "_out"])


def test_oracle_reproduces_euler8_digests():
    p, u = _euler8()
    assert hashlib.sha256(u.tobytes()).hexdigest() == EULER8["sha256_in"]
    assert O.cfl_dt(p, u, 0.3) == EULER8["dt"]
    out = O.step(p, u, EULER8["dt"], 10)
    assert hashlib.sha256(out.tobytes()).hexdigest() == EULER8["sha256_after_10"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(SMALL))
def test_gpu_reproduces_small_golden(name):
    import opendynamo_amd as od
    m = SMALL[name]
    print("CASE golden", name, flush=True)
    with od.Solver(od.params(*m["dims"], *m["lengths"], **m["params"])) as s:
        s.set_state(DATA[name + "_in"])
        s.step(m["dt"], m["steps"])
        np.testing.assert_array_equal(s.get_state(), DATA[name + "_out"])


@pytest.mark.gpu
def test_gpu_reproduces_euler8_digests():
    import opendynamo_amd as od
    p, u = _euler8()
    print("CASE golden euler 8^3 100 steps", flush=True)
    gp = od.params(p.nx, p.ny, p.nz, p.lx, p.ly, p.lz, u_real=8, flux="rusanov", viscous=False)
    with od.Solver(gp) as s:
        s.set_state(u)
        done = 0
        for n in (1, 10, 100):
            s.step(EULER8["dt"], n - done)
            done = n
            got = s.get_state()
            assert hashlib.sha256(got.tobytes()).hexdigest() == EULER8[f"sha256_after_{n}"]
            np.testing.assert_allclose(s.sums(), EULER8[f"sums_after_{n}"], rtol=1e-13, atol=1e-13)
