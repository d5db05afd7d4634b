"""Host layer on the CPU: config.toml keys/defaults/errors (config.rs) and the
HDF5 state files (io.rs): write -> read round trips, attrs, labels,
compression modes and the loader's validation errors."""
import os

import numpy as np
import pytest

from opendynamo_amd import host, ic

CFG = """
# full key set
[mesh]
nx = 3
ny = 4
nz = 2
lx = 12.566370614359172
ly = 2.0
lz = 4.188790204786391

[physics]
mach_bulk = 0.3
initial_condition = "channel"
y_boundary = "wall"       # trailing comment
viscosity_model = "sutherland"
mu_ref = 0.005556
t_ref = 273.15
s_suth = 110.4
perturbation_amplitude = 0.2
perturbation_modes = 16
perturbation_seed = 7
initial_state = "{out}/flow_f32_000000.h5"

[solver]
n_steps = 4
dt = 5.0e-4
flux_scheme = "slau2"
fast_math = false

[io]
save_freq = 2
checkpoint_freq = 3
compression = "gzip4"
output_dir = "{out}"
prefix = "flow"

[precision]
u_real = "f32"
q_real = "f32"
rhs_real = "f32"
compute_real = "f32"
"""

MINIMAL = """
[mesh]
nx = 2
ny = 2
nz = 2
[physics]
mach_bulk = 0.1
[solver]
n_steps = 1
dt = 0.001
flux_scheme = "rusanov"
[io]
save_freq = 1
output_dir = "out"
prefix = "p"
[precision]
u_real = "f64"
q_real = "f64"
rhs_real = "f64"
compute_real = "f64"
"""


def write(tmp_path, text, name="config.toml"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_full_config(tmp_path):
    c = host.config(write(tmp_path, CFG.format(out=tmp_path)))
    assert c["mesh"] == {"nx": 3, "ny": 4, "nz": 2, "lx": float(np.float32(12.566370614359172)), "ly": 2.0,
                         "lz": float(np.float32(4.188790204786391))}
    assert c["physics"]["viscosity_model"] == "sutherland" and c["physics"]["y_boundary"] == "wall"
    assert c["physics"]["mu_ref"] == 0.005556
    assert c["physics"]["t_ref"] == float(np.float32(273.15))
    assert c["solver"]["dt"] == float(np.float32(5e-4)) and c["solver"]["fast_math"] is False
    assert c["io"]["checkpoint_freq"] == 3 and c["io"]["compression"] == "gzip4"
    assert c["precision"]["u_real"] == "f32"


def test_reference_defaults(tmp_path):
    """Omitted keys take config.rs's defaults (f32 arithmetic for lx, lz)."""
    c = host.config(write(tmp_path, MINIMAL))
    assert c["mesh"]["lx"] == float(np.float32(4.0) * np.float32(np.pi))
    assert c["mesh"]["ly"] == 2.0
    assert c["mesh"]["lz"] == float((np.float32(4.0) / np.float32(3.0)) * np.float32(np.pi))
    ph = c["physics"]
    assert ph["initial_condition"] == "channel" and ph["y_boundary"] == "wall" and ph["viscosity_model"] == "fixed"
    assert ph["mu_ref"] == 0.00016 and ph["perturbation_amplitude"] == 0.0
    assert ph["perturbation_modes"] == 16 and ph["perturbation_seed"] == 7 and ph["initial_state"] is None
    assert ph["body_force_x"] == 0.001
    assert c["solver"]["fast_math"] is True
    assert c["io"]["checkpoint_freq"] == 0 and c["io"]["compression"] == "blosc-lz4"
    assert c["precision"]["rk_accum"] == "compute"


@pytest.mark.parametrize("edit,msg", [
    (("nx = 2\n", ""), "missing field `nx`"),
    (('flux_scheme = "rusanov"', 'flux_scheme = "roe"'), "unknown variant `roe`"),
    (('u_real = "f64"', 'u_real = "f16"'), "unknown variant `f16`"),
    (("n_steps = 1", 'n_steps = "one"'), "invalid type"),
    (("dt = 0.001\n", ""), "missing field `dt`"),
])
def test_config_errors(tmp_path, edit, msg):
    text = MINIMAL.replace(*edit, 1)
    with pytest.raises(host.HostError, match=msg):
        host.config(write(tmp_path, text))


def test_cfl_key_replaces_dt(tmp_path):
    c = host.config(write(tmp_path, MINIMAL.replace("dt = 0.001\n", "cfl = 0.3\n")))
    assert c["solver"]["cfl"] == 0.3 and c["solver"]["dt"] is None


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("compression", ["none", "gzip4", "blosc-lz4", "blosc-zstd"])
def test_state_round_trip(tmp_path, dtype, compression):
    nx, ny, nz = 3, 2, 5
    rng = np.random.default_rng(0)
    u = rng.standard_normal(5 * nx * ny * nz * 64).astype(dtype)
    path = host.save_state(12, 0.006, u, nx, ny, nz, 2 * np.pi, 2.0, np.pi, str(tmp_path / "o"), "flow", compression)
    kind = "f32" if dtype == np.float32 else "f64"
    assert os.path.basename(path) == f"flow_{kind}_000012.h5"
    h = host.read_header(path)
    assert (h["nx"], h["ny"], h["nz"], h["fr_order"], h["step"], h["u_real"]) == (nx, ny, nz, 3, 12, kind)
    assert h["t"] == 0.006 and h["lx"] == float(np.float32(2 * np.pi)) and h["lz"] == float(np.float32(np.pi))
    back, step, t = host.load_state(path, nx, ny, nz, kind)
    assert back.dtype == dtype and back.tobytes() == u.tobytes() and step == 12 and t == 0.006


def test_blosc_compresses(tmp_path):
    u = ic.channel_state(8, 8, 8, 2 * np.pi, 2.0, np.pi, 0.1)
    a = host.save_state(0, 0.0, u, 8, 8, 8, 2 * np.pi, 2.0, np.pi, str(tmp_path), "a", "none")
    b = host.save_state(0, 0.0, u, 8, 8, 8, 2 * np.pi, 2.0, np.pi, str(tmp_path), "b", "blosc-lz4")
    assert os.path.getsize(b) < 0.7 * os.path.getsize(a)


def test_loader_validation(tmp_path):
    u = np.ones(5 * 2 * 2 * 2 * 64, dtype=np.float32)
    path = host.save_state(0, 0.0, u, 2, 2, 2, 1.0, 1.0, 1.0, str(tmp_path), "flow", "none")
    with pytest.raises(host.HostError, match="does not match config 2x2x3"):
        host.load_state(path, 2, 2, 3, "f32")
    with pytest.raises(host.HostError, match="u_real 'f32' does not match the configured pipeline"):
        host.load_state(path, 2, 2, 2, "f64")
    with pytest.raises(host.HostError, match="initial_state file not found"):
        host.load_state(str(tmp_path / "absent.h5"), 2, 2, 2, "f32")


def test_make_initial_state(tmp_path):
    from opendynamo_amd import make_initial_state
    cfgp = write(tmp_path, CFG.format(out=tmp_path))
    out = make_initial_state.main([cfgp])
    assert os.path.basename(out) == "flow_f32_000000.h5"
    back, step, t = host.load_state(out, 3, 4, 2, "f32")
    c = host.config(cfgp)
    ref = make_initial_state.build_state(c)
    assert step == 0 and t == 0.0 and back.tobytes() == ref.tobytes()
