"""The HIP library loads without a GPU and exports every symbol that
include/opendynamo.h declares (no compute calls here)."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

import opendynamo_amd as od

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(odh?_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared("opendynamo.h")
    assert "od_step" in names and "od_create" in names
    lib = ctypes.CDLL(od.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(od.EXPORTS) == names


def test_default_params_match_reference_defaults():
    p = od.Params()
    od.lib().od_default_params(ctypes.byref(p))
    assert p.lx == float(ctypes.c_float(4.0 * 3.14159265358979323846).value) or abs(p.lx - 12.566370614359172) < 1e-6
    assert p.ly == 2.0
    assert abs(p.lz - 4.18879020478639) < 1e-6
    assert p.mu_ref == 0.00016 and p.t_ref == 273.15 and p.s_suth == 110.4
    assert p.body_force_x == 0.001
    assert p.flux_scheme == od.FLUX["slau2"] and p.y_wall == 1 and p.fast_math == 1
    assert p.u_real == p.q_real == p.rhs_real == p.compute_real == od.OD_F32


def test_params_struct_layout_matches_header():
    """od_params field order and types as declared in the header."""
    text = open(os.path.join(ROOT, "include", "opendynamo.h")).read()
    body = text[text.index("typedef struct od_params {"):text.index("} od_params;")]
    fields = []
    for line in body.splitlines()[1:]:
        line = line.split("/*")[0].strip().rstrip(";")
        if not line:
            continue
        typ, names = line.split(None, 1)
        fields += [(n.strip(), typ) for n in names.split(",")]
    ours = [(n, t) for n, t in od.Params._fields_]
    assert [f[0] for f in fields] == [f[0] for f in ours]
    cmap = {"int32_t": ctypes.c_int32, "double": ctypes.c_double}
    assert all(cmap[t] is c for (_, t), (_, c) in zip(fields, ours))


def test_oracle_library_loads():
    from oracle import oracle as O
    assert O.lib().orc_step is not None


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    monkeypatch.setattr(od, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(od, "_lib", None)
    with pytest.raises(od.ODError):
        od.lib()


def test_host_library_exports_every_declared_symbol():
    from opendynamo_amd import host
    names = [n for n in declared("odhost.h")]
    lib = ctypes.CDLL(host.HOSTLIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(host.EXPORTS) == sorted(names)


def test_batch_create_rejects_bad_arguments_before_touching_a_device():
    p = od.params(1, 1, 1, 1.0, 1.0, 1.0)
    h = ctypes.c_void_p()
    L = od.lib()
    assert L.od_batch_create(ctypes.byref(p), (od.Box * 1)(), 0, 0, ctypes.byref(h)) == 1
    assert b"at least one box" in L.od_last_error()
    bad = (od.Box * 1)(od.Box(0, 2, 2, 1.0, 1.0, 1.0))
    assert L.od_batch_create(ctypes.byref(p), bad, 1, 0, ctypes.byref(h)) == 1
    bad = (od.Box * 1)(od.Box(2, 2, 2, 1.0, -1.0, 1.0))
    assert L.od_batch_create(ctypes.byref(p), bad, 1, 0, ctypes.byref(h)) == 1
    assert not h.value
