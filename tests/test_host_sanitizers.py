"""The host layer's C++ (config.toml parser, HDF5 state files with gzip and
blosc) built with AddressSanitizer, UndefinedBehaviorSanitizer and
LeakSanitizer into a standalone driver (tests/sanitize/host_driver.cpp) and
run on the CPU: every round trip and error path must pass with no sanitizer
report.  Host code only; no GPU code is involved."""
import os
import shutil
import subprocess

import pytest

from opendynamo_amd import build

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_host_layer_clean_under_asan_ubsan(tmp_path):
    hdf5 = os.path.join(build.HDF5_PREFIX, "lib", "libhdf5.so")
    blosc = os.path.join(build.HDF5_PREFIX, "lib", "libblosc.so")
    if shutil.which("g++") is None or not os.path.exists(hdf5) or not os.path.exists(blosc):
        pytest.skip("g++ or libhdf5/libblosc not available")
    exe = str(tmp_path / "host_driver")
    srcs = [os.path.join(build.HOST, s) for s in ("config.cpp", "odhost_capi.cpp", "state_io.cpp")]
    # the system libstdc++ must win over the older one in the HDF5 prefix
    gcc_lib = subprocess.run(["g++", "-print-file-name=libstdc++.so"], capture_output=True, text=True).stdout.strip()
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-fno-omit-frame-pointer", "-pthread", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(build.HDF5_PREFIX, "include"), "-o", exe,
           os.path.join(HERE, "sanitize", "host_driver.cpp"), *srcs, "-L" + os.path.dirname(gcc_lib), hdf5, blosc,
           "-Wl,-rpath,/usr/lib/x86_64-linux-gnu:" + os.path.join(build.HDF5_PREFIX, "lib")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    scratch = tmp_path / "io"
    scratch.mkdir()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, str(scratch)], capture_output=True, text=True, env=env, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "done, 0 failures" in out
    assert "Sanitizer" not in out and "runtime error" not in out, out[-4000:]
