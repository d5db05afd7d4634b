"""The CLI's dump pipeline with save_freq = 1 and slow writes, on the CPU
under the thread sanitizer: the CLI's own writer pool
(opendynamo_amd/host/writer_pool.hpp), the snapshot slot pool
(opendynamo_amd/csrc/od_slots.hpp) and the real HDF5 writer at gzip level 9,
driven by tests/sanitize/io_pipeline_driver.cpp with a stand-in for the device
(a 16^3 f32 state rewritten every step; the loan-and-return rule of
solver.rs:163-175, 320, 816).  All three slots are out at once, so the step
loop waits for writers; every dump must hold exactly its step's state, with no
sanitizer report.

The check catches the failure it is for: with the slot pool's lend replaced by
the round-robin reuse of slot ticket % 3 that the CLI used to do, 9 of the 12
dumps come out holding another step's state (built without the sanitizer,
which spends minutes reporting the races of that build)."""
import os
import shutil
import subprocess

import pytest

from opendynamo_amd import build

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_cli_dump_pipeline_with_slow_writes_under_thread_sanitizer(tmp_path):
    hdf5 = os.path.join(build.HDF5_PREFIX, "lib", "libhdf5.so")
    blosc = os.path.join(build.HDF5_PREFIX, "lib", "libblosc.so")
    if shutil.which("g++") is None or not os.path.exists(hdf5) or not os.path.exists(blosc):
        pytest.skip("g++ or libhdf5/libblosc not available")
    exe = str(tmp_path / "io_pipeline_driver")
    # the system libstdc++ must win over the older one in the HDF5 prefix
    gcc_lib = subprocess.run(["g++", "-print-file-name=libstdc++.so"], capture_output=True, text=True).stdout.strip()
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fsanitize=thread", "-pthread", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(build.HDF5_PREFIX, "include"), "-o", exe,
           os.path.join(HERE, "sanitize", "io_pipeline_driver.cpp"), os.path.join(build.HOST, "state_io.cpp"),
           "-L" + os.path.dirname(gcc_lib), hdf5, blosc,
           "-Wl,-rpath,/usr/lib/x86_64-linux-gnu:" + os.path.join(build.HDF5_PREFIX, "lib")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    scratch = tmp_path / "dumps"
    scratch.mkdir()
    r = subprocess.run([exe, str(scratch)], capture_output=True, text=True, timeout=300)
    out = r.stdout + r.stderr
    if r.returncode != 0 and "unexpected memory mapping" in out:
        pytest.skip("the thread sanitizer's runtime cannot map its shadow memory on this kernel")
    assert r.returncode == 0, out[-4000:]
    assert "max slots out 3; done, 0 failures" in out, out[-2000:]
    assert "ThreadSanitizer" not in out, out[-4000:]
