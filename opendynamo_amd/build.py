"""Build the gfx950 HIP library and the C++ host tools in-tree.

    python -m opendynamo_amd.build           # everything
    python -m opendynamo_amd.build --lib     # libopendynamo_hip.so only

Outputs (git-ignored, shipped with the tree):
    opendynamo_amd/libopendynamo_hip.so   C ABI of include/opendynamo.h
    opendynamo_amd/libodhost.so           config / state-file / IC host layer
    opendynamo_amd/open_dynamo            CLI (reads ./config.toml)
"""
import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libopendynamo_hip.so")
HOSTLIB = os.path.join(PKG, "libodhost.so")
CLI = os.path.join(PKG, "open_dynamo")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# IEEE arithmetic with no contraction: the kernels reproduce the reference's
# operation order exactly.  Packed FP32 is disabled for correctness.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fno-gpu-rdc", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
HDF5_PREFIX = os.environ.get("OD_HDF5_PREFIX", "/opt/conda")
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-march=x86-64-v3", "-ffp-contract=off", "-pthread", "-Wall"]


def fast_flags(flags):
    """Flags of the fast-math translation unit: contraction allowed, except
    where the source narrows it with #pragma clang fp contract."""
    return [f for f in flags if f != "-ffp-contract=off"] + ["-ffp-contract=fast-honor-pragmas"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    return r


def build_lib():
    os.makedirs(OBJ, exist_ok=True)
    srcs = ["od_kernels.hip", "od_kernels_fast.hip", "od_api.hip", "od_batch.hip", "od_closure.hip"]
    objs = [os.path.join(OBJ, s.replace(".hip", ".o")) for s in srcs]
    jobs = []
    for s, o in zip(srcs, objs):
        flags = list(HIP_FLAGS)
        if s == "od_kernels_fast.hip":
            flags = fast_flags(flags)
        jobs.append([HIPCC, *flags, "-c", os.path.join(CSRC, s), "-o", o])
    with cf.ThreadPoolExecutor(len(jobs)) as ex:
        list(ex.map(_run, jobs))
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB + ".tmp", *objs])
    os.replace(LIB + ".tmp", LIB)
    return LIB


def build_host():
    srcs = [os.path.join(HOST, s) for s in sorted(os.listdir(HOST)) if s.endswith(".cpp") and s != "main.cpp"]
    if not srcs:
        return None
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(HDF5_PREFIX, "include")]
    lib = os.path.join(HDF5_PREFIX, "lib")
    # libhdf5 1.10 and libblosc (with lz4 / zstd) from the image's conda prefix.
    # The system libstdc++ must win over the older one in the conda prefix, at
    # link time (the compiler's own directory ahead of the prefix) and at run
    # time (the system directory first in the rpath).
    gcc_dir = os.path.dirname(subprocess.run(["g++", "-print-file-name=libstdc++.so"], capture_output=True,
                                             text=True, check=True).stdout.strip())
    link = ["-L" + gcc_dir, "-L" + lib, "-Wl,-rpath,/usr/lib/x86_64-linux-gnu:" + lib, "-lhdf5", "-lblosc"]
    _run(["g++", *CXX_FLAGS, "-shared", *inc, "-o", HOSTLIB + ".tmp", *srcs, *link])
    os.replace(HOSTLIB + ".tmp", HOSTLIB)
    main = os.path.join(HOST, "main.cpp")
    if os.path.exists(main):
        rocm = "/opt/rocm/lib"
        _run(["g++", *CXX_FLAGS, *inc, "-o", CLI + ".tmp", main, "-L" + PKG, "-lodhost", "-lopendynamo_hip",
              "-Wl,-rpath,$ORIGIN", "-L" + rocm, "-Wl,-rpath," + rocm, "-Wl,--allow-shlib-undefined", *link])
        os.replace(CLI + ".tmp", CLI)
    return HOSTLIB


def build_all():
    with cf.ThreadPoolExecutor(2) as ex:
        f_lib = ex.submit(build_lib)
        f_lib.result()
        f_host = ex.submit(build_host)
        f_host.result()


if __name__ == "__main__":
    if "--lib" in sys.argv:
        print(build_lib())
    else:
        build_all()
        print("built", LIB)
