"""Compile the native parts in-tree with hipcc for gfx950.

    python -m rtxrmq_amd.build          # librtxrmq.so + the rtxrmq CLI

Outputs (git-ignored, but they travel with the source tree):
    rtxrmq_amd/lib/librtxrmq.so   kernels + C ABI (include/rmq.h)
    rtxrmq_amd/bin/rtxrmq         the ./rtxrmq <n> <q> <lr> <alg> command line
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB_DIR = os.path.join(PKG, "lib")
BIN_DIR = os.path.join(PKG, "bin")
LIB = os.path.join(LIB_DIR, "librtxrmq.so")
CLI = os.path.join(BIN_DIR, "rtxrmq")

ARCH = os.environ.get("RTXRMQ_ARCH", "gfx950")

# -packed-fp32-ops: no v_pk_*_f32 in any kernel (the engine only compares).
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


# Host-only C++ (the CPU HRMQ of algs 1/101): g++, no device code.
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden", "-mpopcnt", "-Wall", f"-I{INCLUDE}"]


def build_lib() -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = [os.path.join(CSRC, f) for f in ("rmq.hip", "batch.hip") if os.path.exists(os.path.join(CSRC, f))]
    host_obj = os.path.join(LIB_DIR, "hrmq.o")
    _run(["g++", *HOST_FLAGS, "-c", "-o", host_obj, os.path.join(CSRC, "hrmq.cpp")])
    tmp = LIB + ".tmp"
    _run([hipcc(), *HIP_FLAGS, "-shared", "-o", tmp, *srcs, "-x", "none", host_obj, "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


def build_cli() -> str:
    os.makedirs(BIN_DIR, exist_ok=True)
    src = os.path.join(CSRC, "cli", "rtxrmq.cpp")
    tmp = CLI + ".tmp"
    _run([hipcc(), "-O2", "-std=c++17", "-ffp-contract=off", f"-I{INCLUDE}", "-o", tmp, src,
          f"-L{LIB_DIR}", "-lrtxrmq", f"-Wl,-rpath,$ORIGIN/../lib", "-lpthread"])
    os.replace(tmp, CLI)
    return CLI


def build_all() -> None:
    build_lib()
    if os.path.exists(os.path.join(CSRC, "cli", "rtxrmq.cpp")):
        build_cli()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
