"""Build recipe of the native libraries (in-tree, no JIT cache).

* ``fluid_amd/libpbf_hip.so``: the HIP kernels and the C ABI of
  ``include/pbf.h``, compiled for gfx950 by hipcc.
* ``fluid_amd/host/pbf_driver``: the C++ host driver over the C ABI (the
  ``-p`` XML import and ``-d`` simulate-and-save path of the reference).

Run ``python -m fluid_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libpbf_hip.so")
DRIVER = os.path.join(PKG, "host", "pbf_driver")

HIP_SOURCES = ["pbf_kernels.hip", "pbf_capi.hip", "pbf_adjoint.hip"]

# IEEE single precision in source order on the device: no FMA contraction,
# correctly rounded division and square root.  Packed FP32 instructions are
# disabled for the whole build.
HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-fvisibility=hidden",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found: the HIP library cannot be built")
    return path


def _run(cmd: list[str], cwd: str) -> None:
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in proc.stdout.splitlines() if "packed-fp32-ops" not in line)
    if proc.returncode != 0:
        raise RuntimeError(f"build failed ({proc.returncode}): {' '.join(cmd)}\n{out}")


def build_hip(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd, CSRC)
    os.replace(tmp, LIB)
    return LIB


def build_driver(verbose: bool = False) -> str | None:
    src = os.path.join(PKG, "host", "pbf_driver.cpp")
    if not os.path.exists(src):
        return None
    cmd = [
        "g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), "-o", DRIVER, src,
        "-L", PKG, "-lpbf_hip", "-Wl,-rpath,$ORIGIN/..",
    ]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd, PKG)
    return DRIVER


def build_all(verbose: bool = False) -> None:
    build_hip(verbose)
    build_driver(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", LIB)
