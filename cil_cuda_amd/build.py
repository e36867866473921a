"""Build recipes: the gfx950 HIP library (in-tree .so) and the CPU oracle.

The HIP library is compiled with -ffp-contract=off so that FP64 paths keep the reference's
operation order bit for bit; FP32 fast paths fuse explicitly with fmaf().  Packed FP32
instructions are disabled (-packed-fp32-ops).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cil_cuda_amd")
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libcil_hip.so")
ORACLE = os.path.join(ROOT, "oracle")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd, cwd=None):
    res = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return res.stdout


def build_hip(verbose: bool = False) -> str:
    src = os.path.join(CSRC, "cil_hip.hip")
    tmp = LIB + ".tmp"
    out = _run([HIPCC, *HIP_FLAGS, "-o", tmp, src])
    os.replace(tmp, LIB)
    if verbose:
        print(out)
    return LIB


def build_oracle() -> None:
    make = shutil.which("make") or "make"
    _run([make, "-s", "-C", ORACLE, "all"])


def build_all(verbose: bool = False) -> None:
    build_oracle()
    build_hip(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", LIB)
