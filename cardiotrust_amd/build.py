"""Builds libcardiotrust_hip.so (gfx950) in-tree with hipcc.

Flags:
  -ffp-contract=off      no implicit multiply-add fusion: the kernels fuse
                         exactly where the reference calls mul_add;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                         no packed FP32 instructions in device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libcardiotrust_hip.so")
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("kernels.hip", "adjoint.hip", "gradcheck.hip",
                                                  "api.cpp", "batch.cpp", "tiles.cpp")]
ARCH = os.environ.get("CT_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = [
        hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
        "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc"),
        *[a for src in SOURCES for a in ("-x", "hip", src)], "-o", tmp,
    ]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
