"""Build recipe for the HIP library libpearl_gpu.so (gfx950).

Compiled in-tree with hipcc so that the shared object travels with the
source tree.  Flags:
  -ffp-contract=off   no fused multiply-add unless the source calls fma, so
                      the FP64 plain form rounds like the reference
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions are generated
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
LIB = os.path.join(PKG, "libpearl_gpu.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PEARL_OFFLOAD_ARCH", "gfx950")

FLAGS = ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
         "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-result"]


def sources():
    return [os.path.join(CSRC, "pearl_gpu.hip")]


def build(verbose: bool = False) -> str:
    """Compile libpearl_gpu.so (always recompiles)."""
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-o", tmp, *sources()]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print("built", LIB)
    sys.exit(0)
