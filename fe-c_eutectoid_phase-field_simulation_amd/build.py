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


HDF5 = os.environ.get("PEARL_HDF5_PREFIX", "/opt/conda")
EXE = os.path.join(PKG, "pearl_gpu")
INPUT_LIB = os.path.join(PKG, "libpearl_input.so")


def build_input(verbose: bool = False) -> str:
    """Compile libpearl_input.so (C++17, host only): the drop-in program's
    input reader behind include/pearl_input.h, which the Python binding
    calls instead of parsing input files itself."""
    tmp = INPUT_LIB + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-fPIC", "-shared", f"-I{INCLUDE}",
           f"-I{os.path.join(CSRC, 'host')}", os.path.join(CSRC, "host", "input_capi.cpp"), "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, INPUT_LIB)
    return INPUT_LIB


def build_host(verbose: bool = False) -> str | None:
    """Compile the host program pearl_gpu (C++17) against libpearl_gpu.so and
    the image's HDF5.  Returns None when no HDF5 is available."""
    if not os.path.exists(os.path.join(HDF5, "include", "hdf5.h")):
        return None
    # The HIP runtime needs the system libstdc++, so the HDF5 library is
    # reached through a private directory holding only a link to it (an rpath
    # to the HDF5 prefix would also pick up that prefix's older libstdc++).
    hdir = os.path.join(PKG, "_hdf5")
    os.makedirs(hdir, exist_ok=True)
    real = os.path.realpath(os.path.join(HDF5, "lib", "libhdf5.so"))
    soname = "libhdf5.so.103"
    link = os.path.join(hdir, soname)
    if os.path.lexists(link):
        os.remove(link)
    os.symlink(real, link)
    tmp = EXE + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", f"-I{INCLUDE}", f"-I{HDF5}/include",
           os.path.join(CSRC, "host", "pearl_host.cpp"), "-o", tmp, f"-L{PKG}", "-lpearl_gpu",
           "-Wl,-rpath,$ORIGIN:$ORIGIN/_hdf5", link]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, EXE)
    return EXE


def build(verbose: bool = False) -> str:
    """Compile libpearl_gpu.so, libpearl_input.so and the host program
    (always recompiles)."""
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-o", tmp, *sources()]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    build_input(verbose)
    build_host(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print("built", LIB)
    sys.exit(0)
