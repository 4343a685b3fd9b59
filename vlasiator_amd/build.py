"""Build libvacc.so (hipcc, gfx950) in-tree: vlasiator_amd/lib/libvacc.so."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libvacc.so")
SOURCES = ["vacc_kernels.hip", "vacc_init.hip", "vacc_trans.hip", "vacc_api.hip"]
CXX_SOURCES = ["vacc_vlsv.cpp", "vacc_pencils.cpp"]   # host-only C++ (no HIP)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("VACC_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: the remap reproduces the reference's per-operation float rounding,
# which fused multiply-adds would change.  Packed FP32 instructions are disabled.
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def build(verbose: bool = False) -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    objs = []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(LIB_DIR, s.replace(".hip", ".o"))
        cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    for s in CXX_SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(LIB_DIR, s.replace(".cpp", ".o"))
        cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-fPIC", "-Wall", "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    # example C++ host program over the adapter (vlasiator_amd/host)
    drv = os.path.join(LIB_DIR, "vacc_acc_driver")
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-I", os.path.join(HERE, "..", "include"),
           os.path.join(HERE, "host", "acc_driver.cpp"), "-o", drv + ".tmp", "-L", LIB_DIR, "-lvacc",
           "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(drv + ".tmp", drv)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
