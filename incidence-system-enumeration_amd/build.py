"""Builds the HIP library libise_amd.so in-tree for gfx950.

Every source under csrc/ is compiled by hipcc into one shared library that
exports the C ABI of include/ise.h.  No CPU fallback is built: the library
only runs on a gfx950 device.
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libise_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = ["ise_kernels.hip", "ise_kfact.hip", "ise_mlp.hip", "ise_runtime.cpp", "ise_enum.cpp",
           "ise_kfact_host.cpp", "ise_mlp_host.cpp", "ise_group.hip",
           "ise_format.cpp", "ise_solver.hip", "ise_noalloc.hip",
           "ise_canon.hip", "ise_solution.cpp",
           "ise_kf_backjump.hip", "ise_random.hip", "ise_cdtools.hip"]
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def build(verbose=False):
    objs = []
    os.makedirs(os.path.join(PKG, "_build"), exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(PKG, "_build", src + ".o")
        cmd = [HIPCC] + [f for f in FLAGS if f != "-shared"] + [
            "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj,
            "-I", os.path.join(ROOT, "include"), "-I", CSRC]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
