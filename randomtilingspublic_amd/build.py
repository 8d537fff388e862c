"""Build librtiling.so (HIP kernels for gfx950 + C++ host) in-tree.

Usage: python -m randomtilingspublic_amd.build   (or build() from Python)
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "librtiling.so")

HIP_SOURCES = ["rt_device.hip", "rt_fused.hip", "rt_batch.hip", "rt_serve.hip", "rt_bibone.hip"]
CXX_SOURCES = ["rt_host.cpp", "rt_bibone_host.cpp", "rt_svg.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RT_OFFLOAD_ARCH", "gfx950")

# Exact IEEE fp32 semantics are part of the sampler's contract (weighted
# acceptance tests compare correctly rounded quotients), so: no fast math,
# no contraction, no packed-fp32 instructions.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-Wno-unused-command-line-argument",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
             "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} {cmd[-1]}")
    return r


def build(verbose=False):
    """Compile every source and link librtiling.so (always recompiles)."""
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(bdir, src + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", path, "-o", obj])
        objs.append(obj)
    for src in CXX_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(bdir, src + ".o")
        _run(["g++", *CXX_FLAGS, "-c", path, "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs,
          "-Wl,-soname,librtiling.so"])
    os.replace(tmp, LIB)
    shutil.rmtree(bdir, ignore_errors=True)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
