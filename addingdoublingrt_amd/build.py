"""Compile the gfx950 HIP library in-tree (``libadrt_hip.so``).

Every HIP source is compiled with packed-FP32 instructions disabled; the
library is a plain shared object with the C-ABI of include/adrt_hip.h.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libadrt_hip.so")
ARCH = os.environ.get("ADRT_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++20", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]

SOURCES = ["adrt_capi.hip"]


def source_digest() -> str:
    """Short SHA-256 of the HIP sources (csrc/, sorted by name): stamps a
    measurement with the kernels it was taken on."""
    import hashlib

    h = hashlib.sha256()
    for name in sorted(os.listdir(CSRC)):
        if name.endswith((".hip", ".hpp")):
            h.update(name.encode())
            with open(os.path.join(CSRC, name), "rb") as f:
                h.update(f.read())
    return h.hexdigest()[:16]


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = OUT + ".tmp"
    cmd = [hipcc] + HIPCC_FLAGS + ["-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    subprocess.run(cmd, check=True, cwd=CSRC)
    os.replace(tmp, OUT)
    return OUT


ROOT = os.path.dirname(HERE)
CPP_CHECK_SRC = os.path.join(ROOT, "tests", "cpp", "adrt_cpp_check.cpp")
CPP_CHECK = os.path.join(ROOT, "tests", "cpp", "adrt_cpp_check")


def build_cpp_check(verbose: bool = False) -> str:
    """The test driver of the C++ host layer (include/adrt_hip.hpp), host code
    linked against libadrt_hip.so and the HIP runtime."""
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cmd = [os.environ.get("CXX", "g++"), "-std=c++17", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(rocm, "include"),
           CPP_CHECK_SRC, "-o", CPP_CHECK + ".tmp", "-L", HERE, "-ladrt_hip",
           "-L", os.path.join(rocm, "lib"), "-lamdhip64",
           "-Wl,-rpath,$ORIGIN/../../addingdoublingrt_amd", "-Wl,-rpath," + os.path.join(rocm, "lib")]
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    subprocess.run(cmd, check=True)
    os.replace(CPP_CHECK + ".tmp", CPP_CHECK)
    return CPP_CHECK


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_cpp_check(verbose=True))
