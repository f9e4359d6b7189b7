"""Build libnuosc_hip.so (gfx950) and the nuosc command-line driver in-tree.

    python -m nuosc_amd.build          # or nuosc_amd.build.build()

hipcc compiles the HIP kernels for gfx950; the host-side C++ (C ABI, velocity
grid, NuOsc class, CLI) is compiled with g++ against the HIP runtime headers
with -ffp-contract=off, so that the setup arithmetic (velocity grid, initial
data) rounds exactly as the reference's x86-64 build does.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(HERE, "libnuosc_hip.so")
CLI = os.path.join(HERE, "bin", "nuosc")

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # no packed-FP32 instructions (v_pk_*_f32) anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]
HOST_FLAGS = [
    "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-result",
    "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include",
]

KERNEL_SRCS = ["nuosc_kernels.hip", "nuosc_flavor3.hip", "nuosc_noalloc.hip"]
HOST_SRCS = ["nuosc_capi.cpp", "vgrid.cpp", "nuosc_serve.cpp"]
CLI_SRCS = ["NuOsc.cpp", "nuosc_main.cpp"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True) -> str:
    """Compile everything; returns the path of libnuosc_hip.so."""
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in KERNEL_SRCS:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    cxx = shutil.which("g++") or "g++"
    for src in HOST_SRCS:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        _run([cxx, *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp,
          f"-Wl,-rpath,{ROCM}/lib", "-lpthread"])
    os.replace(tmp, LIB)
    # command-line driver linked against the library
    if all(os.path.exists(os.path.join(CSRC, s)) for s in CLI_SRCS):
        os.makedirs(os.path.dirname(CLI), exist_ok=True)
        cobjs = []
        for src in CLI_SRCS:
            obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
            _run([cxx, *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
            cobjs.append(obj)
        _run([cxx, *cobjs, "-o", CLI, f"-L{HERE}", "-lnuosc_hip", "-Wl,-rpath,$ORIGIN/..",
              f"-L{ROCM}/lib", "-lamdhip64", f"-Wl,-rpath,{ROCM}/lib"])
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
