"""Compile libibis_gpu.so (gfx950) in-tree.

Flags: no floating-point contraction (the kernels evaluate the reference's
expression trees operation by operation), and no packed FP32 instructions.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libibis_gpu.so")
SOURCES = ["status.cpp", "grid.cpp", "residual.hip", "jfnk.hip", "batch.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc() -> str:
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c
    return "hipcc"


HOST = os.path.join(HERE, "host")
HOST_LIB = os.path.join(LIBDIR, "libibis_host.so")
FV_RUN = os.path.join(LIBDIR, "ibis_fv_run")
# the C++ host layer (FiniteVolume / RungeKutta over the C ABI) and its
# command-line driver: host code only, linked against libibis_gpu.so
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB + ".tmp"
    _run([hipcc(), *HIPCC_FLAGS, "-I", os.path.join(HERE, "..", "include"),
          *[os.path.join(CSRC, s) for s in SOURCES], "-o", tmp], verbose)
    os.replace(tmp, LIB)
    build_host(verbose)
    return LIB


def build_host(verbose: bool = False) -> str:
    """libibis_host.so and ibis_fv_run, linked against libibis_gpu.so (found
    next to them at run time)."""
    tmp = HOST_LIB + ".tmp"
    _run([hipcc(), *HOST_FLAGS, "-shared", os.path.join(HOST, "json.cpp"),
          os.path.join(HOST, "ibis_fv.cpp"), "-L", LIBDIR, "-libis_gpu",
          "-Wl,-rpath,$ORIGIN", "-o", tmp], verbose)
    os.replace(tmp, HOST_LIB)
    tmp = FV_RUN + ".tmp"
    _run([hipcc(), *HOST_FLAGS, os.path.join(HOST, "fv_run.cpp"), "-L", LIBDIR, "-libis_host",
          "-libis_gpu", "-Wl,-rpath,$ORIGIN", "-o", tmp], verbose)
    os.replace(tmp, FV_RUN)
    build_python_api(verbose)
    return HOST_LIB


def python_api_path() -> str:
    import sysconfig
    return os.path.join(LIBDIR, "python_api" + sysconfig.get_config_var("EXT_SUFFIX"))


def build_python_api(verbose: bool = False) -> str:
    """The pybind11 module python_api (ibis_amd/host/python_api.cpp), linked
    against libibis_host.so and libibis_gpu.so next to it."""
    import sysconfig
    import pybind11
    out = python_api_path()
    tmp = out + ".tmp"
    _run([hipcc(), *HOST_FLAGS, "-shared", "-fvisibility=hidden", "-I", pybind11.get_include(),
          "-I", sysconfig.get_paths()["include"], os.path.join(HOST, "python_api.cpp"),
          "-L", LIBDIR, "-libis_host", "-libis_gpu", "-Wl,-rpath,$ORIGIN", "-o", tmp], verbose)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
