"""Build recipe of the native parts: the HIP solver library (gfx950) and the CLI app.

``python -m fastreg_amd.build`` (or ``build()``) compiles, in-tree:

* ``fastreg_amd/libfastreg_hip.so`` — kernels + C-ABI (include/fastreg_c.h), hipcc for gfx950;
* ``fastreg_amd/fastreg_app`` — the params.ini-driven CLI, g++ against the C-ABI library;
* ``tests/cpp/reference_caller`` — a test program written against the reference's C++ classes
  (include/fastreg.hpp, include/solver.hpp), g++ against the same library.

HIP flags: -ffp-contract=off (multiply-adds round twice unless written fmaf, matching the
oracle's operation order), correctly rounded f32 division/sqrt, denormals kept, and packed-FP32
instruction selection disabled.
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
LIB = os.path.join(PKG, "libfastreg_hip.so")
APP = os.path.join(PKG, "fastreg_app")

HIP_SOURCES = ["fr_kernels.hip", "fr_capi.hip", "fr_adjoint.hip"]
APP_SOURCES = ["app.cpp", "nifti.cpp"]
CALLER_SRC = os.path.join(ROOT, "tests", "cpp", "reference_caller.cpp")
CALLER = os.path.join(ROOT, "tests", "cpp", "reference_caller")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return r.stdout


def build_lib(verbose: bool = False) -> str:
    objs = []
    build_dir = os.path.join(PKG, "_build")
    os.makedirs(build_dir, exist_ok=True)
    for src in HIP_SOURCES:
        obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
        out = _run([HIPCC, *HIP_FLAGS, "-I", INCLUDE, "-c", os.path.join(CSRC, src), "-o", obj])
        if verbose and out.strip():
            print(out)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_app() -> str | None:
    srcs = [os.path.join(CSRC, s) for s in APP_SOURCES]
    if not all(os.path.exists(s) for s in srcs):
        return None
    cxx = shutil.which("g++") or "g++"
    tmp = APP + ".tmp"
    _run([cxx, "-O2", "-std=c++17", "-Wall", "-I", INCLUDE, "-I", CSRC, *srcs, "-o", tmp,
          "-L", PKG, "-lfastreg_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, APP)
    return APP


def build_caller() -> str | None:
    if not os.path.exists(CALLER_SRC):
        return None
    cxx = shutil.which("g++") or "g++"
    tmp = CALLER + ".tmp"
    _run([cxx, "-O2", "-std=c++17", "-Wall", "-I", INCLUDE, CALLER_SRC, "-o", tmp, "-L", PKG, "-lfastreg_hip",
          "-Wl,-rpath," + os.path.relpath(PKG, os.path.dirname(CALLER)).join(["$ORIGIN/", ""])])
    os.replace(tmp, CALLER)
    return CALLER


def build(verbose: bool = False) -> None:
    build_lib(verbose)
    build_app()
    build_caller()


if __name__ == "__main__":
    build(verbose=True)
    print("built", LIB)
