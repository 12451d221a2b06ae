"""Builds libcpsfdtd.so (host model in C++, kernels in HIP for gfx950) in-tree.

    python -m cpsfdtd_amd.build

Host set-up code (parsers, painters, coefficient tables) is compiled with g++
against glibc's libm, exactly like the reference, so FP64 tables match it
bit for bit; kernels go through hipcc for gfx950.  Both sides are compiled
without floating-point contraction.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libcpsfdtd.so")
OBJ = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
HIP_FLAGS = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]

HOST_SOURCES = ["model.cpp", "model_capi.cpp", "server.cpp", "image.cpp"]
HIP_SOURCES = ["engine.hip"]
HEADERS = ["model.hpp", "engine.hpp", "kernels.hip"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=False, force=True):
    """Compile every source; with force=False skip when the library is newer than all sources."""
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    srcs = [os.path.join(CSRC, f) for f in HOST_SOURCES + HIP_SOURCES + HEADERS]
    srcs.append(os.path.join(ROOT, "include", "cpsfdtd.h"))
    srcs.append(os.path.join(ROOT, "include", "cpsfdtd.hpp"))
    if not force and os.path.exists(LIB) and all(os.path.getmtime(LIB) >= os.path.getmtime(s) for s in srcs):
        return LIB
    objs = []
    for f in HOST_SOURCES:
        o = os.path.join(OBJ, f + ".o")
        _run(["g++"] + COMMON + ["-Wall", "-pthread", "-c", os.path.join(CSRC, f), "-o", o], verbose)
        objs.append(o)
    for f in HIP_SOURCES:
        o = os.path.join(OBJ, f + ".o")
        _run([HIPCC] + COMMON + HIP_FLAGS + ["-c", os.path.join(CSRC, f), "-o", o], verbose)
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-pthread", "-o", tmp] + objs + ["-lz"], verbose)
    os.replace(tmp, LIB)
    build_facade_driver(verbose)
    return LIB


FACADE_DRIVER_SRC = os.path.join(ROOT, "tests", "facade", "main_loop.cpp")
FACADE_DRIVER = os.path.join(ROOT, "tests", "facade", "_bin", "main_loop")


def build_facade_driver(verbose=False):
    """The main.cpp-shaped driver written against include/cpsfdtd.hpp, linked
    to the in-tree library (used by the facade tests)."""
    if not os.path.exists(FACADE_DRIVER_SRC):
        return None
    os.makedirs(os.path.dirname(FACADE_DRIVER), exist_ok=True)
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"), FACADE_DRIVER_SRC,
          "-o", FACADE_DRIVER, "-L" + LIBDIR, "-lcpsfdtd", "-Wl,-rpath,$ORIGIN/../../../cpsfdtd_amd/lib"], verbose)
    return FACADE_DRIVER


SAN_LIB = os.path.join(LIBDIR, "sanitize", "libcpsfdtd.so")
SANITIZE = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined"]


def build_sanitized(verbose=False):
    """The same library with its host code (model, C ABI of the model,
    server, image export) built with AddressSanitizer and
    UndefinedBehaviorSanitizer by g++, for CPU test runs; the device code is
    the ordinary build.  Load it with gcc's libasan and libubsan preloaded
    (tests/run_sanitized.sh)."""
    out = os.path.dirname(SAN_LIB)
    os.makedirs(out, exist_ok=True)
    objs = []
    for f in HOST_SOURCES:
        o = os.path.join(OBJ, f + ".san.o")
        _run(["g++"] + COMMON + SANITIZE + ["-O1", "-Wall", "-pthread", "-c", os.path.join(CSRC, f), "-o", o], verbose)
        objs.append(o)
    hip_o = os.path.join(OBJ, "engine.hip.o")
    if not os.path.exists(hip_o):
        build(verbose)
    objs.append(hip_o)
    _run(["g++", "-shared", "-fPIC", "-pthread"] + SANITIZE + ["-o", SAN_LIB] + objs +
         ["-L/opt/rocm/lib", "-lamdhip64", "-lz", "-Wl,-rpath,/opt/rocm/lib"], verbose)
    return SAN_LIB


if __name__ == "__main__":
    if "--sanitize" in sys.argv:
        print(build_sanitized(verbose="-q" not in sys.argv))
    else:
        build(verbose="-q" not in sys.argv)
        print(LIB)
