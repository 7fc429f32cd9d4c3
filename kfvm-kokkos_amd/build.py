"""Build the stage library in-tree.

Produces kfvm-kokkos_amd/build/libkfvm_amd.so (HIP kernels for gfx950 plus the
C++17 host driver behind the C ABI of include/kfvm.h), the command-line
driver kfvm-kokkos_amd/build/kfvm.ex, the FP64 throughput probe
kfvm-kokkos_amd/build/fp64_peak and the division / square-root check
kfvm-kokkos_amd/build/div_sqrt_check.

    python kfvm-kokkos_amd/build.py [--force]
"""
import argparse
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "build")
LIB = os.path.join(BUILD, "libkfvm_amd.so")
EXE = os.path.join(BUILD, "kfvm.ex")
PROBE = os.path.join(BUILD, "fp64_peak")
DIV_SQRT_CHECK = os.path.join(BUILD, "div_sqrt_check")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"
# Contraction off: only the explicit fma() calls fuse, so the device results
# follow the same operation order as the CPU oracle.  Packed FP32 disabled.
HIP_FLAGS = ["-x", "hip", f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
             "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-pass-failed"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall"]

HOST_CXX = ["common.cpp", "weights.cpp", "problems.cpp", "writer.cpp"]
HOST_HIP = ["solver.cpp", "indicator.cpp", "capi.cpp"]
DEVICE = ["dispatch.hip", "inst_d2q2.hip", "inst_d2q3.hip", "inst_d3q2.hip", "inst_d3q3.hip", "noalloc.hip"]


def _headers():
    out = []
    for d in (os.path.join(CSRC, "host"), os.path.join(CSRC, "device"),
              os.path.join(HERE, "..", "include")):
        for f in os.listdir(d):
            if f.endswith((".hpp", ".h")):
                out.append(os.path.join(d, f))
    return out


def _stale(obj, src, deps):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(p) > t for p in [src] + deps)


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n  " + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(force=False, verbose=False):
    """Compile what changed (everything with force=True) and link."""
    os.makedirs(BUILD, exist_ok=True)
    deps = _headers()
    jobs = []
    objs = []
    for f in HOST_CXX:
        src, obj = os.path.join(CSRC, "host", f), os.path.join(BUILD, f + ".o")
        objs.append(obj)
        if force or _stale(obj, src, deps):
            jobs.append([CXX] + CXX_FLAGS + ["-c", src, "-o", obj])
    for f in HOST_HIP:
        src, obj = os.path.join(CSRC, "host", f), os.path.join(BUILD, f + ".o")
        objs.append(obj)
        if force or _stale(obj, src, deps):
            jobs.append([HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj])
    for f in DEVICE:
        src, obj = os.path.join(CSRC, "device", f), os.path.join(BUILD, f + ".o")
        objs.append(obj)
        if force or _stale(obj, src, deps):
            jobs.append([HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj])
    workers = max(1, min(len(jobs), os.cpu_count() or 4, 8))
    if jobs:
        with concurrent.futures.ThreadPoolExecutor(workers) as ex:
            for out in ex.map(_run, jobs):
                if verbose and out.strip():
                    print(out)
    relink = force or jobs or not os.path.exists(LIB)
    if relink:
        tmp = LIB + ".tmp"
        _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs +
             ["-lquadmath", "-ldl", "-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx", "-Wl,-rpath,/opt/rocm/lib"])
        os.replace(tmp, LIB)
    main_src = os.path.join(CSRC, "host", "main.cpp")
    if relink or _stale(EXE, main_src, deps):
        _run([CXX, "-O2", "-std=c++17", "-o", EXE, main_src, LIB, "-Wl,-rpath,$ORIGIN"])
    # FP64 vector-FMA vs matrix-FMA throughput probe (run it under rocprofv3)
    probe_src = os.path.join(CSRC, "probe", "fp64_peak.hip")
    if force or _stale(PROBE, probe_src, []):
        _run([HIPCC] + [f for f in HIP_FLAGS if f not in ("-x", "hip", "-fPIC")] + ["-o", PROBE, probe_src])
    # the Riemann solvers' division and square root against IEEE (tests/test_gpu_div_sqrt.py)
    check_src = os.path.join(CSRC, "probe", "div_sqrt_check.hip")
    if force or _stale(DIV_SQRT_CHECK, check_src, [os.path.join(CSRC, "device", "physics.hpp")]):
        _run([HIPCC] + [f for f in HIP_FLAGS if f not in ("-x", "hip", "-fPIC")] + ["-o", DIV_SQRT_CHECK, check_src])
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(force=a.force, verbose=a.verbose))
