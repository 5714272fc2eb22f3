"""Build recipe for libflowsolver.so (gfx950) -- plain hipcc/g++ invocations.

    python flow-solver_amd/build.py [--jobs N]

Host setup (C++17) is compiled with g++, the device kernels and the C-ABI with
hipcc for gfx950.  The FP64 path is compiled with -ffp-contract=off so device
face fluxes keep the reference's rounding; packed FP32 instructions are
disabled for all device code.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libflowsolver.so")
INC = os.path.join(ROOT, "include")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-I" + INC,
]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-I" + INC]

HOST_SRCS = ["csrc/host/fs_setup.cpp", "csrc/host/fs_curved.cpp"]
HIP_SRCS = ["csrc/kernels/fs_kernels.hip", "csrc/kernels/fs_kernels_f32.hip", "csrc/capi/fs_engine.hip",
            "csrc/capi/fs_capi.hip", "csrc/capi/fs_batch.hip", "csrc/capi/fs_serve.hip", "csrc/capi/fs_norms.hip"]


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return r


def build(jobs=8, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs_list = []
    objs = []
    for src in HOST_SRCS:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        jobs_list.append(["g++"] + CXX_FLAGS + ["-c", os.path.join(HERE, src), "-o", obj])
        objs.append(obj)
    hc = hipcc()
    for src in HIP_SRCS:
        path = os.path.join(HERE, src)
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        jobs_list.append([hc] + HIP_FLAGS + ["-c", path, "-o", obj])
        objs.append(obj)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for r in ex.map(_run, jobs_list):
            if verbose and (r.stdout or r.stderr):
                sys.stderr.write(r.stdout + r.stderr)
    tmp = LIB + ".tmp"
    _run([hc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.v))
