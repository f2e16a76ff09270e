"""Builds the native libraries of omnioperator_amd in-tree.

  omnioperator_amd/lib/libomni_gpu.so   HIP kernels + the C ABI of include/omni_gpu.h (hipcc, gfx950)
                                        + the GPU-backed Operator classes (omnioperator_amd/host)
  omnioperator_amd/lib/omni_operator_test   C++ operator-level test driver (tests/cpp/operator_test.cpp)
  omnioperator_amd/lib/omni_operator_bench  end-to-end operator benchmark (omnioperator_amd/drivers/operator_bench.cpp)

Usage: python -m omnioperator_amd.build   (or build() from Python).  Every call recompiles.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_DIR = os.path.join(HERE, "lib")
OBJ_DIR = os.path.join(HERE, "lib", "obj")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-munsafe-fp-atomics",
    # packed FP32 instructions stay off (v_pk_add/mul/fma_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc"), "-I", os.path.join(HERE, "host"),
]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall", "-Wextra", "-Wno-unused-parameter",
             "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "host"),
             "-I", "/opt/rocm/include", "-D__HIP_PLATFORM_AMD__"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ_DIR, exist_ok=True)
    hip_srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
    cpp_srcs = sorted(glob.glob(os.path.join(HERE, "host", "*.cpp")))
    jobs = []
    for s in hip_srcs:
        o = os.path.join(OBJ_DIR, os.path.basename(s) + ".o")
        jobs.append(([HIPCC] + HIP_FLAGS + ["-c", s, "-o", o], o))
    for s in cpp_srcs:
        o = os.path.join(OBJ_DIR, os.path.basename(s) + ".o")
        jobs.append(([CXX] + CXX_FLAGS + ["-c", s, "-o", o], o))
    workers = min(8, max(1, len(jobs)))
    with cf.ThreadPoolExecutor(workers) as ex:
        outs = list(ex.map(lambda j: _run(j[0]), jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    lib = os.path.join(LIB_DIR, "libomni_gpu.so")
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + [j[1] for j in jobs] +
         ["-Wl,-soname,libomni_gpu.so"])
    os.replace(tmp, lib)
    # executables over the operator surface: the C++ operator tests and the end-to-end operator benchmark
    for src, name in ((os.path.join(ROOT, "tests", "cpp", "operator_test.cpp"), "omni_operator_test"),
                      (os.path.join(HERE, "drivers", "operator_bench.cpp"), "omni_operator_bench")):
        if os.path.exists(src):
            _run([CXX] + CXX_FLAGS + ["-O3", src, "-o", os.path.join(LIB_DIR, name), "-L", LIB_DIR, "-lomni_gpu",
                                      "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-lamdhip64",
                                      "-Wl,-rpath,/opt/rocm/lib"])
    return lib


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
