"""Build the in-tree HIP library ``mlfma_amd/libmlfma_amd.so`` for gfx950 and the
``mlfma`` driver executable.

Objects are compiled in parallel with hipcc (device code) and g++ (host-only
code) into ``mlfma_amd/_build/``; the library lands next to this file so that
it travels with the source tree."""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmlfma_amd.so")
DRIVER = os.path.join(HERE, "mlfma")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall"]

DEVICE_SOURCES = ["kernels.hip", "kernels_mv.hip", "plan.hip", "solver.hip"]
HOST_SOURCES = ["host.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _obj(src):
    return os.path.join(OUT, os.path.splitext(src)[0] + ".o")


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = _obj(src)
    if src.endswith(".hip"):
        cmd = [HIPCC, *HIP_FLAGS, "-I", CSRC, "-c", path, "-o", obj]
    else:
        cmd = ["g++", *CXX_FLAGS, "-I", CSRC, "-c", path, "-o", obj]
    _run(cmd)
    return obj


def build(verbose: bool = True, jobs: int | None = None):
    os.makedirs(OUT, exist_ok=True)
    srcs = [s for s in DEVICE_SOURCES + HOST_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = jobs or min(8, len(srcs))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    drv = os.path.join(CSRC, "driver.cpp")
    if os.path.exists(drv):
        dobj = os.path.join(OUT, "driver.o")
        _run(["g++", *CXX_FLAGS, "-I", os.path.join(ROOT, "include"), "-c", drv, "-o", dobj])
        _run(["g++", "-o", DRIVER, dobj, "-L", HERE, "-lmlfma_amd", f"-Wl,-rpath,$ORIGIN"])
    if verbose:
        print("built", LIB)
    return LIB


def clean():
    shutil.rmtree(OUT, ignore_errors=True)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
