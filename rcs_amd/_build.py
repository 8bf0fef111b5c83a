"""Build librcs_amd.so for gfx950 with hipcc (one object per source, then one link)."""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librcs_amd.so")
SOURCES = ["entries.hip", "dense.hip", "rrqr.hip", "aca.hip", "mlp.hip", "tree.cpp", "refine.cpp", "capi.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # no packed FP32 instructions: some of their forms have given wrong results on this
    # hardware without any error
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
ROOT = os.path.dirname(HERE)
DRIVER_SRC = os.path.join(ROOT, "tests", "cpp", "driver_sphere.cpp")
DRIVER_BIN = os.path.join(ROOT, "tests", "cpp", "driver_sphere")


def build_driver():
    """The C++ driver test over include/rcs_amd.hpp (host code only, g++)."""
    cmd = ["g++", "-std=c++17", "-O2", "-I" + os.path.join(ROOT, "include"), DRIVER_SRC, "-L" + HERE, "-lrcs_amd",
           "-Wl,-rpath,$ORIGIN/../../rcs_amd", "-o", DRIVER_BIN]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("driver build failed:\n" + r.stderr)
    return DRIVER_BIN


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, LIB)
    build_driver()
    if verbose:
        print("built", LIB, "and", DRIVER_BIN)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
