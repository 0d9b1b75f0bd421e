"""Build libclothmi.so in-tree with hipcc for gfx950.

    python -m licenta_amd.build          # compile every source, link the library

All kernels are compiled with -ffp-contract=off (FP32 results reproducible
against the oracle) and with packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libclothmi.so")
SOURCES = ["k_prims.hip", "k_lbvh.hip", "k_narrow.hip", "k_pbd.hip", "cm_body.hip",
           "cm_noalloc.hip", "k_springs.hip", "k_response.hip"]
ARCH = os.environ.get("CM_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines()
                     if "packed-fp32-ops" not in l and l.strip())
    if warn and os.environ.get("CM_BUILD_VERBOSE"):
        print(warn, file=sys.stderr)
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = min(len(SOURCES), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    build_wrapper_demo()
    if verbose:
        print(f"built {OUT}")
    return OUT


DEMO_SRC = os.path.join(os.path.dirname(HERE), "tests", "cpp", "wrapper_demo.cpp")
DEMO = os.path.join(os.path.dirname(HERE), "tests", "cpp", "wrapper_demo")


def build_wrapper_demo() -> str:
    """The C++ wrapper's test driver (include/clothmi.hpp), a plain g++ program
    linked against the library."""
    rocm_lib = "/opt/rocm/lib"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", DEMO_SRC, "-o", DEMO, f"-L{HERE}", "-lclothmi",
           f"-L{rocm_lib}", "-lamdhip64", "-Wl,-rpath,$ORIGIN/../../licenta_amd",
           f"-Wl,-rpath,{rocm_lib}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"wrapper demo build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return DEMO


if __name__ == "__main__":
    build(verbose=True)
