"""Compile the HIP library libsphugo_amd.so for gfx950 (in-tree).

Every source is compiled with
  -ffp-contract=off   no product is fused into an add: the FP64 path
                      evaluates the reference's expressions with one
                      rounding per operation, as Go does on amd64;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions are generated.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libsphugo_amd.so")
SOURCES = ["sph_sort.hip", "sph_step.hip", "sph_api.hip", "sph_server.hip", "sph_noalloc.hip"]
BUILD_DIR = os.path.join(HERE, "_build")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I", INCLUDE, "-I", CSRC,
]


def hipcc() -> str:
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD_DIR, exist_ok=True)
    cc = hipcc()

    def compile_one(src):
        obj = os.path.join(BUILD_DIR, src.replace(".hip", ".o"))
        cmd = [cc, "-c", *FLAGS, os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
        if verbose and r.stderr.strip():
            print(r.stderr, file=sys.stderr)
        return obj

    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    build_demo()
    return LIB


DEMO = os.path.join(HERE, "sim_demo")


def build_demo() -> str:
    """The C++ mirror of package `sim` (csrc/sim.hpp) is header-only over the
    C ABI; sim_demo exercises it (host code only, g++)."""
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", INCLUDE, "-I", CSRC,
           os.path.join(CSRC, "sim_demo.cpp"), "-o", DEMO, "-L", HERE, "-lsphugo_amd",
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed on sim_demo.cpp:\n{r.stderr}")
    return DEMO


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
