"""Compile libh2w.so in-tree with hipcc for gfx950.

Every HIP source is built with -ffp-contract=off (kernel entries and ACA
residuals round like the reference's scalar code; the streaming loops use
explicit fma) and with packed-FP32 instructions disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libh2w.so")
SOURCES = ["plan.cpp", "kernels.hip", "h2w_ops.hip", "multi.hip", "batch.hip", "serialize.hip", "noalloc.hip", "complex.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("H2W_OFFLOAD_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-result",
         "-Wno-unused-function", "-I", os.path.join(HERE, "..", "include")]


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    msgs = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    return out, msgs


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(len(SOURCES)) as ex:
        res = list(ex.map(_compile, SOURCES))
    for _, m in res:
        if m and verbose:
            print(m, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + [o for o, _ in res]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # the command-line driver over include/h2w_initialisation.hpp (links libh2w in-tree)
    drv = os.path.join(HERE, "h2w_main")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", os.path.join(CSRC, "h2w_main.cpp"), "-I",
           os.path.join(HERE, "..", "include"), "-L", HERE, "-lh2w", "-Wl,-rpath,$ORIGIN", "-o", drv + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(drv + ".tmp", drv)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
