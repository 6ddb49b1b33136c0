"""Builds hpgl_amd/hpgl.so and hpgl_amd/_cvariogram.so for gfx950 with
hipcc (no CMake, no JIT).

Usage: python hpgl_amd/build.py [--jobs N]
Objects go to hpgl_amd/_build/.
"""
import argparse
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "hpgl.so")
# the experimental-variogram module, a library of its own as in the reference
VARIOGRAM_OUT = os.path.join(HERE, "_cvariogram.so")
VARIOGRAM_SOURCES = ["variogram.hip"]

SOURCES = ["model.cpp", "propio.cpp", "solve.hip", "solve_quad.hip", "search.hip", "cluster.hip", "engine.hip", "api.hip", "sgs.hip", "server.hip", "cokriging.hip"]

# -ffp-contract=off: every product and sum rounds separately, as in the
# reference's FP64 code; kernels that want fused multiply-adds call fma()
# explicitly.  Packed FP32 instructions are disabled for the device.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden", "-Wno-unused-result",
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    path = os.path.join(SRC, src)
    if src.endswith(".cpp"):
        cmd = ["g++", "-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fvisibility=hidden", "-c", path, "-o", obj]
    else:
        cmd = ["hipcc"] + HIP_FLAGS + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, SOURCES + VARIOGRAM_SOURCES))
    n = len(SOURCES)
    for out, o in ((OUT, objs[:n]), (VARIOGRAM_OUT, objs[n:])):
        tmp = out + ".tmp"
        cmd = ["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + o
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
        os.replace(tmp, out)
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    a = ap.parse_args()
    print(build(a.jobs))
    sys.exit(0)
