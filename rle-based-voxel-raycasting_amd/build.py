"""Builds librlevox.so (HIP kernels for gfx950 + host C++) in-tree.

    python rle-based-voxel-raycasting_amd/build.py [--jobs N]

Device code is compiled for strict IEEE fp32: no FMA contraction, no
fast-math, correctly rounded division and square root, denormals kept, and
packed fp32 instructions disabled.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librlevox.so")
ARCH = os.environ.get("RLEVOX_ARCH", "gfx950")

HIP_SOURCES = ["kernels.hip", "api.hip", "batch.hip", "server.hip", "tune.hip", "noalloc.hip"]
HOST_SOURCES = ["host.cpp", "raymap.cpp", "scene.cpp"]

FP_FLAGS = ["-ffp-contract=off", "-fno-fast-math"]
HIPCC_FLAGS = [
    "-c", "-fPIC", "-O3", "-std=c++17", "--offload-arch=" + ARCH,
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result", "-Wno-unused-value",
] + FP_FLAGS
CXX_FLAGS = ["-c", "-fPIC", "-O3", "-std=c++17", "-pthread", "-Wall",
             "-Wno-unused-function"] + FP_FLAGS


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(jobs=None, verbose=False):
    os.makedirs(OUT, exist_ok=True)
    hipcc = _hipcc()
    cxx = os.environ.get("CXX", "g++")
    cmds = []
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OUT, src + ".o")
        cmds.append([hipcc] + HIPCC_FLAGS + ["-x", "hip", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(OUT, src + ".o")
        cmds.append([cxx] + CXX_FLAGS + [os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    jobs = jobs or min(8, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-pthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    j = None
    if "--jobs" in sys.argv:
        j = int(sys.argv[sys.argv.index("--jobs") + 1])
    print(build(jobs=j, verbose=True))
