"""Build libhydro_hip.so (HIP kernels + C ABI) in-tree for gfx950.

Every translation unit is compiled with hipcc in parallel into _build/, then
linked into hydrodynamicsgpu_amd/libhydro_hip.so.  Packed FP32 instructions
are disabled for all device code.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libhydro_hip.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

ARCH = os.environ.get("HYDRO_OFFLOAD_ARCH", "gfx950")
# -ffp-contract=on: a*b+c is fused where one expression says so, never across
# statements, so a cell's arithmetic does not depend on where the compiler
# inlined it: results are the same for every work decomposition (chunk
# height, strips per task) and for a grid alone or inside a batch.
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH, "-ffp-contract=on",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", CSRC, "-I", INCLUDE, "-Wno-unused-command-line-argument",
]

SOURCES = ["stage2d_f64.hip", "stage2d_f32.hip", "stage2d_f32k.hip", "stage3d.hip", "aux_kernels.hip",
           "batch_f64.hip", "batch_f32.hip", "batch_f32k.hip", "batch_aux.hip", "run1d.hip", "batch.cpp", "capi.cpp",
           "noalloc.hip", "burgers.hip", "gravity.hip"]


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c
    return "hipcc"


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [_hipcc()] + COMMON + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


HOST_SRC = os.path.join(HERE, "host", "src")
HOST_INC = os.path.join(HERE, "host", "include")
HOST_LIB = os.path.join(HERE, "libhydrogpu.so")


def build_host(verbose=False):
    """The C++ host layer (HydroGPU::Solver & co.) over the C ABI: plain g++,
    linked against libhydro_hip.so next to it."""
    srcs = sorted(os.path.join(HOST_SRC, f) for f in os.listdir(HOST_SRC) if f.endswith(".cpp"))
    tmp = HOST_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra", "-I", HOST_INC, "-I", INCLUDE,
           "-o", tmp] + srcs + ["-L", HERE, "-lhydro_hip", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("host build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, HOST_LIB)
    if verbose:
        print("built", HOST_LIB)
    return HOST_LIB


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = min(len(SOURCES), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_host(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
