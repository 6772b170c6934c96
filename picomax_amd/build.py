"""Builds libpmx.so (HIP, gfx950) and the `picomax` CLI in-tree.

    python -m picomax_amd.build          # or picomax_amd.build.build()

Outputs: picomax_amd/lib/libpmx.so and picomax_amd/bin/picomax (rpath to
../lib).  Every HIP source is compiled for gfx950 only, with packed-FP32
instruction selection disabled (the hipcc host pass warns that it does not
know the feature; that warning is expected).
"""
import concurrent.futures
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB_DIR = os.path.join(PKG, "lib")
BIN_DIR = os.path.join(PKG, "bin")
OBJ_DIR = os.path.join(ROOT, "build", "obj")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

LIB_SOURCES = ["pmx_input.cpp", "pmx_api.hip", "pmx_eig.hip", "pmx_chi.hip", "pmx_post.hip", "pmx_run.cpp"]
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-result", "-Wno-unused-function",
             "-I", os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n%s\n%s" % (" ".join(cmd), out))
    return out


def _compile(src):
    obj = os.path.join(OBJ_DIR, os.path.basename(src) + ".o")
    cmd = [HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-x", "hip"] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    _run(cmd)
    return obj


def build(verbose=False):
    """Compile libpmx.so and bin/picomax; returns the library path."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    os.makedirs(BIN_DIR, exist_ok=True)
    jobs = min(len(LIB_SOURCES), int(os.environ.get("MAX_JOBS", "8")))
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, LIB_SOURCES))
    lib = os.path.join(LIB_DIR, "libpmx.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, lib)
    exe = os.path.join(BIN_DIR, "picomax")
    _run([HIPCC, "-x", "hip", "--offload-arch=gfx950", "-O2", "-std=c++17",
          "-I", os.path.join(ROOT, "include"), os.path.join(CSRC, "picomax_main.cpp"),
          "-L", LIB_DIR, "-lpmx", "-Wl,-rpath,$ORIGIN/../lib", "-o", exe + ".tmp"])
    os.replace(exe + ".tmp", exe)
    if verbose:
        print("built", lib, "and", exe)
    return lib


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
