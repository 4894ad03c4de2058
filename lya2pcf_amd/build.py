"""Build the lya2pcf HIP library in-tree: lya2pcf_amd/_build/liblya2pcf.so.

HIP sources are compiled for gfx950 only; host sources with g++ and OpenMP.
All FP64 paths are compiled without floating-point contraction so that the
reference formula is evaluated exactly (no fused multiply-add); explicit
fmaf() calls in the mixed-precision kernel are unaffected by that flag.
Packed FP32 instructions are disabled for all device code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "liblya2pcf.so")

HIP_SOURCES = ["xi.hip", "distortion.hip", "pairs.hip"]
HOST_SOURCES = ["host.cpp", "errors.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
             "-munsafe-fp-atomics", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include"), "-Wno-unused-result"]
HOST_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fopenmp", "-I" + os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + os.path.basename(cmd[-1]))
    return r


def build(verbose=False, extra_hip_flags=(), out_dir=None):
    """Compile and link; ``out_dir`` (default lya2pcf_amd/_build) receives the objects and the .so."""
    global OUT, LIB
    if out_dir:
        OUT, LIB = out_dir, os.path.join(out_dir, "liblya2pcf.so")
    os.makedirs(OUT, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OUT, src + ".o")
        _run([HIPCC, *HIP_FLAGS, *extra_hip_flags, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(OUT, src + ".o")
        _run(["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fopenmp", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
