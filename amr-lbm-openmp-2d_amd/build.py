"""Builds libamrlbm.so (HIP kernels + runtime + C-ABI) for gfx950 in-tree.

    python amr-lbm-openmp-2d_amd/build.py

The library lands in amr-lbm-openmp-2d_amd/lib/libamrlbm.so. Device code is
compiled with -ffp-contract=off (no fused multiply-add contraction, so FP64
results follow the reference's operation order) and without packed FP32 ops.
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libamrlbm.so")
SOURCES = ["amrlbm_kernels.hip", "amrlbm_cnn.hip", "amrlbm_runtime.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("AMRLBM_ARCH", "gfx950")

DEVICE_FLAGS = ["-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(LIBDIR, src.replace(".hip", ".o"))
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-c", "-fopenmp",
               "-I", os.path.join(ROOT, "include"), "-I", CSRC] + DEVICE_FLAGS + \
              [os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        subprocess.check_call(cmd)
        objs.append(obj)
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-fopenmp", "-o", LIB + ".tmp"] + objs
    subprocess.check_call(cmd)
    os.replace(LIB + ".tmp", LIB)
    for o in objs:
        os.unlink(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
