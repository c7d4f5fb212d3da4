"""Builds the gfx950 HIP library ``libtomograpy_amd.so`` in-tree with hipcc.

The library is compiled with ``-ffp-contract=off`` so that ``a * b + c`` is two
roundings, as in the reference's x86-64 build of tomograpy/C_siddon.c.template,
which the forward projector reproduces bit for bit.  Packed FP32 instructions
are disabled (``-packed-fp32-ops``).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libtomograpy_amd.so")
SOURCES = ["siddon.hip", "siddon_det.hip", "siddon_cg.hip", "siddon_batch.hip", "siddon_plan.hip",
           "siddon_static.hip"]
ARCH = os.environ.get("TOMOGRAPY_AMD_ARCH", "gfx950")


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return os.path.join(rocm, "bin", "hipcc")


def flags():
    return [
        "--offload-arch=" + ARCH,
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
        "-I", INCLUDE,
        "-I", CSRC,
    ]


def build(verbose=False):
    """Compiles every source to an object and links the shared library."""
    objs = []
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(build_dir, src.replace(".hip", ".o"))
        cmd = [hipcc()] + flags() + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
