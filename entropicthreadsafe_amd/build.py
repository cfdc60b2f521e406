"""Build the HIP library (libtslb.so) and the Fortran driver for gfx950.

The library is compiled in-tree with hipcc; the Fortran driver (when its
sources are present) with amdflang against the same library.
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libtslb.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build_library(verbose=False):
    srcs = [os.path.join(CSRC, f) for f in sorted(os.listdir(CSRC)) if f.endswith(".hip")]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I" + INCLUDE, "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        raise RuntimeError("hipcc failed building libtslb.so")
    os.replace(tmp, LIB)
    return LIB


def build_fortran(verbose=False):
    fdir = os.path.join(PKG, "fortran")
    mk = os.path.join(fdir, "Makefile")
    if not os.path.exists(mk):
        return None
    subprocess.run(["make", "-s", "-C", fdir], check=True)
    return fdir


def build_all(verbose=False):
    build_library(verbose)
    build_fortran(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
    print(LIB)
