"""Builds the HIP library hsspack_amd/libhsdc.so for gfx950 (in-tree)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libhsdc.so")
SOURCES = ["hsdc_kern.hip", "hsdc_hss.hip", "hsdc_mlp.hip", "hsdc_api.hip", "hsdc_svd.hip", "hsdc_noalloc.hip",
           "hsdc_band.hip", "hsdc_mat2hss.hip"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]


FLANG = os.environ.get("FLANG", "/opt/rocm/llvm/bin/flang")
FSRC = os.path.join(HERE, "fortran", "hsdc_mdstedc.f90")
FLIB = os.path.join(HERE, "libhsdc_f.so")
FDRIVE_SRC = os.path.join(os.path.dirname(HERE), "tests", "fortran", "fdrive.f90")
FDRIVE = os.path.join(os.path.dirname(HERE), "tests", "fortran", "fdrive")
BUILD_DIR = os.path.join(HERE, "_build")


def _run(cmd, verbose, cwd=None):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=cwd)


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    _run([hipcc] + FLAGS + ["-I", CSRC, "-o", tmp] + srcs, verbose)
    os.replace(tmp, LIB)
    return LIB


def build_fortran(verbose=False):
    """The Fortran drop-ins (libhsdc_f.so, linked to libhsdc.so) and their test driver."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    tmp = FLIB + ".tmp"
    _run([FLANG, "-O2", "-fPIC", "-shared", "-module-dir", BUILD_DIR, "-o", tmp, FSRC, "-L" + HERE, "-lhsdc",
          "-Wl,-rpath,$ORIGIN"], verbose, cwd=BUILD_DIR)
    os.replace(tmp, FLIB)
    tmp = FDRIVE + ".tmp"
    _run([FLANG, "-O2", "-module-dir", BUILD_DIR, "-o", tmp, FDRIVE_SRC, "-L" + HERE, "-lhsdc_f", "-lhsdc",
          "-Wl,-rpath," + HERE, "-Wl,-rpath,$ORIGIN/../../hsspack_amd"], verbose, cwd=BUILD_DIR)
    os.replace(tmp, FDRIVE)
    return FLIB


if __name__ == "__main__":
    build(verbose=True)
    build_fortran(verbose=True)
    print("built", LIB, FLIB)
