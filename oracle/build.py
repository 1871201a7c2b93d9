"""ORACLE -- test infrastructure only.  Builds liboracle.so with g++.

Same floating-point rules as the device build: no contraction, no fast-math,
SSE arithmetic (x86-64 baseline, so the library also runs on other hosts).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
SOURCES = ["oracle.cpp"]

FLAGS = ["-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
         "-fno-math-errno", "-Wno-unknown-pragmas"]


LIB_OMP = os.path.join(HERE, "liboracle_omp.so")   # the same with its pure Jacobi passes over OpenMP threads
LIB_F64 = os.path.join(HERE, "liboracle_f64.so")   # the same steps in double precision
LIB_PLANT = os.path.join(HERE, "liboracle_plant.so")   # the float build with orc_plant(k): planted mistakes


def build(verbose=False):
    srcs = [os.path.join(HERE, s) for s in SOURCES]
    for lib, extra in ((LIB, []), (LIB_OMP, ["-fopenmp"]), (LIB_F64, ["-DORACLE_FP64"]),
                       (LIB_PLANT, ["-DORACLE_PLANTS"])):
        tmp = lib + ".tmp"
        cmd = ["g++"] + FLAGS + extra + srcs + ["-o", tmp]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(tmp, lib)
    return LIB


def build_native(out_dir, verbose=False):
    """The CPU-baseline builds for the host this runs on: the same source
    with -O3 -march=native (still no contraction and no fast-math, so the
    arithmetic is the parity build's), serial and OpenMP, into out_dir.
    Returns the two paths."""
    srcs = [os.path.join(HERE, s) for s in SOURCES]
    flags = [f for f in FLAGS if f != "-O2"] + ["-O3", "-march=native"]
    out = []
    for name, extra in (("liboracle_native.so", []), ("liboracle_native_omp.so", ["-fopenmp"])):
        lib = os.path.join(out_dir, name)
        cmd = ["g++"] + flags + extra + srcs + ["-o", lib]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        out.append(lib)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
