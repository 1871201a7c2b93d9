"""Builds libspatial_cell.so (gfx950) in place with hipcc.

Flags that matter for results: -ffp-contract=off (no fused multiply-add, so
device arithmetic rounds exactly like the host oracle), no fast-math, and
packed-fp32 instructions disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libspatial_cell.so")
SOURCES = ["sc_api.hip", "sc_kernels.hip", "sc_bin.hip", "sc_setup.cpp", "sc_json.cpp", "sc_hostapi.cpp", "sc_server.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unknown-pragmas", "-Wno-unused-result", "-Wno-unused-value",
]


# spatial_cell_sim: the command-line driver (host code only) over the library
DRIVER = os.path.join(HERE, "spatial_cell_sim")
DRIVER_SOURCES = ["sc_main.cpp", "sc_json.cpp"]
CXX = os.environ.get("CXX", "g++")
ROCM_LIB = "/opt/rocm/lib"


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    dcmd = [CXX, "-O2", "-std=c++17", "-pthread"] + [os.path.join(CSRC, s) for s in DRIVER_SOURCES] + [
        "-L" + HERE, "-lspatial_cell", "-Wl,-rpath,$ORIGIN", "-L" + ROCM_LIB, "-Wl,-rpath-link," + ROCM_LIB,
        "-o", DRIVER + ".tmp"]
    if verbose:
        print(" ".join(dcmd), flush=True)
    subprocess.run(dcmd, check=True)
    os.replace(DRIVER + ".tmp", DRIVER)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
