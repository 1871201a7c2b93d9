"""Writes philox_rocrand.json: Philox4x32-10 blocks produced by rocRAND's
header-only engine (/opt/rocm/include/rocrand/rocrand_philox4x32_10.h, MIT),
compiled here as a host program from philox_rocrand_kat.cpp."""
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "kat")
        subprocess.run(["/opt/rocm/bin/hipcc", "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                        os.path.join(HERE, "philox_rocrand_kat.cpp"), "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    rows = []
    for line in out.split("\n"):
        if not line.strip():
            continue
        seed, sub, blk, a, b, c, e = line.split()
        seed, sub, blk = int(seed), int(sub), int(blk)
        rows.append({"key": [seed & 0xffffffff, seed >> 32],
                     "counter": [blk, 0, sub & 0xffffffff, sub >> 32],
                     "out": [int(a), int(b), int(c), int(e)]})
    with open(os.path.join(HERE, "philox_rocrand.json"), "w") as f:
        json.dump({"source": "rocRAND rocrand_philox4x32_10.h via philox_rocrand_kat.cpp", "blocks": rows}, f,
                  indent=1)


if __name__ == "__main__":
    main()
