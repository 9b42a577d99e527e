"""Build recipe for the FP64 CPU oracle (test infrastructure only).

Compiles oracle/iceemdan_oracle.c into oracle/liboracle.so with gcc.  The
flags matter for parity: -ffp-contract=off (no FMA contraction: the only
fused multiply-adds are the explicit fma() calls of the envelope evaluation,
which -mfma turns into single instructions) and no fast-math, so every
double operation rounds exactly as the HIP kernels do.

The reference implementation itself (cudaICEEMDAN.cu) needs nvcc, cuRAND,
cuSPARSE and CUB, none of which exist in this image, so no oracle/_ref build
is made.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "iceemdan_oracle.c")
OUT = os.path.join(HERE, "liboracle.so")

CFLAGS = ["-O2", "-fPIC", "-shared", "-std=c11", "-ffp-contract=off",
          "-fno-fast-math", "-mfma", "-fopenmp", "-Wall", "-Wno-unused-function"]


def build(verbose: bool = False) -> str:
    cmd = ["gcc", *CFLAGS, SRC, "-o", OUT + ".tmp", "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
