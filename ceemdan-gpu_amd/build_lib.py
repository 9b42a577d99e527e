"""Build libiceemdan_mi355x.so (HIP, gfx950) in-tree.

Every translation unit under csrc/ is compiled with hipcc for gfx950 and
linked into one shared library next to this file.  Flags that matter:
  -ffp-contract=off       no FMA contraction: FP64 results round exactly like
                          the CPU oracle
  -Xclang -target-feature -Xclang -packed-fp32-ops
                          no packed-FP32 instructions
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libiceemdan_mi355x.so")
EXAMPLES = os.path.join(HERE, "examples")
BIN = os.path.join(HERE, "bin")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

ARCH = "gfx950"
FLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function"]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in p.stderr.splitlines() if "packed-fp32-ops" not in l
                    and "is not a recognized feature" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{err}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_examples(verbose)
    return LIB


def build_examples(verbose: bool = False):
    """Host programs calling the library (linked with an rpath to it)."""
    os.makedirs(BIN, exist_ok=True)
    for f in sorted(os.listdir(EXAMPLES)):
        if not f.endswith(".cpp"):
            continue
        exe = os.path.join(BIN, f[:-4])
        cmd = [HIPCC, "-O2", "-std=c++17", "-Wall", os.path.join(EXAMPLES, f), "-o", exe + ".tmp",
               f"-L{HERE}", "-liceemdan_mi355x", "-Wl,-rpath,$ORIGIN/.."]
        if verbose:
            print(" ".join(cmd))
        p = subprocess.run(cmd, capture_output=True, text=True)
        if p.returncode != 0:
            raise RuntimeError(f"example build failed on {f}:\n{p.stderr}")
        os.replace(exe + ".tmp", exe)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
