"""Build recipe of libswm_hip.so (gfx950).

Strict IEEE arithmetic (no FMA contraction, no fast-math; HIP's default correctly rounded
float division and sqrt) keeps the kernels bitwise comparable to the CPU restatement.
Packed-FP32 instruction selection is disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
OUT = os.path.join(PKG, "libswm_hip.so")
SOURCES = [os.path.join(HERE, "swm.hip")]
DEPS = SOURCES + [os.path.join(HERE, h) for h in ("swm_common.h", "k_stage.h", "k_dissipation.h", "k_misc.h", "k_adjoint.h", "adj_core.h")] + [
    os.path.join(ROOT, "include", "swm.h")]

FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-ffp-contract=off", "-fno-fast-math",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument"]


def build_library(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [hipcc, *FLAGS, "-I", os.path.join(ROOT, "include"), "-o", OUT + ".tmp", *SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("hipcc failed")
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    build_library(verbose=True)
