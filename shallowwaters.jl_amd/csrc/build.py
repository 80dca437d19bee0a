"""Build recipe of libswm_hip.so (gfx950).

Strict IEEE arithmetic (no FMA contraction, no fast-math; HIP's default correctly rounded
float division and sqrt) keeps the kernels bitwise comparable to the CPU restatement.
Packed-FP32 instruction selection is disabled.

The library is five translation units: the C ABI (swm.hip) and one unit per number-format
bundle (swm_f64.hip, swm_f32.hip, swm_f16.hip, swm_bf16.hip) holding that bundle's kernels
and launches.  They compile in parallel and link into one shared object.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
OUT = os.path.join(PKG, "libswm_hip.so")
OBJ_DIR = os.path.join(HERE, "_build")
UNITS = ["swm.hip", "swm_f32.hip", "swm_f64.hip", "swm_f16.hip", "swm_bf16.hip", "swm_f16a.hip", "swm_bf16a.hip"]
SOURCES = [os.path.join(HERE, u) for u in UNITS]
HEADERS = ["swm_host.h", "swm_launch.h", "swm_common.h", "k_setup.h", "k_stage.h", "k_dissipation.h", "k_misc.h", "k_persist.h", "k_adjoint.h",
           "adj_core.h"]
DEPS = SOURCES + [os.path.join(HERE, h) for h in HEADERS] + [os.path.join(ROOT, "include", "swm.h")]

FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-ffp-contract=off", "-fno-fast-math",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    return r.returncode, r.stdout


def build_library(verbose: bool = False, jobs: int = 0) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJ_DIR, exist_ok=True)
    objs = [os.path.join(OBJ_DIR, os.path.splitext(u)[0] + ".o") for u in UNITS]
    cmds = [[hipcc, *FLAGS, "-I", os.path.join(ROOT, "include"), "-c", "-o", o, s] for s, o in zip(SOURCES, objs)]
    jobs = jobs or min(len(cmds), max(1, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda c: _run(c, verbose), cmds))
    for (rc, out), c in zip(results, cmds):
        if rc != 0:
            sys.stderr.write(out)
            raise RuntimeError(f"hipcc failed: {c[-1]}")
    rc, out = _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", OUT + ".tmp", *objs], verbose)
    if rc != 0:
        sys.stderr.write(out)
        raise RuntimeError("link failed")
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    build_library(verbose=True)
