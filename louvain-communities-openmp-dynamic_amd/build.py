"""Build the native libraries in place (hipcc for gfx950, g++ for host code).

    liblouvain_hip.so  csrc/lv_kernels.hip + lv_update.hip + lv_capi.hip + lv_serve.hip   (HIP, gfx950)
    liblvgraph.so      csrc/lvgraph.cpp                          (C++17 + OpenMP)

Usage: python build.py [--quiet]
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INC = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: deltaModularity without FMA contraction (bit-exact with
# the reference's x86-64 arithmetic).  Packed FP32 instructions are disabled.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
             "-munsafe-fp-atomics", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]


def _run(cmd, quiet):
    if not quiet:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=HERE, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in (r.stderr or "").splitlines() if noise not in l and "(ignoring feature)" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed ({r.returncode}): {' '.join(cmd)}\n{err}")
    if err.strip() and not quiet:
        print(err)


def build(quiet=False):
    hip_src = [os.path.join(HERE, "csrc", f) for f in ("lv_kernels.hip", "lv_update.hip", "lv_capi.hip",
                                                          "lv_serve.hip")]
    out = os.path.join(HERE, "liblouvain_hip.so")
    _run([HIPCC, *HIP_FLAGS, f"-I{INC}", "-o", out + ".tmp", *hip_src], quiet)
    os.replace(out + ".tmp", out)
    gout = os.path.join(HERE, "liblvgraph.so")
    _run(["g++", "-std=c++17", "-O3", "-fopenmp", "-fPIC", "-shared", f"-I{INC}", "-o", gout + ".tmp",
          os.path.join(HERE, "csrc", "lvgraph.cpp")], quiet)
    os.replace(gout + ".tmp", gout)
    return [out, gout]


if __name__ == "__main__":
    for p in build(quiet="--quiet" in sys.argv):
        print("built", p)
