"""Builds libpgpu.so (HIP kernels + C ABI) for gfx950 with hipcc, in-tree.

Usage: python -m preclib_amd.build   (or preclib_amd.build.build()).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpgpu.so")
SOURCES = ["pg_ntt.hip", "pg_carry.hip", "pg_runtime.hip", "pg_api.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-ffp-contract=off", "-Wno-unused-result",
         "-I" + CSRC, "-I" + os.path.join(ROOT, "include")]


def build(verbose=False, out=LIB, defines=()):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = out + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-D" + d for d in defines] + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for libpgpu.so")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
