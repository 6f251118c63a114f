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
SOURCES = ["pg_ntt.hip", "pg_ntt_v1.hip", "pg_ntt_v2.hip", "pg_ntt_v3.hip", "pg_ntt_v4.hip", "pg_ntt_v5.hip",
           "pg_ntt_v6.hip", "pg_ntt_v7.hip",
           "pg_ntt_v8.hip", "pg_ntt_v9.hip", "pg_ntt_v10.hip", "pg_ntt_v11.hip", "pg_ntt_v12.hip", "pg_ntt_v13.hip", "pg_ntt_v14.hip", "pg_ntt_v15.hip", "pg_fft.hip", "pg_small.hip", "pg_carry.hip", "pg_batch.hip", "pg_runtime.hip", "pg_dint.hip",
           "pg_pi.hip", "pg_serve.hip", "pg_noalloc.hip", "pg_api.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-ffp-contract=off", "-Wno-unused-result",
         "-I" + CSRC, "-I" + os.path.join(ROOT, "include")]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed: " + " ".join(cmd[-3:]))


def build(verbose=False, out=LIB, defines=()):
    """Compiles every source to an object in parallel, then links libpgpu.so."""
    import concurrent.futures
    import tempfile
    defs = ["-D" + d for d in defines]
    cflags = [f for f in FLAGS if f != "-shared"]
    with tempfile.TemporaryDirectory(prefix="pgpu_build_") as tmpd:
        objs = [os.path.join(tmpd, os.path.splitext(s)[0] + ".o") for s in SOURCES]
        jobs = [[HIPCC] + cflags + defs + ["-c", os.path.join(CSRC, s), "-o", o] for s, o in zip(SOURCES, objs)]
        with concurrent.futures.ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 4)) as ex:
            for f in [ex.submit(_run, j, verbose) for j in jobs]:
                f.result()
        tmp = out + ".tmp"
        _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", tmp], verbose)
        os.replace(tmp, out)
    return out


HOST_LIB = os.path.join(HERE, "libprecn.so")
CPP_TEST = os.path.join(ROOT, "tests", "cpp", "test_precn")
CPP_BRANCH = os.path.join(ROOT, "tests", "cpp", "integration_branch")
CXX = os.environ.get("CXX", "g++")


def build_host(verbose=False):
    """The precn_t layer (host C++17 over libpgpu.so) and its C++ test
    program (which also links the oracle, as its checker)."""
    inc = os.path.join(ROOT, "include")
    _run([CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-I" + inc, os.path.join(HERE, "host", "precn.cpp"),
          os.path.join(HERE, "host", "precz.cpp"), "-L" + HERE, "-lpgpu", "-Wl,-rpath,$ORIGIN", "-o", HOST_LIB],
         verbose)
    orc = os.path.join(ROOT, "oracle")
    _run([CXX, "-O2", "-std=c++17", "-I" + inc, "-I" + orc, os.path.join(ROOT, "tests", "cpp", "test_precn.cpp"),
          "-L" + HERE, "-lprecn", "-lpgpu", "-L" + orc, "-loracle",
          "-Wl,-rpath," + HERE + ":" + orc + ":$ORIGIN/../../preclib_amd:$ORIGIN/../../oracle", "-o", CPP_TEST],
         verbose)
    # the dispatcher branch INTEGRATION.md shows for preclib's mul_into
    _run([CXX, "-O2", "-std=c++17", "-I" + inc, os.path.join(ROOT, "tests", "cpp", "integration_branch.cpp"),
          "-L" + HERE, "-lprecn", "-lpgpu", "-Wl,-rpath," + HERE + ":$ORIGIN/../../preclib_amd", "-o", CPP_BRANCH],
         verbose)
    return HOST_LIB


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_host(verbose=True))
