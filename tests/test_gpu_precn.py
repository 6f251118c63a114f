"""The precn_t C++ layer (include/precn.hpp, preclib_amd/host/precn.cpp) run
through its C++ test program, which checks every product, division, square
root and conversion entry point against the oracle."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "test_precn")
BRANCH = os.path.join(ROOT, "tests", "cpp", "integration_branch")


@pytest.mark.gpu
def test_precn_cpp_program():
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/test_precn not built (preclib_amd.build.build_host)")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=240)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert "all passed" in r.stdout


@pytest.mark.gpu
def test_integration_branch_program():
    """The mul_into branch printed in INTEGRATION.md (option 2), compiled as
    tests/cpp/integration_branch.cpp, against schoolbook products."""
    if not os.path.exists(BRANCH):
        pytest.fail("tests/cpp/integration_branch not built (preclib_amd.build.build_host)")
    r = subprocess.run([BRANCH], capture_output=True, text=True, timeout=240)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "integration branch ok" in r.stdout


def test_integration_branch_matches_doc():
    """The function INTEGRATION.md prints is the one the program compiles."""
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    src = open(BRANCH + ".cpp").read()
    body = src[src.index("static bool mul_gpu("):]
    body = body[: body.index("\n}\n") + 2]
    assert body.replace("reserve_limbs", "mul_reserve") in doc


def test_precn_library_links():
    # the host layer resolves against libpgpu.so without a device call
    import ctypes
    lib = os.path.join(ROOT, "preclib_amd", "libprecn.so")
    if not os.path.exists(lib):
        pytest.skip("libprecn.so not built")
    ctypes.CDLL(os.path.join(ROOT, "preclib_amd", "libpgpu.so"), mode=ctypes.RTLD_GLOBAL)
    L = ctypes.CDLL(lib)
    assert L._Z14precn_gpu_initim is not None   # precn_gpu_init(int, unsigned long)
