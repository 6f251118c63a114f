"""The reference's own programs, unchanged, built against the drop-in header
include/precn.hpp and linked with libprecn.so / libpgpu.so (oracle/Makefile
target `dropin`, into oracle/_ref/): its unit tests (test/test_prec.cpp
--checks-only: init, compare/shift, base conversion, add/sub, mul_u32,
mul_basic, divexact, division, gcd, precz_t over every sign, all
multiplication algorithms equal), its optimisation smoke test, its FFT
torture test (88 adversarial products, mul_fft == mul_ntt, COUNT_FFTS
statistics) and its pi program, whose digit string must equal the
reference program's.  Every product they make runs on the GPU."""
import hashlib
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_BIN = os.path.join(ROOT, "oracle", "_ref")


def run(name, *args, timeout=300):
    exe = os.path.join(REF_BIN, name)
    if not os.path.exists(exe):
        pytest.skip("%s not built (make -f oracle/Makefile dropin, where the reference tree is present)" % name)
    print("case", name, *args, flush=True)
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=timeout)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    return r


def test_reference_unit_tests():
    r = run("test_prec_gpu", "--checks-only")
    assert "ok" in r.stdout.split()


def test_reference_opt_smoke():
    run("test_opt_smoke_gpu")


def test_reference_fft_torture():
    r = run("fft_torture_gpu")
    out = r.stdout.split("\n")
    # 87 FFT products: the popcount operands of shape 193 trim to 192 limbs,
    # which mul_fft multiplies with mul_basic (mul_fft.cpp:388), uncounted
    assert "cases 88" in out and "total_fftmuls 87" in out and "ok" in out
    assert "danger_fftmuls 0" in out


@pytest.mark.parametrize("digits", [1000, 100000])
def test_reference_pi_program(tmp_path, digits):
    gold = {e["digits"]: e for e in json.load(open(os.path.join(ROOT, "tests", "golden", "pi.json")))["pi"]}
    out = tmp_path / "pi.txt"
    run("pi_chudnovsky_gpu", str(digits), "--file", str(out), timeout=600)
    assert hashlib.sha256(out.read_bytes()).hexdigest() == gold[digits]["sha256"]
