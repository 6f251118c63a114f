"""The drop-in header against the reference's own sources (CPU, no device):
every source file of the reference library and its test and pi programs
compiles, unchanged and where it lies, with `#include "../prec.hpp"`
redirected to include/precn.hpp (a clang VFS overlay maps the reference's
prec.hpp to a one-line shim); and every function prec.hpp declares is
declared by include/precn.hpp.  Skipped where the reference tree or the
ROCm clang is absent."""
import glob
import json
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
CLANG = "/opt/rocm/llvm/bin/clang++"

pytestmark = pytest.mark.skipif(not (os.path.exists(os.path.join(REF, "prec.hpp")) and os.path.exists(CLANG)),
                                reason="reference tree or ROCm clang absent")


def sources():
    out = sorted(glob.glob(os.path.join(REF, "src", "*.cpp")))
    out += [os.path.join(REF, "pi", "pi_chudnovsky.cpp")]
    out += [os.path.join(REF, "test", f) for f in ("test_prec.cpp", "test_opt_smoke.cpp", "fft_torture.cpp",
                                                    "bench_ntt.cpp", "bench_div.cpp")]
    return out


@pytest.fixture(scope="module")
def overlay(tmp_path_factory):
    d = tmp_path_factory.mktemp("dropin")
    shim = d / "prec.hpp"
    shim.write_text('#include "%s"\n' % os.path.join(ROOT, "include", "precn.hpp"))
    ov = d / "overlay.yaml"
    ov.write_text(json.dumps({"version": 0, "case-sensitive": "true", "overlay-relative": False,
                              "roots": [{"type": "directory", "name": REF,
                                         "contents": [{"type": "file", "name": "prec.hpp",
                                                       "external-contents": str(shim)}]}]}))
    return str(ov)


@pytest.mark.parametrize("src", sources(), ids=lambda p: os.path.relpath(p, REF))
def test_reference_source_compiles_against_dropin(overlay, src):
    r = subprocess.run([CLANG, "-std=c++17", "-fsyntax-only", "-DCOUNT_FFTS=1", "-ivfsoverlay", overlay, src],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and not r.stderr.strip(), r.stderr[-3000:]


def test_overlay_is_what_compiles(overlay, tmp_path):
    # the same overlay aimed at a shim that refuses to compile: the reference
    # source then fails, so the passing runs above used include/precn.hpp
    bad = tmp_path / "bad.hpp"
    bad.write_text("#error dropin-shim\n")
    ov = tmp_path / "bad.yaml"
    ov.write_text(open(overlay).read().replace(json.loads(open(overlay).read())["roots"][0]["contents"][0]
                                               ["external-contents"], str(bad)))
    r = subprocess.run([CLANG, "-std=c++17", "-fsyntax-only", "-ivfsoverlay", str(ov),
                        os.path.join(REF, "src", "add.cpp")], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "dropin-shim" in r.stderr


def _declared(path):
    src = open(path).read()
    src = re.sub(r"//.*", "", src)
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*|operator\s*[^\s(]+)\s*\([^;{]*\)\s*(?:const\s*)?[;{]", src))
    return {re.sub(r"\s+", "", n) for n in names}


def test_every_reference_declaration_present():
    ref = _declared(os.path.join(REF, "prec.hpp"))
    ours = _declared(os.path.join(ROOT, "include", "precn.hpp"))
    skip = {"if", "while", "for", "return", "sizeof", "malloc", "realloc", "free", "abort"}   # calls in inline bodies
    missing = sorted(n for n in ref - ours if n not in skip)
    assert not missing, missing
