"""CPU checks of the C-ABI library: it loads, and exports every symbol that
include/pgpu.h declares (no device calls)."""
import ctypes
import os
import re

import pytest

from preclib_amd import pgpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pg_[a-z_]+)\s*\(", src)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(pgpu.LIB_PATH):
        from preclib_amd import build
        build.build()
    return ctypes.CDLL(pgpu.LIB_PATH)


def test_exports_every_declared_symbol(lib):
    names = declared("pgpu.h")
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(pgpu.EXPORTS) == names


def test_calls_before_init_fail_cleanly(lib):
    pgpu.load()
    r = (ctypes.c_uint64 * 4)()
    st = pgpu._lib.pg_mul(r, 4, None, r, 2, r, 2, 0, None)
    assert st == pgpu.PG_ENOINIT
    assert b"pg_init" in pgpu._lib.pg_last_error()
