"""Failure handling through the C ABI, each in a fresh process (the budget
is set at pg_init): device-memory budget exhaustion returns PG_ENOMEM from
pg_buf_alloc and from a product whose scratch does not fit, leaves no memory
behind (a smaller product then succeeds, repeatedly); bad arguments return
PG_EINVAL."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r'''
import ctypes, sys
sys.path.insert(0, ROOT)
import numpy as np
from preclib_amd import pgpu, workloads as W
L = pgpu.load()
P64 = ctypes.POINTER(ctypes.c_uint64)
assert L.pg_init(0, 48 << 20) == 0                       # a 48 MiB device budget
h = ctypes.c_void_p()
assert L.pg_buf_alloc(ctypes.byref(h), (96 << 20) // 8) == pgpu.PG_ENOMEM
assert b"budget" in L.pg_last_error()
big_a, big_b = W.pair(1 << 20, 1 << 20, "uniform", 1)   # 8 MiB operands, ~300 MiB of spectra
small_a, small_b = W.pair(3000, 2000, "uniform", 2)
for rep in range(3):
    r = np.zeros(2 << 20, np.uint64)
    n = ctypes.c_size_t(0)
    st = L.pg_mul(r.ctypes.data_as(P64), len(r), ctypes.byref(n), big_a.ctypes.data_as(P64), len(big_a),
                  big_b.ctypes.data_as(P64), len(big_b), 0, None)
    assert st == pgpu.PG_ENOMEM, st
    r = np.zeros(5000, np.uint64)
    st = L.pg_mul(r.ctypes.data_as(P64), len(r), ctypes.byref(n), small_a.ctypes.data_as(P64), len(small_a),
                  small_b.ctypes.data_as(P64), len(small_b), 0, None)
    assert st == 0 and n.value in (4999, 5000), (st, L.pg_last_error())
# bad arguments
assert L.pg_mul(None, 10, None, small_a.ctypes.data_as(P64), 3, small_b.ctypes.data_as(P64), 3, 0, None) == pgpu.PG_EINVAL
r = np.zeros(4, np.uint64)
assert L.pg_mul(r.ctypes.data_as(P64), 4, None, small_a.ctypes.data_as(P64), 3, small_b.ctypes.data_as(P64), 3, 0,
                None) == pgpu.PG_EINVAL                   # r_cap < an + bn
print("failures ok")
'''


def test_budget_exhaustion_and_bad_arguments():
    r = subprocess.run([sys.executable, "-c", SCRIPT.replace("ROOT", repr(ROOT))], capture_output=True,
                       text=True, timeout=240)
    print(r.stdout[-2000:], r.stderr[-3000:])
    assert r.returncode == 0 and "failures ok" in r.stdout
