"""GPU binary splitting, square root, decimal conversion and the pi program
against the oracle (oracle/or_pi.py, or_sqrt), Python integers, GMP digests
and the reference pi program's own output digests.  Bit-exact."""
import hashlib
import json
import math
import os
import sys

import numpy as np
import pytest

from preclib_amd import workloads as W
from tests.oracle_py import trim

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import or_pi  # noqa: E402

pytestmark = pytest.mark.gpu
if hasattr(sys, "set_int_max_str_digits"):
    sys.set_int_max_str_digits(0)


@pytest.fixture(scope="module")
def pg():
    from preclib_amd import pgpu
    pgpu.init(0)
    return pgpu


@pytest.fixture(scope="module")
def golden_pi():
    with open(os.path.join(ROOT, "tests", "golden", "pi.json")) as f:
        return {e["digits"]: e for e in json.load(f)["pi"]}


def to_int(x):
    return int.from_bytes(np.asarray(x, dtype="<u8").tobytes(), "little")


def from_int(v):
    n = max(1, (v.bit_length() + 63) // 64)
    return trim(np.frombuffer(v.to_bytes(n * 8, "little"), dtype="<u8").copy())


def sha(x):
    return hashlib.sha256(trim(x).astype("<u8").tobytes()).hexdigest()


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 31, 32, 33, 63, 64, 65, 100, 127, 129, 1000, 4097])
def test_chud_split_vs_oracle(pg, n):
    p, q, t = pg.chud_split(n)
    ep, eq, et = or_pi.chud_bs(0, n)
    assert to_int(p) == ep and to_int(q) == eq and to_int(t) == et


def test_chud_split_root_without_p(pg):
    n = 20000
    p, q, t = pg.chud_split(n, need_p=False)
    _, eq, et = or_pi.chud_bs(0, n, False)
    assert p is None and to_int(q) == eq and to_int(t) == et


@pytest.mark.parametrize("an", [1, 2, 3, 4, 5, 6, 7, 8, 9, 17, 64, 65, 333, 1000, 4096])
def test_sqrt_vs_python(pg, an):
    a = W.operand(an, W.SEED_A + 800 + an, "uniform")
    assert to_int(pg.isqrt(a)) == math.isqrt(to_int(a))


def test_sqrt_edges(pg, oracle):
    for v in [1, 2, 3, 4, 15, 16, 17, (1 << 128) - 1, 1 << 128, ((1 << 900) + 7) ** 2, ((1 << 900) + 7) ** 2 - 1,
              10 ** 3000 * 10005]:
        assert to_int(pg.isqrt(from_int(v))) == math.isqrt(v)
    ones = np.full(300, 0xFFFFFFFFFFFFFFFF, np.uint64)
    assert np.array_equal(pg.isqrt(ones), oracle.sqrt(ones))


def test_sqrt_golden(pg, golden_div):
    for c in golden_div["sqrt"]:
        a = W.operand(c["an"], W.SEED_A + c["k"], "uniform")
        x = pg.isqrt(a)
        assert len(x) == c["len"] and sha(x) == c["sha256"]


@pytest.mark.parametrize("v", [1, 9, 10, 10 ** 19 - 1, 10 ** 19, 10 ** 38 - 1, 10 ** 38, 2 ** 64 - 1, 2 ** 64,
                               10 ** 608 - 1, 10 ** 608, 10 ** 1216 + 1, 3 ** 2000, 7 ** 20000])
def test_to_decimal_values(pg, v):
    assert pg.to_decimal(from_int(v)) == str(v)


def test_to_decimal_zero_and_random(pg):
    assert pg.to_decimal(np.zeros(0, np.uint64)) == ""
    for n in (1, 2, 33, 500, 3000, 20000):
        a = W.operand(n, W.SEED_B + 900 + n, "uniform")
        assert pg.to_decimal(a) == str(to_int(a))
    ones = np.full(5000, 0xFFFFFFFFFFFFFFFF, np.uint64)
    assert pg.to_decimal(ones) == str((1 << (64 * 5000)) - 1)


@pytest.mark.parametrize("d", [0, 1, 2, 9, 10, 100, 1000, 12345, 100000, 1000000])
def test_pi_digits(pg, golden_pi, d):
    s = pg.pi(d)
    g = golden_pi[d]
    assert len(s) == g["len"] and s[:100] == g["head"]
    assert hashlib.sha256((s + "\n").encode()).hexdigest() == g["sha256"]


def test_pi_ten_million(pg, golden_pi):
    s, ph = pg.pi(10000000, want_phases=True)
    g = golden_pi[10000000]
    print("pi 10^7 phases (ms): binary_split %.1f sqrt_scale %.1f final_div %.1f to_decimal %.1f" % tuple(ph))
    assert len(s) == g["len"] and s[-100:] == g["tail"]
    assert hashlib.sha256((s + "\n").encode()).hexdigest() == g["sha256"]
