"""Decimal parse on the GPU (pg_from_decimal, precn_t(std::string),
src/init_basic.cpp:66-92): against Python's own integer parse up to 10^5
digits, against the string's residues modulo three primes (the oracle's
Horner over the digits) up to 10^7 digits, as the inverse of
pg_to_decimal (round trips), and the reference's character rule (non-digits
skipped)."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

from preclib_amd import workloads as W

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def pg():
    from preclib_amd import pgpu
    pgpu.init(0)
    if hasattr(sys, "set_int_max_str_digits"):
        sys.set_int_max_str_digits(0)
    return pgpu


def limbs_of(x):
    n = (x.bit_length() + 63) // 64
    return np.frombuffer(x.to_bytes(8 * n, "little"), dtype="<u8").astype(np.uint64) if n else np.zeros(0, np.uint64)


def digits(n, seed):
    rng = np.random.default_rng(seed)
    d = rng.integers(0, 10, n, dtype=np.uint8) + ord("0")
    if n:
        d[0] = ord("1") + d[0] % 9
    return d.tobytes().decode()


@pytest.mark.parametrize("n", [1, 2, 18, 19, 20, 37, 38, 39, 607, 608, 609, 1216, 1217, 5000, 19456, 100000])
def test_parse_against_python(pg, n):
    s = digits(n, n)
    print("case parse", n, flush=True)
    got = pg.from_decimal(s)
    assert np.array_equal(got, limbs_of(int(s)))


def test_parse_edges(pg):
    assert len(pg.from_decimal("")) == 0 and len(pg.from_decimal("0000")) == 0
    assert np.array_equal(pg.from_decimal("12a34"), limbs_of(1234))          # init_basic.cpp:69-72
    assert np.array_equal(pg.from_decimal("-1,000,000 x"), limbs_of(1000000))
    assert np.array_equal(pg.from_decimal("0" * 700 + "5"), limbs_of(5))
    assert np.array_equal(pg.from_decimal("18446744073709551616"), limbs_of(1 << 64))
    assert np.array_equal(pg.from_decimal("9" * 608), limbs_of(10 ** 608 - 1))   # one leaf, full
    assert np.array_equal(pg.from_decimal("1" + "0" * 608), limbs_of(10 ** 608))


@pytest.mark.parametrize("n", [1000, 123457, 1000000, 10000000])
def test_parse_residues_and_round_trip(pg, oracle, n):
    """No second big-integer parser at these sizes: the parsed limbs' residues
    equal the digit string's, and converting back gives the string."""
    s = digits(n, 7 + n)
    print("case parse residues", n, flush=True)
    got = pg.from_decimal(s)
    assert oracle.mod_p3(got) == oracle.dec_mod_p3(s)
    assert pg.to_decimal(got) == s


def test_parse_pi_digits(pg, oracle):
    """The 10^7-digit string of the pi program (its SHA-256 is the reference
    program's), parsed and printed back."""
    gold = {e["digits"]: e for e in json.load(open(os.path.join(ROOT, "tests", "golden", "pi.json")))["pi"]}
    s = pg.pi(10 ** 7)
    assert hashlib.sha256((s + "\n").encode()).hexdigest() == gold[10 ** 7]["sha256"]
    d = s.replace(".", "")
    x = pg.from_decimal(s)                       # the '.' is skipped
    assert oracle.mod_p3(x) == oracle.dec_mod_p3(d)
    assert pg.to_decimal(x) == d


def test_limbs_round_trip(pg):
    for k, n in enumerate([1, 7, 31, 32, 33, 1000, 70000]):
        a = W.operand(n, 600 + k)
        assert np.array_equal(pg.from_decimal(pg.to_decimal(a)), a)
