"""Pins the oracle (CPU restatement of mul_ntt / mul_fft / mul_into).

Each check is independent of the restatement's own formulas: GMP digests
(tests/golden/products.json), closed forms, textbook transforms computed with
Python integers, and the reference's compiled product where oracle/_ref exists.
"""
import os
import random

import numpy as np
import pytest

from preclib_amd import workloads as W
from tests.oracle_py import ROOT, digest, trim

PRIMES = [469762049, 1811939329, 2013265921]
GENS = [3, 13, 31]


def py_int(limbs):
    return sum(int(x) << (64 * i) for i, x in enumerate(limbs))


def to_limbs(v):
    out = []
    while v:
        out.append(v & 0xFFFFFFFFFFFFFFFF)
        v >>= 64
    return np.array(out, dtype=np.uint64)


def bitrev(i, bits):
    return int(format(i, "0%db" % bits)[::-1], 2) if bits else 0


# ---- Montgomery constants and arithmetic (mul_ntt.cpp:37-59) -------------

@pytest.mark.parametrize("k", [0, 1, 2])
def test_montgomery_constants(oracle, k):
    rc, (p, g, ninv, r2, one) = oracle.ctx(k)
    assert rc == 0, "derived constants differ from the frozen table"
    assert p == PRIMES[k] and g == GENS[k]
    assert (ninv * p) % 2**32 == 2**32 - 1           # ninv = -p^-1 mod 2^32
    assert r2 == pow(2, 64, p) and one == pow(2, 32, p)
    # g is a primitive root: g^((p-1)/q) != 1 for every prime q | p-1
    m, qs, q = p - 1, set(), 2
    while q * q <= m:
        while m % q == 0:
            qs.add(q)
            m //= q
        q += 1
    if m > 1:
        qs.add(m)
    assert all(pow(g, (p - 1) // q, p) != 1 for q in qs)
    assert (p - 1) % 2**26 == 0                      # supports n up to 2^26


@pytest.mark.parametrize("k", [0, 1, 2])
def test_mont_mul(oracle, k):
    p = PRIMES[k]
    rinv = pow(2**32, -1, p)
    rng = random.Random(k)
    for _ in range(2000):
        a, b = rng.randrange(p), rng.randrange(p)
        assert oracle.mont_mul(k, a, b) == a * b * rinv % p
    assert oracle.mont_mul(k, p - 1, p - 1) == (p - 1) ** 2 * rinv % p


# ---- transforms vs naive DFT (mul_ntt.cpp:193-329) -----------------------

@pytest.mark.parametrize("k", [0, 1, 2])
@pytest.mark.parametrize("n", [2, 8, 64, 256])
def test_forward_is_bitreversed_dft(oracle, k, n):
    p, g = PRIMES[k], GENS[k]
    R = 2**32 % p
    w = pow(g, (p - 1) // n, p)
    rng = random.Random(n * 7 + k)
    x = [rng.randrange(p) for _ in range(n)]
    got = oracle.ntt_forward(k, np.array([v * R % p for v in x], np.uint32))
    bits = n.bit_length() - 1
    for kk in range(n):
        X = sum(x[j] * pow(w, j * kk, p) for j in range(n)) % p
        assert int(got[bitrev(kk, bits)]) == X * R % p


@pytest.mark.parametrize("k", [0, 1, 2])
@pytest.mark.parametrize("n", [2, 16, 1024, 4096])
def test_inverse_undoes_forward(oracle, k, n):
    p = PRIMES[k]
    rng = np.random.default_rng(n + k)
    x = rng.integers(0, p, n, dtype=np.uint64).astype(np.uint32)
    assert np.array_equal(oracle.ntt_inverse(k, oracle.ntt_forward(k, x)), x)


@pytest.mark.parametrize("k", [0, 1, 2])
def test_convolution_theorem(oracle, k):
    p = PRIMES[k]
    rng = random.Random(99 + k)
    for (na, nb, n) in [(5, 3, 8), (100, 150, 256), (300, 200, 512), (1000, 1000, 2048)]:
        a = [rng.randrange(65536) for _ in range(na)]
        b = [rng.randrange(65536) for _ in range(nb)]
        ref = [0] * n
        for i, x in enumerate(a):
            for j, y in enumerate(b):
                ref[(i + j) % n] += x * y
        got = oracle.convolve(k, a, b, n)
        assert [int(v) for v in got] == [v % p for v in ref]


def test_crt2(oracle):
    p1, p2 = PRIMES[0], PRIMES[1]
    rng = random.Random(5)
    vals = [0, 1, p1 - 1, p1, p1 * p2 - 1, 0xFFFF**2 * 12345] + [rng.randrange(p1 * p2) for _ in range(3000)]
    for x in vals:
        assert oracle.crt2(x % p1, x % p2) == x


def test_digits_trim(oracle):
    a = np.array([0x0001000200030004, 0x0000000000000005], np.uint64)
    assert list(oracle.digits(a)) == [4, 3, 2, 1, 5]
    a = np.array([0xFFFF000000000000], np.uint64)
    assert list(oracle.digits(a)) == [0, 0, 0, 0xFFFF]


def test_plan_two_primes_bound(oracle):
    # two_mod_ok (mul_ntt.cpp:489-492): terms <= (p1 p2 - 1) / 0xFFFF^2
    assert (PRIMES[0] * PRIMES[1] - 1) // 0xFFFF**2 == 198186912
    a, b = W.pair(16384, 16384)
    assert oracle.ntt_plan(a, b) == (1 << 17, 2)
    a, b = W.pair(193, 193)
    assert oracle.ntt_plan(a, b) == (2048, 2)
    a, b = W.pair(192, 192)
    assert oracle.ntt_plan(a, b)[0] == 0


# ---- whole products vs GMP digests ---------------------------------------

def _wl(golden, an, bn, dist):
    for e in golden["workload"]:
        if e["an"] == an and e["bn"] == bn and e["dist"] == dist:
            return e
    raise KeyError((an, bn, dist))


@pytest.mark.parametrize("dist", ["uniform", "ones", "sparse"])
def test_ntt_16384_vs_gmp(oracle, golden, dist):
    a, b = W.pair(16384, 16384, dist)
    r = oracle.mul_ntt(a, b)
    e = _wl(golden, 16384, 16384, dist)
    assert len(r) == e["len"] and digest(r) == e["sha256"]


@pytest.mark.parametrize("dist", ["uniform", "ones", "sparse"])
def test_fft_4096_vs_gmp(oracle, golden, dist):
    a, b = W.pair(4096, 4096, dist)
    r, err = oracle.mul_fft(a, b)
    e = _wl(golden, 4096, 4096, dist)
    assert digest(r) == e["sha256"]
    assert 0.0 <= err < 0.125


@pytest.mark.slow
def test_ntt_2p20_uniform_vs_gmp(oracle, golden):
    a, b = W.pair(1 << 20, 1 << 20, "uniform")
    r = oracle.mul_ntt(a, b)
    assert digest(r) == _wl(golden, 1 << 20, 1 << 20, "uniform")["sha256"]


def test_torture_fft_and_ntt_vs_gmp(oracle, golden):
    from tests.golden.gen_golden import TORTURE_PATTERNS, torture_operands
    pats = {p[0]: p for p in TORTURE_PATTERNS}
    for e in golden["torture"]:
        a, b = torture_operands(e["an"], e["bn"], pats[e["pattern"]])
        r, err = oracle.mul_fft(a, b)
        assert digest(r) == e["sha256"], e
        assert err < 0.5
        if max(e["an"], e["bn"]) > 1024:
            assert digest(oracle.mul_ntt(a, b)) == e["sha256"], e


def test_all_mul_eq_grid_vs_gmp(oracle, golden):
    for e in golden["all_mul_eq"]:
        a = W.reference_pattern32(e["an"], e["seed"])
        b = W.reference_pattern32(e["bn"], e["seed"] + 1009)
        assert digest(oracle.mul_into(a, b)) == e["sha256"], e
        assert digest(oracle.mul_basic(a, b)) == e["sha256"], e
        if max(e["an"], e["bn"]) > 192:
            assert digest(oracle.mul_ntt(a, b)) == e["sha256"], e
            assert digest(oracle.mul_fft(a, b)[0]) == e["sha256"], e


def test_bench_gmp_rows_vs_gmp(oracle, golden):
    for e in golden["bench_gmp"]:
        a = W.reference_pattern(e["an"], 11 + e["p"])
        b = W.reference_pattern(e["bn"], 19 + e["p"])
        assert digest(oracle.mul_into(a, b)) == e["sha256"], e


def test_batch_first_products_vs_gmp(oracle, golden):
    shapes = W.batch_shapes()
    for k, want in enumerate(golden["batch"]["first"][:24]):
        an, bn = shapes[k]
        a, b = W.pair(an, bn, "uniform", k)
        assert digest(oracle.mul_into(a, b)) == want, (k, an, bn)


# ---- closed forms ---------------------------------------------------------

@pytest.mark.parametrize("n", [193, 1000, 6145, 20000])
def test_all_ones_closed_form(oracle, n):
    a = np.full(n, 0xFFFFFFFFFFFFFFFF, np.uint64)
    want = np.concatenate([np.array([1], np.uint64), np.zeros(n - 1, np.uint64),
                           np.array([0xFFFFFFFFFFFFFFFE], np.uint64),
                           np.full(n - 1, 0xFFFFFFFFFFFFFFFF, np.uint64)])
    assert np.array_equal(oracle.mul_ntt(a, a), want)
    assert np.array_equal(oracle.mul_fft(a, a)[0], want)
    assert np.array_equal(oracle.mul_into(a, a, sqr=True), want)


def test_sparse_brute_force(oracle):
    a, b = W.pair(9000, 7000, "sparse")
    want = to_limbs(py_int(a) * py_int(b))
    assert np.array_equal(oracle.mul_ntt(a, b), want)
    assert np.array_equal(oracle.mul_into(a, b), want)


def test_blocked_path_vs_python(oracle):
    for (an, bn) in [(30, 100), (193, 1000), (500, 7000), (7000, 20000), (25, 50000)]:
        a, b = W.pair(an, bn, "uniform", an)
        assert np.array_equal(oracle.mul_into(a, b), to_limbs(py_int(a) * py_int(b))), (an, bn)


def test_small_edges(oracle):
    one = np.array([1], np.uint64)
    m = np.array([0xFFFFFFFFFFFFFFFF], np.uint64)
    assert list(oracle.mul_into(m, m)) == [1, 0xFFFFFFFFFFFFFFFE]
    assert list(oracle.mul_into(one, one)) == [1]
    z = np.zeros(0, np.uint64)
    assert len(oracle.mul_into(z, one)) == 0
    # (2^32-1)^2 KAT through every entry (test_prec.cpp:184-195)
    x = np.array([0xFFFFFFFF], np.uint64)
    for r in (oracle.mul_ntt(x, x), oracle.mul_fft(x, x)[0], oracle.mul_basic(x, x)):
        assert list(r) == [0xFFFFFFFE00000001]


# ---- against the reference's own compiled product (oracle/_ref) ----------

REF = os.path.join(ROOT, "oracle", "_ref", "libpreclib_ref.so")


@pytest.mark.skipif(not os.path.exists(REF), reason="reference build absent")
def test_against_reference_build(oracle):
    import ctypes
    from tests.oracle_py import P64, SZ, _p64
    R = ctypes.CDLL(REF)
    R.ref_mul_into.restype = SZ
    R.ref_mul_into.argtypes = [P64, SZ, P64, SZ, P64, SZ, ctypes.c_int]
    R.ref_mul_fft.restype = SZ
    R.ref_mul_fft.argtypes = [P64, SZ, P64, SZ, P64, SZ, ctypes.POINTER(ctypes.c_double)]
    for (an, bn, dist) in [(200, 300, 0), (1000, 1000, 1), (2049, 2039, 0), (7000, 7000, 2),
                           (16384, 16384, 0), (6145, 20000, 1), (5000, 9999, 0), (4096, 4096, 1)]:
        a, b = W.pair(an, bn, dist)
        r = np.zeros(an + bn, np.uint64)
        n = R.ref_mul_into(_p64(r), an + bn, _p64(a), an, _p64(b), bn, 0)
        assert np.array_equal(oracle.mul_into(a, b), r[:n])
        e = ctypes.c_double()
        n = R.ref_mul_fft(_p64(r), an + bn, _p64(a), an, _p64(b), bn, ctypes.byref(e))
        mine, err = oracle.mul_fft(a, b)
        assert np.array_equal(mine, r[:n])
        assert err == e.value        # same operation order -> same doubles


def _is_prime64(n):
    """Deterministic Miller-Rabin for n < 2^64 (bases 2..37)."""
    if n < 2:
        return False
    small = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)
    for q in small:
        if n % q == 0:
            return n == q
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in small:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def test_mod_checks_against_python_ints(oracle):
    """or_mod_p / or_mod_p3 (the serving tests' residue checks) against
    Python's integer remainder, on random, all-ones and empty inputs; the
    three moduli are prime."""
    import random
    rng = random.Random(11)
    for p in oracle.MODS3:
        assert _is_prime64(p)
    assert not _is_prime64((1 << 64) - 57)   # the check can fail
    for n in (0, 1, 2, 3, 17, 1000):
        for a in (np.array([rng.getrandbits(64) for _ in range(n)], np.uint64),
                  np.full(n, 0xFFFFFFFFFFFFFFFF, np.uint64)):
            x = int.from_bytes(a.astype("<u8").tobytes(), "little")
            assert oracle.mod_p3(a) == tuple(x % p for p in oracle.MODS3)
            for k, c in ((61, 1), (63, 25), (64, 59), (40, 87)):
                assert oracle.mod_p(a, k, c) == x % ((1 << k) - c)


def test_decimal_residues_against_python_ints(oracle):
    """or_dec_mod_p3 (the decimal parse's residue check) against Python's
    own integer parse and remainder."""
    import random
    import sys
    if hasattr(sys, "set_int_max_str_digits"):
        sys.set_int_max_str_digits(0)
    rng = random.Random(5)
    for n in (0, 1, 2, 19, 20, 700, 5000):
        s = "".join(rng.choice("0123456789") for _ in range(n))
        x = int(s) if s else 0
        assert oracle.dec_mod_p3(s) == tuple(x % p for p in oracle.MODS3)
    assert oracle.dec_mod_p3("9" * 3000) == tuple((10 ** 3000 - 1) % p for p in oracle.MODS3)
