"""GPU Newton division (pg_div / pg_recip / pg_div_dev) against the oracle,
GMP digests and the reference's own identities.  Bit-exact."""
import hashlib

import numpy as np
import pytest

from preclib_amd import workloads as W
from tests.oracle_py import trim

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pg():
    from preclib_amd import pgpu
    pgpu.init(0)
    return pgpu


def to_int(x):
    return int.from_bytes(np.asarray(x, dtype="<u8").tobytes(), "little")


def from_int(v):
    n = max(1, (v.bit_length() + 63) // 64)
    return trim(np.frombuffer(v.to_bytes(n * 8, "little"), dtype="<u8").copy())


def sha(x):
    return hashlib.sha256(trim(x).astype("<u8").tobytes()).hexdigest()


def test_reference_known_answers(pg):
    # test_prec.cpp:243-248
    assert list(pg.recip(from_int(3), 8)) == [85]
    assert list(pg.recip(from_int(5), 3)) == [1]
    assert len(pg.recip(from_int(9), 3)) == 0
    assert len(pg.recip(np.zeros(0, np.uint64), 128)) == 0
    q, r = pg.divmod_(from_int(7), from_int(3))
    assert list(q) == [2] and list(r) == [1]
    q, r = pg.divmod_(from_int(3), from_int(7))
    assert len(q) == 0 and list(r) == [3]
    q, r = pg.divmod_(from_int(9), np.zeros(0, np.uint64))
    assert len(q) == 0 and len(r) == 0


SHAPES = [(2, 1), (3, 2), (10, 3), (40, 20), (80, 40), (193, 192), (300, 100), (1000, 999), (2000, 700),
          (5000, 2000), (8000, 25), (13000, 6500)]


@pytest.mark.parametrize("an,bn", SHAPES)
@pytest.mark.parametrize("dist", ["uniform", "ones", "sparse"])
def test_vs_oracle(pg, oracle, an, bn, dist):
    a, b = W.pair(an, bn, dist, 700 + an)
    q, r = pg.divmod_(a, b)
    eq, er = oracle.divmod(a, b)
    assert np.array_equal(q, eq) and np.array_equal(r, er)


@pytest.mark.parametrize("top", [0x8000000000000000, 0xFFFFFFFFFFFFFFFF, 1])
def test_divisor_tops(pg, oracle, top):
    b = np.full(700, 0xFFFFFFFFFFFFFFFF, np.uint64)
    b[-1] = np.uint64(top)
    b[0] = np.uint64(0)
    a = np.full(2100, 0xFFFFFFFFFFFFFFFF, np.uint64)
    q, r = pg.divmod_(a, b)
    eq, er = oracle.divmod(a, b)
    assert np.array_equal(q, eq) and np.array_equal(r, er)


@pytest.mark.parametrize("n", [0, 1, 4, 8, 10, 12, 14, 15])
def test_bench_div_identity(pg, oracle, n):
    # test_prec.cpp:628-653: dividend = divisor * quotient + (divisor - 1)
    limbs = 1 << n
    d = W.reference_pattern32(limbs, 12000 + n)
    q = W.reference_pattern32(limbs, 13000 + n)
    p = from_int(to_int(oracle.mul_into(d, q)) + to_int(d) - 1)
    qq, rr = pg.divmod_(p, d)
    assert np.array_equal(qq, trim(q))
    assert to_int(rr) == to_int(d) - 1


def test_golden_divisions(pg, golden_div):
    for c in golden_div["div"]:
        a, b = W.pair(c["an"], c["bn"], "uniform", c["k"])
        q, r = pg.divmod_(a, b)
        assert len(q) == c["q_len"] and sha(q) == c["q_sha256"], (c["an"], c["bn"])
        assert len(r) == c["r_len"] and sha(r) == c["r_sha256"], (c["an"], c["bn"])


def test_golden_reciprocals(pg, golden_div):
    for c in golden_div["recip"]:
        b = W.operand(c["bn"], W.SEED_B + c["k"], "uniform")
        x = pg.recip(b, c["bits"])
        assert len(x) == c["len"] and sha(x) == c["sha256"]


def test_recip_vs_oracle(pg, oracle):
    rng = np.random.default_rng(13)
    for bn, bits in [(1, 64), (1, 65), (2, 129), (3, 500), (20, 4096), (40, 4096), (7, 449), (500, 64 * 3000)]:
        b = rng.integers(1, 2**64, bn, dtype=np.uint64)
        assert np.array_equal(pg.recip(b, bits), oracle.recip(b, bits))


def test_device_resident(pg, oracle):
    a, b = W.pair(9000, 4000, "uniform", 901)
    da, db = pg.DevBuf.of(a), pg.DevBuf.of(b)
    dq, dr = pg.DevBuf(9000 - 4000 + 1), pg.DevBuf(4000)
    ql, rl = pg.div_dev(dq, dr, da, len(a), db, len(b))
    eq, er = oracle.divmod(a, b)
    assert np.array_equal(dq.download(ql), eq) and np.array_equal(dr.download(rl), er)
    # quotient only, operands with zero limbs above their length
    a2 = np.concatenate([a, np.zeros(5, np.uint64)])
    da2 = pg.DevBuf.of(a2)
    ql, _ = pg.div_dev(dq, None, da2, len(a2), db, len(b))
    assert np.array_equal(dq.download(ql), eq)


@pytest.mark.parametrize("amount", [1000, -1000, 10 ** 6, -(10 ** 6), 1 << 40, 3])
@pytest.mark.parametrize("shape", [(20000, 7000), (300, 60), (9000, 4500)])
def test_quotient_correction_from_far(pg, oracle, shape, amount):
    """The first quotient estimate moved by `amount` units (test hook): the
    correction of div_newton.cpp:114-125 (delta = floor(|r| / b), not unit
    steps) still ends at the exact floor."""
    an, bn = shape
    a, b = W.pair(an, bn, "uniform", 91 + an)
    wq, wr = oracle.divmod(a, b)
    print("case perturb q", shape, amount, flush=True)
    pg.debug_perturb(0, amount)
    q, r = pg.divmod_(a, b)
    assert np.array_equal(q, wq) and np.array_equal(r, wr)
    q2, r2 = pg.divmod_(a, b)   # one-shot: the next division is unperturbed
    assert np.array_equal(q2, wq) and np.array_equal(r2, wr)


@pytest.mark.parametrize("mul", [2, 3, 6, 1000])
def test_reciprocal_overshoot_branch(pg, oracle, mul):
    """A Newton seed `mul` times too large (test hook) takes the halving
    branch of div_newton.cpp:57-62, with the unsigned difference saturating
    at zero when halving is not enough; the reciprocal and the quotient are
    still exact."""
    b = W.operand(3000, 5)
    bits = 3000 * 64 + 5000
    print("case overshoot", mul, flush=True)
    pg.debug_perturb(1, mul)
    assert np.array_equal(pg.recip(b, bits), oracle.recip(b, bits))
    a = W.operand(9000, 6)
    wq, wr = oracle.divmod(a, b)
    pg.debug_perturb(1, mul)
    q, r = pg.divmod_(a, b)
    assert np.array_equal(q, wq) and np.array_equal(r, wr)


def test_large_division_perturbed(pg, golden_div):
    """pi's shape (2^21 / 2^20 limbs) with the estimate 1000 units low:
    the GMP digests still match."""
    c = [c for c in golden_div["div"] if c["an"] == 1 << 21][0]
    a, b = W.pair(c["an"], c["bn"], "uniform", c["k"])
    pg.debug_perturb(0, -1000)
    q, r = pg.divmod_(a, b)
    assert sha(q) == c["q_sha256"] and sha(r) == c["r_sha256"]
