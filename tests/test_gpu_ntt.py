"""GPU parity of the NTT product path (pg_mul through the C ABI) against the
oracle and the GMP digests; bit-exact, limb for limb."""
import numpy as np
import pytest

from preclib_amd import workloads as W
from tests.oracle_py import digest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pg():
    from preclib_amd import pgpu
    pgpu.init(0)
    return pgpu


def _wl(golden, an, bn, dist):
    return next(e for e in golden["workload"] if e["an"] == an and e["bn"] == bn and e["dist"] == dist)


@pytest.mark.parametrize("shape", [(193, 193), (300, 255), (1000, 1000), (2049, 2039), (4097, 4097),
                                   (6145, 6145), (6145, 12000), (9000, 5000), (16384, 12289)])
@pytest.mark.parametrize("dist", ["uniform", "ones", "sparse"])
@pytest.mark.parametrize("mode", ["wide", "ref"])
def test_ntt_vs_oracle(pg, oracle, shape, dist, mode):
    an, bn = shape
    a, b = W.pair(an, bn, dist, k=an)
    print("case", an, bn, dist, mode, flush=True)
    got = pg.mul(a, b, path=pg.PATH_NTT if mode == "wide" else pg.PATH_NTT_REF)
    want = oracle.mul_ntt(a, b)
    assert len(got) == len(want)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("dist", ["uniform", "ones", "sparse"])
def test_ntt_16384_golden(pg, golden, dist):
    a, b = W.pair(16384, 16384, dist)
    got = pg.mul(a, b)
    e = _wl(golden, 16384, 16384, dist)
    assert len(got) == e["len"] and digest(got) == e["sha256"]


@pytest.mark.parametrize("logn", [20, 22])
@pytest.mark.parametrize("dist", ["uniform", "ones", "sparse"])
@pytest.mark.parametrize("mode", ["wide", "ref"])
def test_ntt_large_golden(pg, golden, logn, dist, mode):
    n = 1 << logn
    a, b = W.pair(n, n, dist)
    print("case", n, dist, mode, flush=True)
    got = pg.mul(a, b, path=pg.PATH_NTT if mode == "wide" else pg.PATH_NTT_REF)
    e = _wl(golden, n, n, dist)
    assert len(got) == e["len"] and digest(got) == e["sha256"]
    if dist == "ones":   # (2^64N - 1)^2 closed form
        assert got[0] == 1 and not got[1:n].any()
        assert got[n] == 0xFFFFFFFFFFFFFFFE and (got[n + 1:] == 0xFFFFFFFFFFFFFFFF).all()


def test_ntt_torture_patterns(pg, golden):
    from tests.golden.gen_golden import TORTURE_PATTERNS, torture_operands
    pats = {p[0]: p for p in TORTURE_PATTERNS}
    for e in golden["torture"]:
        a, b = torture_operands(e["an"], e["bn"], pats[e["pattern"]])
        if max(len(a), len(b)) <= 192:
            continue
        for path in (pg.PATH_NTT, pg.PATH_NTT_REF):
            got = pg.mul(a, b, path=path)
            assert digest(got) == e["sha256"], (e, path)


def test_device_resident_alias(pg, oracle):
    a, b = W.pair(20000, 15000, "uniform", k=7)
    want = oracle.mul_ntt(a, b)
    da = pg.DevBuf(len(a) + len(b))
    da.upload(a)
    db = pg.DevBuf.of(b)
    n, _ = pg.mul_dev(da, da, len(a), db, len(b), path=pg.PATH_NTT)   # r aliases a
    assert n == len(want)
    assert np.array_equal(da.download(n), want)
    da.free()
    db.free()


def test_repeat_is_deterministic(pg):
    a, b = W.pair(50000, 50000, "uniform", k=3)
    r1 = pg.mul(a, b)
    r2 = pg.mul(a, b)
    assert np.array_equal(r1, r2)


@pytest.mark.parametrize("v", list(range(1, 16)))
def test_plan_variants(pg, oracle, golden, v):
    # every autotuner plan variant gives the same product as the oracle and GMP
    import hashlib
    saved = pg.ntt_plans()
    try:
        pg.set_ntt_plans([v] * 27)
        for k, (an, bn) in enumerate([(1000, 1100), (3000, 3000), (7000, 9000), (30000, 30000), (100000, 70000),
                                      (200000, 300000)]):
            a, b = W.pair(an, bn, "uniform", 40 + k)
            assert np.array_equal(pg.mul(a, b, path=pg.PATH_NTT), oracle.mul_ntt(a, b)), (v, an, bn)
        for c in golden["workload"]:
            if c["an"] in (1 << 20, 1 << 22) and c["dist"] in ("uniform", "ones"):
                a, b = W.pair(c["an"], c["bn"], c["dist"])
                r = pg.mul(a, b, path=pg.PATH_NTT)
                assert hashlib.sha256(r.astype("<u8").tobytes()).hexdigest() == c["sha256"], (v, c["an"], c["dist"])
    finally:
        pg.set_ntt_plans(saved)


@pytest.mark.parametrize("n", [193, 1000, 4096, 6144, 20000])
def test_auto_dispatch_squares_and_mid_sizes(pg, oracle, n):
    """The dispatcher sends squares and balanced products above 192 limbs to
    the NTT (no rounding distance reported); each equals the oracle's
    mul_into, including the aliased square precn_sqr would take."""
    a, b = W.pair(n, n, "uniform", k=n)
    print("case", n, flush=True)
    r, err = pg.mul(a, b, want_round=True)
    assert err == 0.0 and np.array_equal(r, oracle.mul_into(a, b))
    s, serr = pg.mul(a, a, want_round=True)
    assert serr == 0.0 and np.array_equal(s, oracle.mul_into(a, a, sqr=True))


@pytest.mark.parametrize("logn", list(range(12, 26)))
def test_shipped_plan_every_length(pg, oracle, logn):
    """One product per wide-mode transform length 2^12 .. 2^25 under the
    shipped plan table (ntt_plans.txt: non-default variants at several of
    these lengths): equal to the reference-decomposition product
    (PG_PATH_NTT_REF, other kernels and digits), to the oracle up to 2^20
    points, and to A B modulo three primes at every length."""
    an, bn = (1 << (logn - 2)) - 3, 1 << (logn - 2)     # 2 (an + bn) base-2^32 digits -> n = 2^logn
    a, b = W.pair(an, bn, "uniform", 700 + logn)
    print("case logn", logn, "variant", pg.ntt_plans()[logn], flush=True)
    r = pg.mul(a, b, path=pg.PATH_NTT)
    assert np.array_equal(r, pg.mul(a, b, path=pg.PATH_NTT_REF))
    M = oracle.MODS3
    ra, rb, rc = oracle.mod_p3(a), oracle.mod_p3(b), oracle.mod_p3(r)
    assert all(rc[i] == ra[i] * rb[i] % M[i] for i in range(3))
    if logn <= 20:
        assert np.array_equal(r, oracle.mul_ntt(a, b))
