"""pg_mul_dev_static, the product that allocates nothing (its scratch is one
static device array), against pg_mul_dev with PG_PATH_NTT on the same device
operands: bit-exact, on calls in a row whose second is smaller than the
first (the array keeps the first call's contents), and its size limits."""
import ctypes

import numpy as np
import pytest

from preclib_amd import workloads as W

pytestmark = pytest.mark.gpu

GARBAGE = 0xDEADBEEFCAFEF00D


@pytest.fixture(scope="module")
def pg():
    from preclib_amd import pgpu
    pgpu.init(0)
    return pgpu


def both(pg, an, bn, k):
    a, b = W.pair(an, bn, "uniform", k)
    da, db = pg.DevBuf.of(a), pg.DevBuf.of(b)
    rs, rp = pg.DevBuf(an + bn), pg.DevBuf(an + bn)
    rs.upload(np.full(an + bn, GARBAGE, np.uint64))   # every limb must be written
    print("case static", an, bn, flush=True)
    pg.mul_dev_static(rs, da, an, db, bn)
    pg.mul_dev(rp, da, an, db, bn, path=pg.PATH_NTT, want_len=False)
    pg.sync()
    x, y = rs.download(an + bn), rp.download(an + bn)
    for buf in (da, db, rs, rp):
        buf.free()
    return a, b, x, y


def test_static_equals_pooled_calls_in_a_row(pg, oracle):
    # the largest product the array holds, then smaller ones (unbalanced too)
    for k, (an, bn) in enumerate([(1 << 18, (1 << 18) - 5), (30000, 1000), (1000, 777), (193, 193)]):
        a, b, x, y = both(pg, an, bn, 300 + k)
        assert np.array_equal(x, y), (an, bn)
        if an + bn <= 40000:
            want = oracle.mul_ntt(a, b)
            assert np.array_equal(x[: len(want)], want) and not x[len(want):].any(), (an, bn)


def test_static_size_limits(pg):
    L = pg.load()
    L.pg_mul_dev_static.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                    ctypes.c_size_t]
    a = pg.DevBuf((1 << 19) + 1)
    r = pg.DevBuf((1 << 19) + 2)
    try:
        p, q = L.pg_buf_ptr(a.h), L.pg_buf_ptr(r.h)
        assert L.pg_mul_dev_static(q, p, (1 << 18) + 1, p, 1 << 18) == pg.PG_ERANGE
        assert b"2^19" in L.pg_last_error()
        assert L.pg_mul_dev_static(q, p, 192, p, 100) == pg.PG_ERANGE
        assert b"192" in L.pg_last_error()
        assert L.pg_mul_dev_static(None, p, 300, p, 300) == pg.PG_EINVAL
        assert L.pg_mul_dev_static(q, p, 0, p, 300) == pg.PG_EINVAL
    finally:
        a.free()
        r.free()
