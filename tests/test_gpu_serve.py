"""The continuous-batching server (pg_server_*): every result bit-identical
to the same request alone (pg_mul / pg_div) and to the oracle, under mixes,
admission limits and requests arriving while others run."""
import threading
import time

import numpy as np
import pytest

from preclib_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pg():
    from preclib_amd import pgpu
    pgpu.init(0)
    return pgpu


def mixed_requests(count, seed, hi=1 << 15):
    """(op, a, b) with sizes log-uniform in [1, hi]; divisions with a > b."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        op = 1 if rng.uniform() < 0.3 else 0
        n1 = int(np.exp(rng.uniform(0, np.log(hi))))
        n2 = int(np.exp(rng.uniform(0, np.log(hi))))
        if op == 1:
            n1, n2 = max(n1, n2) + 1, min(n1, n2)
        dist = ["uniform", "ones", "sparse"][k % 3]
        a, b = W.pair(n1, n2, dist, 2000 + k)
        out.append((op, a, b))
    return out


def alone(pg, op, a, b):
    return pg.mul(a, b) if op == 0 else pg.divmod_(a, b)


def same(x, y):
    if isinstance(x, tuple):
        return all(np.array_equal(u, v) for u, v in zip(x, y))
    return np.array_equal(x, y)


def test_mix_equals_alone(pg):
    reqs = mixed_requests(60, 1)
    srv = pg.Server()
    try:
        handles = [srv.submit(op, a, b) for op, a, b in reqs]
        srv.drain()
        for (op, a, b), h in zip(reqs, handles):
            assert same(h.result(), alone(pg, op, a, b)), (op, len(a), len(b))
        st = srv.stats()
        assert st["completed"] == len(reqs)
        # shared rounds: far fewer rounds than request steps
        assert st["rounds"] < len(reqs)
    finally:
        srv.stop()


def test_against_oracle(pg, oracle):
    reqs = mixed_requests(20, 2, hi=3000)
    srv = pg.Server()
    try:
        handles = [srv.submit(op, a, b) for op, a, b in reqs]
        for (op, a, b), h in zip(reqs, handles):
            got = srv.wait(h)
            want = oracle.mul_into(a, b) if op == 0 else oracle.divmod(a, b)
            assert same(got, want)
    finally:
        srv.stop()


def test_edge_requests(pg):
    srv = pg.Server()
    try:
        z = np.zeros(0, np.uint64)
        one = np.array([1], np.uint64)
        a, b = W.pair(40000, 33000, "uniform", 77)
        cases = [(0, z, a), (0, a, z), (1, z, b), (1, a, z), (1, b[:100], a), (0, one, one), (1, a, a), (1, a, one)]
        hs = [srv.submit(op, x, y) for op, x, y in cases]
        srv.drain()
        for (op, x, y), h in zip(cases, hs):
            assert same(h.result(), alone(pg, op, x, y))
    finally:
        srv.stop()


def test_admission_limits(pg):
    # a budget that holds a few requests at a time and 4 slots: everything
    # still completes, never more than 4 in flight
    reqs = mixed_requests(30, 3, hi=1 << 14)
    srv = pg.Server(budget=64 << 20, max_active=4)
    try:
        hs = [srv.submit(op, a, b) for op, a, b in reqs]
        srv.drain()
        for (op, a, b), h in zip(reqs, hs):
            assert same(h.result(), alone(pg, op, a, b))
        assert srv.stats()["max_active"] <= 4
    finally:
        srv.stop()


def test_arrivals_while_running(pg):
    # a submitter thread keeps adding requests while the server works: the
    # late ones are admitted into running rounds and finish correctly
    reqs = mixed_requests(40, 4)
    srv = pg.Server(max_active=16)
    hs = []
    try:
        def feed():
            for op, a, b in reqs:
                hs.append(srv.submit(op, a, b))
                time.sleep(0.002)
        t = threading.Thread(target=feed)
        t.start()
        t.join()
        srv.drain()
        for (op, a, b), h in zip(reqs, hs):
            assert same(h.result(), alone(pg, op, a, b))
            assert h.req.t_submit <= h.req.t_admit <= h.req.t_done
    finally:
        srv.stop()


def test_large_division_requests(pg, golden_div):
    # divisor >= 32768 limbs, dividend above twice the divisor (GMP digests)
    import hashlib
    srv = pg.Server()
    try:
        cs = golden_div["div"][:3]
        pairs = [W.pair(c["an"], c["bn"], "uniform", c["k"]) for c in cs]
        hs = [srv.submit(1, a, b) for a, b in pairs]
        srv.drain()
        for c, h in zip(cs, hs):
            q, r = h.result()
            assert hashlib.sha256(q.astype("<u8").tobytes()).hexdigest() == c["q_sha256"]
            assert hashlib.sha256(r.astype("<u8").tobytes()).hexdigest() == c["r_sha256"]
    finally:
        srv.stop()


MASK64 = (1 << 64) - 1


def _below(r, b):
    """r < b for trimmed limb arrays."""
    if len(r) != len(b):
        return len(r) < len(b)
    d = np.nonzero(r != b)[0]
    return len(d) > 0 and r[d[-1]] < b[d[-1]]


def check_result(oracle, op, a, b, res):
    """A serving result checked without a second multiplier: lengths, the low
    limb, and C = A B (or A = Q B + R with R < B) modulo three primes."""
    M = oracle.MODS3
    ra, rb = oracle.mod_p3(a), oracle.mod_p3(b)
    an, bn = len(a), len(b)
    if op == 0:
        c = res
        assert len(c) in (an + bn - 1, an + bn) and c[-1] != 0, (an, bn, len(c))
        assert int(c[0]) == (int(a[0]) * int(b[0])) & MASK64
        rc = oracle.mod_p3(c)
        assert all(rc[i] == ra[i] * rb[i] % M[i] for i in range(3)), (an, bn)
    else:
        q, r = res
        assert len(q) in (an - bn, an - bn + 1) and q[-1] != 0, (an, bn, len(q))
        assert len(r) == 0 or r[-1] != 0
        assert _below(r, b), (an, bn)
        rq, rr = oracle.mod_p3(q), oracle.mod_p3(r)
        assert all(ra[i] == (rq[i] * rb[i] + rr[i]) % M[i] for i in range(3)), (an, bn)


@pytest.mark.timeout(900)
def test_c5_stream_three_rates(pg, oracle):
    """The serving configuration: 240 requests, 80% multiplies with both sizes
    log-uniform in [256, 2^21] limbs, 20% Newton divisions with divisors of
    at least 32768 limbs and dividends above twice the divisor (uniform
    limbs).  Every request alone through pg_mul / pg_div is checked modulo
    three primes (plus lengths, low limb, remainder bound); then the stream
    runs through the server all at once (saturation) and with Poisson
    arrivals at 0.5x, 0.8x and 1.1x the measured saturated rate, and every
    result of every run must equal the request alone, limb for limb."""
    reqs = W.serve_stream(240)
    assert sum(op for op, _, _ in reqs) >= 30
    for op, a, b in reqs:
        assert 256 <= min(len(a), len(b)) and max(len(a), len(b)) <= 1 << 21
        if op == 1:
            assert len(b) >= 32768 and len(a) > 2 * len(b)
    ref = []
    for op, a, b in reqs:
        res = alone(pg, op, a, b)
        check_result(oracle, op, a, b, res)
        ref.append(res)
    print("c5: %d requests alone checked" % len(reqs), flush=True)

    srv = pg.Server()
    try:
        t0 = time.perf_counter()
        hs = [srv.submit(op, a, b) for op, a, b in reqs]
        srv.drain()
        sat = len(reqs) / (time.perf_counter() - t0)
        for i, h in enumerate(hs):
            assert same(h.result(), ref[i]), ("saturated", i)
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["rounds"] < len(reqs)
    finally:
        srv.stop()
    del hs
    print("c5: saturated %.1f req/s, %d rounds" % (sat, st["rounds"]), flush=True)

    for frac in (0.5, 0.8, 1.1):
        rng = np.random.default_rng(int(frac * 1000))
        gaps = rng.exponential(1.0 / (frac * sat), len(reqs))
        srv = pg.Server()
        hs = [None] * len(reqs)
        try:
            def feed():
                t = time.perf_counter()
                for i, (op, a, b) in enumerate(reqs):
                    t += gaps[i]
                    dt = t - time.perf_counter()
                    if dt > 0:
                        time.sleep(dt)
                    hs[i] = srv.submit(op, a, b)
            th = threading.Thread(target=feed)
            th.start()
            th.join()
            srv.drain()
            lat = np.array([h.latency for h in hs]) * 1e3
            for i, h in enumerate(hs):
                assert same(h.result(), ref[i]), (frac, i)
                assert h.req.t_submit <= h.req.t_admit <= h.req.t_done
            st = srv.stats()
            assert st["completed"] == len(reqs)
            # requests that arrive while others run are admitted beside them
            assert st["max_active"] >= 2
        finally:
            srv.stop()
        print("c5: %.1fx load: p50 %.1f ms, p99 %.1f ms, %d rounds, max %d active"
              % (frac, np.percentile(lat, 50), np.percentile(lat, 99), st["rounds"], st["max_active"]), flush=True)
        del hs
