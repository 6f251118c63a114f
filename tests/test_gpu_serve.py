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
