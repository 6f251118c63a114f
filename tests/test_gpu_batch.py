"""GPU parity of the dispatcher paths that need more than one transform:
small products, blocked (unbalanced) products, and batches of uneven
products in one submission (plain host API and planned device batches
replayed from a hipGraph)."""
import hashlib
import json

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


@pytest.mark.parametrize("shape", [(1, 1), (1, 500), (5, 300), (24, 24), (25, 192), (192, 192), (30, 100),
                                   (193, 1000), (500, 7000), (25, 50000), (300, 100), (1000, 2100),
                                   (2049, 9000), (7000, 20000), (6145, 30000), (200, 100000)])
@pytest.mark.parametrize("dist", ["uniform", "ones"])
def test_auto_vs_oracle(pg, oracle, shape, dist):
    """mul_into's full dispatch (schoolbook, blocked, NTT, FFT) on the device."""
    an, bn = shape
    a, b = W.pair(an, bn, dist, k=an + 7)
    print("case", an, bn, dist, flush=True)
    got = pg.mul(a, b)
    assert np.array_equal(got, oracle.mul_into(a, b))


def test_reference_all_mul_eq_grid(pg, golden):
    """test_prec.cpp:429-478: sizes 32..2048 x ratios 1..3 (32-bit LCG limbs)."""
    for e in golden["all_mul_eq"]:
        a = W.reference_pattern32(e["an"], e["seed"])
        b = W.reference_pattern32(e["bn"], e["seed"] + 1009)
        assert digest(pg.mul(a, b)) == e["sha256"], e


def test_bench_gmp_rows(pg, golden):
    """bench_gmp.cpp:104-129 rows (64-bit LCG limbs) through mul_into."""
    for e in golden["bench_gmp"]:
        a = W.reference_pattern(e["an"], 11 + e["p"])
        b = W.reference_pattern(e["bn"], 19 + e["p"])
        assert digest(pg.mul(a, b)) == e["sha256"], e


def test_batch_mixed_equals_alone(pg, oracle):
    """Every product of a mixed batch equals the same product alone."""
    shapes = [(30, 100), (193, 193), (500, 700), (1000, 2500), (7000, 7000), (100, 50), (600, 6000),
              (8000, 3000), (20000, 9000), (5, 70000), (65536, 65536), (300, 300)]
    pairs = [W.pair(an, bn, "uniform", k=i) for i, (an, bn) in enumerate(shapes)]
    res, errs = pg.mul_batch(pairs, want_round=True)
    for (a, b), r, e in zip(pairs, res, errs):
        alone = pg.mul(a, b)
        assert np.array_equal(r, alone)
        assert np.array_equal(r, oracle.mul_into(a, b))
        assert 0.0 <= e < 0.125


def test_uneven_batch_4096_golden(pg, golden):
    """The uneven batch: 4096 products, sizes log-uniform in [193, 65536],
    30% unbalanced up to 4:1; digest of the per-product digests."""
    shapes = W.batch_shapes()
    assert hashlib.sha256(json.dumps(shapes).encode()).hexdigest() == golden["batch"]["shapes_sha256"]
    pairs = [W.pair(an, bn, "uniform", k) for k, (an, bn) in enumerate(shapes)]
    res = pg.mul_batch(pairs)
    per = [digest(r) for r in res]
    assert per[:64] == golden["batch"]["first"]
    assert hashlib.sha256("".join(per).encode()).hexdigest() == golden["batch"]["digest_of_digests"]


def test_device_batch_graph_replays(pg, oracle):
    """A planned device batch is captured into a hipGraph on its first run and
    replayed; every replay, with eager work and syncs in between, equals the
    eager result."""
    shapes = [(300, 700), (5000, 5000), (9000, 2500), (20, 4000), (70000, 70000)]
    bufs, trip, want = [], [], []
    for i, (an, bn) in enumerate(shapes):
        a, b = W.pair(an, bn, "uniform", k=100 + i)
        da, db, dr = pg.DevBuf.of(a), pg.DevBuf.of(b), pg.DevBuf(an + bn)
        bufs += [da, db, dr]
        trip.append((dr, da, an, db, bn))
        want.append(oracle.mul_into(a, b))
    batch = pg.DevBatch(trip)
    for rep in range(3):
        for (dr, _, an, _, bn) in trip:            # clobber the outputs between replays
            dr.upload(np.full(an + bn, 0xAB, np.uint64))
        batch.run()
        pg.sync()
        for (dr, _, an, _, bn), w in zip(trip, want):
            got = dr.download(an + bn)
            assert np.array_equal(got[: len(w)], w) and not got[len(w):].any(), rep
        x, y = W.pair(3000, 3000, "uniform", k=rep)   # eager work between replays
        assert np.array_equal(pg.mul(x, y), oracle.mul_into(x, y))
    batch.free()


def test_batch_plan_table(pg, oracle):
    """A batched launch of 2^21 transform points or more takes the batch plan
    table (pg_ntt_batch_plan_set; -1: the single-product plan).  Every plan
    variant of one such bucket (128 products, each a 2^14-point transform)
    gives the oracle's products, for tiles of 16 to 64 columns (packed limb
    sums in pieces of 8 limbs) and product lengths that are not whole pieces;
    an out-of-range variant is refused and leaves the table as it was."""
    single, bt = pg.ntt_plans(), pg.ntt_batch_plans()
    rng = np.random.default_rng(77)
    shapes = [(int(rng.integers(2100, 4097)), int(rng.integers(2100, 4097))) for _ in range(128)]
    shapes[0] = (4096, 4096)
    trip, want, bufs = [], [], []
    for i, (an, bn) in enumerate(shapes):
        a, b = W.pair(an, bn, "uniform", k=500 + i)
        da, db, dr = pg.DevBuf.of(a), pg.DevBuf.of(b), pg.DevBuf(an + bn)
        bufs += [da, db, dr]
        trip.append((dr, da, an, db, bn))
        want.append(oracle.mul_into(a, b))
    try:
        for v in (-1, 0, 3, 7, 13, 15):
            t = list(bt)
            t[14] = v
            pg.set_ntt_batch_plans(t)
            assert pg.ntt_batch_plans()[14] == v
            batch = pg.DevBatch(trip)
            for dr, _, an, _, bn in trip:
                dr.upload(np.full(an + bn, 0xEF, np.uint64))
            batch.run()
            pg.sync()
            for (dr, _, an, _, bn), w in zip(trip, want):
                got = dr.download(an + bn)
                assert np.array_equal(got[: len(w)], w) and not got[len(w):].any(), (v, an, bn)
            batch.free()
        before = pg.ntt_batch_plans()
        with pytest.raises(pg.PgError):
            pg.set_ntt_batch_plans([16] * 27)
        assert pg.ntt_batch_plans() == before
    finally:
        pg.set_ntt_plans(single)
        pg.set_ntt_batch_plans(bt)


def test_shared_operand_spectra(pg, oracle):
    """Products of one device batch that share an operand (same buffer and
    length, same transform length) transform it once: a in both operand
    slots, a square, and the short operand of a blocked product whose last
    block lands in the same transform length.  Every product, on the eager
    run and on two graph replays, equals the oracle's."""
    a, b = W.pair(3000, 3000, "uniform", k=410)
    c, big = W.pair(2900, 9000, "uniform", k=411)
    da, db, dc, dbig = pg.DevBuf.of(a), pg.DevBuf.of(b), pg.DevBuf.of(c), pg.DevBuf.of(big)
    ops = {"a": (da, a), "b": (db, b), "c": (dc, c), "big": (dbig, big)}
    pairs = [("a", "b"), ("a", "c"), ("c", "a"), ("a", "a"), ("b", "a"), ("big", "a"), ("c", "c")]
    trip, want, outs = [], [], []
    for x, y in pairs:
        (dx, hx), (dy, hy) = ops[x], ops[y]
        dr = pg.DevBuf(len(hx) + len(hy))
        outs.append(dr)
        trip.append((dr, dx, len(hx), dy, len(hy)))
        want.append(oracle.mul_into(hx, hy))
    batch = pg.DevBatch(trip)
    for rep in range(3):
        for dr, _, an, _, bn in trip:
            dr.upload(np.full(an + bn, 0xCD, np.uint64))
        batch.run()
        pg.sync()
        for (dr, _, an, _, bn), w, pq in zip(trip, want, pairs):
            got = dr.download(an + bn)
            assert np.array_equal(got[: len(w)], w) and not got[len(w):].any(), (rep, pq)
    batch.free()


def test_square_transforms_once(pg, oracle):
    """A square through pg_mul_dev (one buffer in both operand slots) reads
    its second operand's spectra from the first's, at sizes on every plan
    geometry class."""
    for n in [200, 3000, 40000, 300000]:
        a = W.operand(n, 4200 + n)
        da, dr = pg.DevBuf.of(a), pg.DevBuf(2 * n)
        rl, _ = pg.mul_dev(dr, da, n, da, n)
        assert np.array_equal(dr.download(rl), oracle.mul_into(a, a)), n


def test_staging_images_recycled(pg):
    """Small and unbalanced products upload page-locked descriptor images;
    every synchronous call recycles them, so a long run of pg_mul calls
    keeps the staging memory bounded (held 0 after each call, spare at most
    256 MiB) and the products stay correct."""
    cap = 256 << 20
    for k in range(300):
        an, bn = [(40, 30), (30, 20000), (200, 900), (5, 7000)][k % 4]
        a, b = W.pair(an, bn, "uniform", 9000 + k)
        r = pg.mul(a, b)
        if k % 37 == 0:
            ref = pg.mul(a, b, path=pg.PATH_NTT)
            assert np.array_equal(r, ref)
        held, spare = pg.host_staging()
        assert held == 0, (k, held)
        assert spare <= cap, (k, spare)


def test_uneven_batch_4096_graph_replays(pg, golden):
    """The whole 4096-product batch as a planned device batch
    (pg_batch_plan / pg_batch_run): captured into a hipGraph on its first
    run and replayed twice, with the outputs clobbered and eager work in
    between; every run matches the golden digest of digests."""
    shapes = W.batch_shapes()
    trip, bufs = [], []
    for k, (an, bn) in enumerate(shapes):
        a, b = W.pair(an, bn, "uniform", k)
        da, db, dr = pg.DevBuf.of(a), pg.DevBuf.of(b), pg.DevBuf(an + bn)
        bufs += [da, db, dr]
        trip.append((dr, da, an, db, bn))
    batch = pg.DevBatch(trip)
    try:
        for rep in range(3):
            for (dr, _, an, _, bn) in trip:
                dr.upload(np.full(an + bn, 0xCD, np.uint64))
            print("case batch run", rep, flush=True)
            batch.run()
            pg.sync()
            per = []
            for (dr, _, an, _, bn) in trip:
                x = dr.download(an + bn)
                n = len(x)
                while n and x[n - 1] == 0:
                    n -= 1
                per.append(digest(x[:n]))
            assert per[:64] == golden["batch"]["first"], rep
            assert hashlib.sha256("".join(per).encode()).hexdigest() == golden["batch"]["digest_of_digests"], rep
            x, y = W.pair(5000, 3000, "uniform", k=rep)   # eager work between replays
            assert np.array_equal(pg.mul(x, y), pg.mul(x, y, path=pg.PATH_NTT))
    finally:
        batch.free()
        for d in bufs:
            d.free()
