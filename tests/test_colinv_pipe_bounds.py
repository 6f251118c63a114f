"""CPU bound check of the pipelined inverse column kernel's indexing
(pg_ntt_impl.h: col_inv_wide_pipe, stage_dma, stage_store, fill_inv_tw,
pass_inv_lt), for every (transform length, plan variant) that takes it.

The plan constants come from the kernels' own constexpr functions, printed
by tests/cpp/plan_dump.cpp (built here with hipcc, host side only); the
index arithmetic of the kernels is replayed on them:
  - LDS: transform tile [0, NE), staging tile [NE, 2 NE), three primes' pass
    twiddles [2 NE, 2 NE + 3 TWW), all within the launch's dynamic LDS and
    the 156 KiB limit;
  - every 16-byte DMA lands inside the staging tile, reads 16 aligned bytes
    inside the prime's n residues, and the DMAs cover every staging word
    exactly once with the element the staging read expects there;
  - every twiddle fill reads inside its R x M table and writes inside its
    pass's LDS range, each entry once; every 16-byte twiddle read of a pass
    stays inside that range;
  - the persistent schedule visits every tile of the transform exactly once;
  - the forward column kernel's workgroup (col_threads) covers its tile,
    keeps one column per thread, and its stepped twiddles follow the
    frequencies of its rows (every plan).
Skipped where hipcc is absent."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"
pytestmark = pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc absent")
LDS_LIMIT = 156 * 1024


@pytest.fixture(scope="module")
def plans(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("plans") / "plan_dump")
    subprocess.check_call([HIPCC, "-std=c++17", "-O1", "--offload-arch=gfx950", "-I" + os.path.join(ROOT, "preclib_amd", "csrc"),
                           os.path.join(ROOT, "tests", "cpp", "plan_dump.cpp"), "-o", exe])
    out = subprocess.check_output([exe], text=True).splitlines()
    threads = int(out[0].split("threads=")[1])
    rows = []
    for line in out[1:]:
        v = [int(x) for x in line.split()]
        rows.append(dict(zip("logn v l1 l2 lt rb step d waves pipe lds tww fnt fstep fd".split(), v)))
    return threads, rows


def split_lr(l, i, rb):
    # the pass split of pg_kernels.h, checked below against the TW words the
    # kernel reports
    n = (l + rb - 1) // rb if l else 0
    return 0 if n == 0 else (l // n + (1 if i < l % n else 0) if i < n else 0)


def passes(l1, rb):
    """(R, M, word offset) of the inverse column passes with twiddles."""
    out, logb, off, i = [], l1, 0, 0
    while split_lr(l1, i, rb) > 0:
        lr = split_lr(l1, i, rb)
        m = 1 << (logb - lr)
        if m > 1:
            out.append((1 << lr, m, off))
            off += (1 << lr) * m
        logb -= lr
        i += 1
    return out, off


def test_pipelined_plans_exist(plans):
    threads, rows = plans
    assert threads == 1024
    # the 2^22 x 2^22 product's transform (2^24 points) under its shipped plan
    assert any(r["pipe"] for r in rows if r["logn"] == 24)
    assert all(not r["pipe"] or r["waves"] == 4 for r in rows)


def test_lds_layout(plans):
    threads, rows = plans
    for r in rows:
        if not r["pipe"]:
            continue
        ne = (1 << r["l1"]) << r["lt"]
        tw = passes(r["l1"], r["rb"])
        assert tw[1] == r["tww"], r
        assert r["lds"] == (2 * ne + 3 * r["tww"]) * 4, r
        assert r["lds"] <= LDS_LIMIT, r
        assert r["step"] == 1 and r["lt"] >= 2 and ne % 256 == 0 and ne % threads == 0, r


def test_dma_and_staging(plans):
    threads, rows = plans
    nw = threads // 64
    for r in rows:
        if not r["pipe"]:
            continue
        l1, l2, lt, logn = r["l1"], r["l2"], r["lt"], r["logn"]
        ne, n, T = (1 << l1) << lt, 1 << logn, 1 << lt
        ntiles = (1 << l2) >> lt
        for c0 in {0, (ntiles - 1) * T}:
            got = {}
            for tid in range(threads):
                w, l = tid >> 6, tid & 63
                for k in range((ne // 256 + nw - 1) // nw):
                    chunk = k * nw + w
                    if not (ne % (256 * nw) == 0 or chunk < ne // 256):
                        continue
                    e = chunk * 256 + 4 * l
                    off = ((e >> lt) << l2) + c0 + (e & (T - 1))
                    dst = chunk * 256 + 4 * l            # lane l of the DMA: 16 bytes at 16 l
                    assert 0 <= off and off + 4 <= n and off % 4 == 0, r
                    assert 0 <= dst and dst + 4 <= ne, r
                    for j in range(4):
                        assert dst + j not in got
                        got[dst + j] = off + j
            assert len(got) == ne
            # the staging read of element e (row e >> lt, column c0 + e mod T)
            for e in range(0, ne, max(1, ne // 4096)):
                assert got[e] == ((e >> lt) << l2) + c0 + (e & (T - 1)), r


def test_twiddle_fill_and_reads(plans):
    threads, rows = plans
    for r in rows:
        if not r["pipe"]:
            continue
        tw, total = passes(r["l1"], r["rb"])
        written = set()
        for R, M, off in tw:
            W = R * M
            for tid in range(threads):
                for k in range((W + threads - 1) // threads):
                    i = tid + k * threads
                    if not (W % threads == 0 or i < W):
                        continue
                    src = (i & (R - 1)) * M + i // R       # table entry k1 M + j2, k1 = i mod R, j2 = i / R
                    assert 0 <= src < R * M
                    assert off + i not in written and off + i < total
                    written.add(off + i)
            # pass_inv_lt: j2 < M reads 16-byte groups R j2 + [0, R)
            assert R >= 4 and R % 4 == 0
            for j2 in (0, M - 1):
                assert off + j2 * R + R <= off + W and (off + j2 * R) % 4 == 0
        assert written == set(range(total)), r


@pytest.mark.parametrize("grid", [256, 304, 80, 1])
def test_persistent_schedule(plans, grid):
    threads, rows = plans
    for r in rows:
        if not r["pipe"]:
            continue
        ntiles = (1 << r["l2"]) >> r["lt"]
        g = min(ntiles, grid)
        seen = []
        for b in range(g):
            nsteps = 3 * ((ntiles - b + g - 1) // g)
            for step in range(nsteps):
                i = b + step // 3 * g
                t = i if ntiles & 7 else (i & 7) * (ntiles >> 3) + (i >> 3)   # xcd_tile
                if step % 3 == 0:
                    seen.append(t)
        assert sorted(seen) == list(range(ntiles)), (r, grid)


def freq_of_pos(rho, l1, rb):
    """Frequency of tile row rho after the forward column passes: pass i's
    digit (the i-th from the top of rho) becomes the i-th from the bottom."""
    k, w, logb, i = 0, 0, l1, 0
    while split_lr(l1, i, rb) > 0:
        lr = split_lr(l1, i, rb)
        logb -= lr
        k |= ((rho >> logb) & ((1 << lr) - 1)) << w
        w += lr
        i += 1
    return k


def test_forward_column_workgroups(plans):
    """k_col_fwd's workgroup (col_threads): whole waves, at least one column
    per thread (a thread keeps its column C = c0 + tid mod T), the elements
    e = tid + k NT cover the tile once; where the twiddles are stepped, the
    frequency of a thread's k-th row is the first one's plus k 2^D, so the
    product w = w0 * (omega^(C 2^D))^k is the element's own twiddle."""
    _, rows = plans
    stepped = 0
    for r in rows:
        l1, lt, rb, nt = r["l1"], r["lt"], r["rb"], r["fnt"]
        if l1 == 0:
            continue
        ne, T = (1 << l1) << lt, 1 << lt
        assert nt % 64 == 0 and T <= nt <= 1024 and ne % nt == 0, r
        E = ne // nt
        if not r["fstep"]:
            continue
        stepped += 1
        assert E > 1 and r["fd"] >= 0, r
        for tid in range(nt):
            f0 = freq_of_pos(tid >> lt, l1, rb)
            for k in range(E):
                e = tid + k * nt
                assert e & (T - 1) == tid & (T - 1)
                assert freq_of_pos(e >> lt, l1, rb) == f0 + (k << r["fd"]), (r, tid, k)
    assert stepped > 0
    # short transforms run narrower workgroups; tiles of 2^13 elements and
    # more (the 2^22-limb product's 2^24 points) keep 1024 threads
    assert min(r["fnt"] for r in rows if r["l1"] > 0 and r["logn"] == 12) < 1024
    assert all(r["fnt"] == 1024 for r in rows if r["l1"] + r["lt"] >= 13)
