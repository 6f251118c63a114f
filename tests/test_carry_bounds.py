"""Index bounds of the carry kernels (pg_carry.hip: k_carry_spec, k_carry_fix)
on the CPU.

The kernels' global accesses are replayed index by index for launch shapes
built as the host builds them (lin_run and Batch::build: a node's chunks
start at chunk0 = the sum of the chunks before it, and a node covers
ceil(max(m (+1 with a flag word), out_len) / CHUNK) chunks): every chunk,
wave, lane and iteration, with the kernel constants (CHUNK = 2048 limbs,
8 groups of 64 per wave, 4 waves).  Each read and write must fall inside its
array: the limb sums (m entries, 2 m for half-limb parts), the two layers of
a blocked node (m entries each), the linear operands (xn / yn), the output
(out_len), the chunk map and aggregates (nchunks); the fix-up's look-back
reads only aggregates below its chunk and ends (at the latest at its node's
first chunk, which the first pass always resolves).  Node-table launches
read plain whole-limb nodes with the product form's clamped loads, and
blocked nodes with both layers' clamped loads at every lane's limb plus the
limb below the wave.

Linear nodes are read with both operands' word pairs at every lane's limb
(shift_words; index 0 where a word lies outside the operand) plus the limb
below the wave.

The blocked layout itself is checked on values: block k of a blocked
multiply (Batch::plan) is summed into layer k & 1 at limb k B, the blocks of
a layer never overlap, and limb j's sum read through blocked_cover equals the
sum of the block products covering j, whatever the unwritten words hold; the
loader's lanes (the high word of limb j - 1 from the lane below, lane 63 of
the previous iteration, or the limb below the wave) give the same t, g, p as
limb-by-limb arithmetic.
"""
import random

import numpy as np
import pytest

CHUNK, ITERS, WAVES = 2048, 8, 4          # pg_kernels.h, pg_carry.hip
PLAIN, BLOCKED, LIN = 0, 1, 2


class Node:
    def __init__(self, kind, m, out_len, parts=1, B=0, blen=0, nblocks=0, xn=0, xsh=0, yn=0, ysh=0, flag=False,
                 pack_log=0):
        self.kind, self.m, self.out_len, self.parts = kind, m, out_len, parts
        self.pack_log = pack_log
        self.B, self.blen, self.nblocks = B, blen, nblocks
        self.xn, self.xsh, self.yn, self.ysh, self.flag = xn, xsh, yn, ysh, flag
        self.chunk0 = 0


def ceil_chunks(v):
    return (v + CHUNK - 1) // CHUNK


def chunks_lin(n):       # pg_dint.hip node_chunks (lin_run)
    return max(1, ceil_chunks(max(n.m + (1 if n.flag else 0), n.out_len)))


def chunks_batch(n):     # pg_batch.hip Batch::build, and the single plain carry of pg_runtime.hip
    return ceil_chunks(max(n.m, n.out_len))


def launch(nodes, chunks=chunks_lin):
    """chunk0 of every node and the chunk map, as the host writes them."""
    cmap, c = [], 0
    for i, n in enumerate(nodes):
        n.chunk0 = c
        k = chunks(n)
        cmap += [i] * k
        c += k
    return cmap


def shifted_reads(n, sh, j):
    """Indices shifted_limb(v, n, sh, j) reads (pg_carry.hip)."""
    out = []
    if sh >= 0:
        i = j + (sh >> 6)
        if i < n:
            out.append(i)
        if sh & 63 and i + 1 < n:
            out.append(i + 1)
    else:
        q, r = (-sh) >> 6, (-sh) & 63
        if j >= q and j - q < n:
            out.append(j - q)
        if r and j >= q + 1 and j - q - 1 < n:
            out.append(j - q - 1)
    return out


def limb_sum_reads(N, j):
    """(array, index) pairs limb_sum(N, j) reads."""
    if N.kind == LIN:
        return [("x", i) for i in shifted_reads(N.xn, N.xsh, j)] + [("y", i) for i in shifted_reads(N.yn, N.ysh, j)]
    if N.kind == PLAIN:
        if N.parts == 1:
            return [("sums", j)]
        return [("sums", 2 * j), ("sums", 2 * j + 1)]
    c0, c1 = blocked_cover(N.B, N.blen, N.nblocks, j)
    return ([("l0", j)] if c0 else []) + ([("l1", j)] if c1 else [])


def blocked_cover(B, blen, nblocks, j):
    """blocked_cover (pg_carry.hip): the layers whose block covers limb j."""
    k, r = divmod(j, B)
    ck = k < nblocks
    cp = k >= 1 and (not ck or r < blen - B)
    return (cp, ck) if k & 1 else (ck, cp)


def check_node_access(N, arr, i):
    size = {"x": N.xn, "y": N.yn, "sums": N.m * N.parts, "out": N.out_len, "out_hi": N.out_len,
            "aux": N.m >> N.pack_log, "l0": N.m, "l1": N.m}.get(arr)
    assert 0 <= i < size, (arr, i, size)


def blocked_wave_reads(N, base0, iters):
    """limbs_t_blocked for the wave whose lane l starts at limb base0 + l:
    both layers at min(j, m - 1) for every lane and iteration, and the limb
    below the wave (blocked_sum) when it is in [0, m)."""
    out = []
    if N.m:
        for it in range(iters):
            for lane in range(64):
                jl = min(base0 + lane + 64 * it, N.m - 1)
                out += [("l0", jl), ("l1", jl)]
    jb = base0 - 1
    if 0 <= jb < N.m:
        out += limb_sum_reads(N, jb)
    return out


def shift_words(n, sh, j):
    """shift_words (pg_carry.hip): the indices of the two words shifted_limb
    combines, -1 outside [0, n)."""
    if sh >= 0:
        ia = j + (sh >> 6)
        ib = ia + 1
    else:
        ia = j - ((-sh) >> 6)
        ib = ia - 1
    return (ia if 0 <= ia < n else -1), (ib if 0 <= ib < n else -1)


def lin_wave_reads(N, base0, iters):
    """limbs_t_lin for the wave whose lane l starts at limb base0 + l: both
    operands' two words at every lane and iteration (index 0 where the word
    is outside the operand), and limb_sum of the limb below the wave."""
    out = []
    for it in range(iters):
        for lane in range(64):
            j = base0 + lane + 64 * it
            for arr, n, sh in (("x", N.xn, N.xsh), ("y", N.yn, N.ysh)):
                if n:
                    ia, ib = shift_words(n, sh, j)
                    out += [(arr, max(ia, 0)), (arr, max(ib, 0))]
    jb = base0 - 1
    if 0 <= jb < N.m:
        out += limb_sum_reads(N, jb)
    return out


def t_reads(N, j, plain):
    """Reads of limbs_t_plain (the product launch, and plain whole-limb nodes
    of node-table launches) or limb_t for limb j of the window."""
    if plain or (N.kind == PLAIN and N.parts == 1):
        # limbs_t_plain: indices clamped into [0, m) before the loads
        if not N.m:
            return []
        jl = 0 if j < 0 else (j if j < N.m else N.m - 1)
        if N.pack_log:   # limbs_t_packed: the limb word and its piece's (or the previous piece's) word
            pc, q = jl >> N.pack_log, jl & ((1 << N.pack_log) - 1)
            return [("sums", jl), ("aux", pc if (q or pc == 0) else pc - 1)]
        jh = 0 if j < 1 else (j - 1 if j - 1 < N.m else N.m - 1)
        return [("sums", jl), ("sums", jh)]
    if j < 0:
        return []
    out = []
    if 1 <= j <= N.m:
        out += limb_sum_reads(N, j - 1)
    if j < N.m:
        out += limb_sum_reads(N, j)
    return out


def writes(N, j):
    return [("out", j), ("out_hi", j)] if j < N.out_len else []


def spec_accesses(nodes, cmap, plain, reads=t_reads):
    for chunk in range(len(cmap)):
        assert 0 <= cmap[chunk] < len(nodes)
        N = nodes[cmap[chunk]]
        start = chunk == 0 if plain else N.chunk0 == chunk
        cbase = (chunk - N.chunk0) * CHUNK
        assert cbase >= 0
        for wave in range(WAVES):
            base = cbase + wave * 64 * ITERS + np.arange(64)
            window = wave == 0 and not start
            b0 = base - 64 if window else base
            if N.kind in (BLOCKED, LIN) and not plain and reads is t_reads:
                wave_reads = blocked_wave_reads if N.kind == BLOCKED else lin_wave_reads
                for a, i in wave_reads(N, int(b0[0]), ITERS + (1 if window else 0)):
                    check_node_access(N, a, i)
            else:
                for it in range(ITERS + (1 if window else 0)):
                    for j in (b0 + 64 * it).tolist():
                        for a, i in reads(N, j, plain):
                            check_node_access(N, a, i)
            for it in range(ITERS):
                for j in (base + 64 * it).tolist():
                    for a, i in writes(N, j):
                        check_node_access(N, a, i)


def fix_accesses(nodes, cmap, agg_resolved, plain):
    """k_carry_fix: the look-back reads agg[i] only for 0 <= i < chunk and
    ends at the latest at the node's first chunk; the ripple reads and
    writes the chunk's own limbs."""
    nch = len(cmap)
    for blk in range((nch + WAVES - 1) // WAVES):
        for wave in range(WAVES):
            chunk = blk * WAVES + wave
            if chunk >= nch or agg_resolved[chunk]:
                continue
            N = nodes[cmap[chunk]]
            cbase = (chunk - N.chunk0) * CHUNK
            if cbase >= N.out_len and not (N.flag and cbase <= N.m < cbase + CHUNK):
                continue   # stores nothing: k_carry_fix returns before its look-back
            top = chunk - 1
            while True:
                idx = [top - lane for lane in range(64)]
                assert all(i < chunk for i in idx)
                # read only where i >= 0; lanes below 0 decide (no chunk below)
                dec = [i < 0 or agg_resolved[i] for i in idx]
                if any(dec):
                    first = idx[dec.index(True)]
                    assert first >= N.chunk0, (chunk, first, N.chunk0)
                    break
                top -= 64
            cbase = (chunk - N.chunk0) * CHUNK
            if N.kind in (BLOCKED, LIN) and not plain:
                wave_reads = blocked_wave_reads if N.kind == BLOCKED else lin_wave_reads
                for grp in range(CHUNK // 64):
                    for a, i in wave_reads(N, cbase + 64 * grp, 1):
                        check_node_access(N, a, i)
            for j in range(cbase, cbase + CHUNK):
                reads = [] if (N.kind in (BLOCKED, LIN) and not plain) else t_reads(N, j, plain)
                for a, i in reads + writes(N, j):
                    check_node_access(N, a, i)


def worst_case_resolution(nodes, cmap):
    """Only node starts resolved: the longest look-backs the fix-up can take."""
    return [nodes[cmap[c]].chunk0 == c for c in range(len(cmap))]


def check_launch(nodes, cmap, plain):
    spec_accesses(nodes, cmap, plain)
    fix_accesses(nodes, cmap, worst_case_resolution(nodes, cmap), plain)


@pytest.mark.parametrize("m,out_len", [(1, 1), (1, 2), (2047, 2048), (2048, 2048), (2048, 2049), (4096, 4097),
                                       (5000, 4999), (6143, 6144), (10240, 10241), (3, 9000)])
def test_single_plain_launch(m, out_len):
    # the product's carry: one plain node with whole-limb sums, no node table
    nodes = [Node(PLAIN, m, out_len)]
    check_launch(nodes, launch(nodes, chunks_batch), plain=True)


@pytest.mark.parametrize("pack_log,m", [(1, 2), (2, 4), (2, 2048), (2, 4096), (3, 8), (3, 2048 * 3), (1, 10240)])
def test_packed_product_launch(pack_log, m):
    # the product's carry over packed sums (m limbs of sums, out_len <= m);
    # the host packs only when the pieces tile the sums (m a multiple of 2^pack_log)
    assert m % (1 << pack_log) == 0
    for out_len in (m, m - 1):
        nodes = [Node(PLAIN, m, out_len, pack_log=pack_log)]
        check_launch(nodes, launch(nodes, chunks_batch), plain=True)


def random_lin(rng, m):
    return Node(LIN, m, rng.choice([m, max(1, m - 1)]), xn=rng.randrange(0, m + 3), yn=rng.randrange(0, m + 3),
                xsh=rng.choice([0, 1, 63, 64, 65, 200, -1, -64, -130]), ysh=rng.choice([0, 5, -5, 128, -128]),
                flag=rng.random() < 0.5)


def random_batch_node(rng, m):
    if rng.random() < 0.6:
        return Node(PLAIN, m, m + rng.choice([0, 1, 40]), parts=rng.choice([1, 2]))
    # mul_unbalanced_into's blocks: B = 2 sn limbs of the longer operand, each
    # block product pn + sn long (pg_batch.hip, blocked plan)
    sn = rng.randrange(1, 800)
    ln = rng.randrange(sn, 6 * sn + 1)
    B = 2 * sn
    plen = ln + sn
    return Node(BLOCKED, plen, plen + rng.choice([0, 0, 1]), B=B, blen=B + sn, nblocks=(ln + B - 1) // B)


SIZES = [1, 2, 63, 64, 65, 2047, 2048, 2049, 4095]


def test_linear_launches():
    rng = random.Random(7)
    for _ in range(10):
        nodes = [random_lin(rng, rng.choice(SIZES + [rng.randrange(1, 7000)])) for _ in range(rng.randrange(1, 5))]
        check_launch(nodes, launch(nodes, chunks_lin), plain=False)


def test_batch_levels():
    rng = random.Random(11)
    for _ in range(10):
        nodes = [random_batch_node(rng, rng.choice(SIZES + [rng.randrange(1, 7000)]))
                 for _ in range(rng.randrange(1, 5))]
        check_launch(nodes, launch(nodes, chunks_batch), plain=False)


def test_model_catches_an_unclamped_window():
    """The bound check has teeth: limb loads without the clamp of
    limbs_t_plain (a plausible slip) read sums[-1] at limb 0 and sums[m] and
    beyond in the chunk that holds only the top word and out_len."""
    def unclamped(N, j, plain):
        return [("sums", j), ("sums", j - 1)]
    nodes = [Node(PLAIN, 4096, 4097)]
    with pytest.raises(AssertionError):
        spec_accesses(nodes, launch(nodes, chunks_batch), True, reads=unclamped)


M64 = (1 << 64) - 1


def blocked_layers(rng, ln, sn):
    """Batch::plan's blocked layout with random block products: block k
    (pn + sn limbs of 128-bit sums below 2^121) summed into layer k & 1 at
    limb k B; unwritten words hold garbage.  Returns the layers, the node and
    the exact limb sums (the blocks covering each limb added)."""
    B = 2 * sn
    m = ln + sn
    nblocks = (ln + B - 1) // B
    layers = [[rng.getrandbits(128) | (1 << 127) for _ in range(m)] for _ in range(2)]
    written = [set(), set()]
    exact = [0] * m
    for off in range(0, ln, B):
        k = off // B
        pn = min(B, ln - off)
        for i in range(pn + sn):
            v = rng.getrandbits(121)
            assert off + i not in written[k & 1], "blocks of one layer overlap"
            written[k & 1].add(off + i)
            layers[k & 1][off + i] = v
            exact[off + i] += v
    return layers, Node(BLOCKED, m, m, B=B, blen=B + sn, nblocks=nblocks), exact


def layer_sum(layers, N, j, cover=blocked_cover):
    c0, c1 = cover(N.B, N.blen, N.nblocks, j)
    return (layers[0][j] if c0 else 0) + (layers[1][j] if c1 else 0)


@pytest.mark.parametrize("ln,sn", [(3, 1), (4, 2), (5, 2), (100, 50), (101, 50), (149, 50), (150, 50),
                                   (199, 50), (200, 50), (201, 50), (4000, 1000), (3999, 1000), (4001, 999)])
def test_blocked_layers_sum(ln, sn):
    """Every limb's sum read through blocked_cover equals the sum of the block
    products that cover it (last block full or short, one block or many)."""
    rng = random.Random(ln * 7 + sn)
    layers, N, exact = blocked_layers(rng, ln, sn)
    for j in range(N.m):
        assert layer_sum(layers, N, j) == exact[j], j


def test_blocked_layers_catch_an_off_by_one():
    """Covering block k - 1 while r <= blen - B (one limb too far) reads a
    word of the next block of that layer's gap: the sums then differ."""
    def cover_slip(B, blen, nblocks, j):
        k, r = divmod(j, B)
        ck = k < nblocks
        cp = k >= 1 and (not ck or r <= blen - B)
        return (cp, ck) if k & 1 else (ck, cp)
    rng = random.Random(3)
    layers, N, exact = blocked_layers(rng, 400, 50)
    assert any(layer_sum(layers, N, j, cover_slip) != exact[j] for j in range(N.m))


def t_g_p(S, j):
    """Limb-by-limb t_j = lo_j + hi_(j-1) (mod 2^64), generate and propagate,
    of limb sums S (the top word t_m = hi_(m-1))."""
    m = len(S)
    hprev = (S[j - 1] >> 64) if 1 <= j <= m else 0
    if 0 <= j < m:
        s = (S[j] & M64) + hprev
        t = s & M64
        return t, s >> 64, int(t == M64)
    return (hprev if j == m else 0), 0, 0


def blocked_wave(layers, N, base0, iters):
    """limbs_t_blocked's lanes for the wave at limb base0 + lane: each lane's
    (lo, hi) from the covered layer words; limb j - 1's high word from the
    lane below, lane 0 from lane 63 of the previous iteration or, first, the
    limb below the wave."""
    m = N.m
    jb = base0 - 1
    below = (layer_sum(layers, N, jb) >> 64) if 0 <= jb < m else 0
    last = below
    out = []
    for it in range(iters):
        los, his = [], []
        for lane in range(64):
            j = base0 + lane + 64 * it
            if j < m:
                c0, c1 = blocked_cover(N.B, N.blen, N.nblocks, j)
                w0 = layers[0][min(j, m - 1)] if c0 else 0
                w1 = layers[1][min(j, m - 1)] if c1 else 0
                x0, x1 = w0 & M64, w1 & M64
                lo = (x0 + x1) & M64
                hi = ((w0 >> 64) + (w1 >> 64) + ((x0 + x1) >> 64)) & M64
            else:
                lo, hi = 0, 0
            los.append(lo)
            his.append(hi)
        for lane in range(64):
            j = base0 + lane + 64 * it
            hprev = his[lane - 1] if lane else last
            if j < m:
                s = los[lane] + hprev
                out.append((j, (s & M64, s >> 64, int((s & M64) == M64))))
            else:
                out.append((j, (hprev if j == m else 0, 0, 0)))
        last = his[63]
    return out


@pytest.mark.parametrize("ln,sn", [(150, 50), (4001, 999), (5000, 1200), (300, 2)])
def test_blocked_loader_lanes(ln, sn):
    """The loader's cross-lane high words give limb-by-limb t, g, p for every
    wave of every chunk (windows included), up to limb m + 64."""
    rng = random.Random(ln + sn)
    layers, N, exact = blocked_layers(rng, ln, sn)
    for base0 in list(range(0, N.m + 64, 512)) + [CHUNK - 64, N.m - 3]:
        if base0 < 0:
            continue
        for j, tgp in blocked_wave(layers, N, base0, ITERS + 1):
            assert tgp == t_g_p(exact, j), (base0, j)


def shifted_limb(v, sh, j):
    """shifted_limb (pg_carry.hip) from the value: limb j of v >> sh (sh >= 0)
    or v << -sh."""
    x = to_int(v)
    x = x >> sh if sh >= 0 else x << (-sh)
    return (x >> (64 * j)) & M64 if j >= 0 else 0


def to_int(limbs):
    return sum(int(w) << (64 * i) for i, w in enumerate(limbs))


def lin_sums(x, xsh, xmul, y, ysh, yneg, pow2, cond, m):
    """Limb sums S_j = lo_j + 2^64 hi_j of a linear node, from its definition."""
    S = []
    for j in range(m):
        s = shifted_limb(x, xsh, j) * xmul
        if pow2 >= 0 and j == pow2 >> 6:
            s += 1 << (pow2 & 63)
        if cond:
            yv = shifted_limb(y, ysh, j) if y else 0
            s += (~yv & M64) if yneg else yv
        S.append(s)
    return S


def lin_wave(x, xsh, xmul, y, ysh, yneg, pow2, cond, m, base0, iters):
    """limbs_t_lin's lanes: words by shift_words / shift_combine, sums per
    lane, then lanes_t (limb j - 1's high word from the lane below)."""
    def words(v, sh, j):
        n = len(v) if v else 0
        ia, ib = shift_words(n, sh, j)
        a = v[ia] if (n and ia >= 0) else 0
        b = v[ib] if (n and ib >= 0) else 0
        if sh >= 0:
            r = sh & 63
            return ((a >> r) | (b << (64 - r))) & M64 if r else a
        r = (-sh) & 63
        return ((a << r) | (b >> (64 - r))) & M64 if r else a
    def limb(j):
        if not (0 <= j < m):
            return 0, 0
        s = words(x, xsh, j) * xmul
        if pow2 >= 0 and j == pow2 >> 6:
            s += 1 << (pow2 & 63)
        if cond:
            yv = words(y, ysh, j)
            s += (~yv & M64) if yneg else yv
        return s & M64, s >> 64
    below = limb(base0 - 1)[1]
    last = below
    out = []
    for it in range(iters):
        lo_hi = [limb(base0 + lane + 64 * it) for lane in range(64)]
        for lane in range(64):
            j = base0 + lane + 64 * it
            hprev = lo_hi[lane - 1][1] if lane else last
            if j < m:
                s = lo_hi[lane][0] + hprev
                out.append((j, (s & M64, s >> 64, int((s & M64) == M64))))
            else:
                out.append((j, (hprev if j == m else 0, 0, 0)))
        last = lo_hi[63][1]
    return out


@pytest.mark.parametrize("xsh,ysh", [(0, 0), (1, 63), (64, -64), (130, -5), (-1, 200), (-130, 0), (5000, -70)])
@pytest.mark.parametrize("yneg,cond,ynull", [(0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 0, 0), (1, 0, 0)])
def test_lin_loader_lanes(xsh, ysh, yneg, cond, ynull):
    """limbs_t_lin's shifted words and cross-lane high words give the t, g, p
    of the node's definition xmul X' + 2^pow2 + (yneg ? ~Y' : Y') (the y term
    only when the condition holds; y null reads as zero), up to limb m + 64."""
    rng = random.Random(xsh * 7 + ysh + 100 * yneg + 10 * cond)
    m = 300
    x = [rng.choice([M64, 0, rng.getrandbits(64)]) for _ in range(rng.randrange(1, 400))]
    y = None if ynull else [rng.choice([M64, 0, rng.getrandbits(64)]) for _ in range(rng.randrange(1, 400))]
    xmul = rng.choice([1, 10005, (1 << 59) - 1])
    pow2 = rng.choice([-1, 0, 63, 64 * 150 + 7])
    S = lin_sums(x, xsh, xmul, y, ysh, yneg, pow2, cond, m)
    for base0 in (0, 64, 1984, 256):
        for j, tgp in lin_wave(x, xsh, xmul, y, ysh, yneg, pow2, cond, m, base0, ITERS + 1):
            assert tgp == t_g_p(S, j), (base0, j)
