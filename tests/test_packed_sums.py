"""Packed limb sums (NttLeaf::pack_log; pg_ntt_impl.h col_inv_wide writes
them, pg_carry.hip limbs_t_packed reads them) against the plain form.

The inverse column kernel holds limb sums S_j = lo_j + 2^64 hi_j in pieces of
PL = min(T / 2, 8) consecutive limbs (half a row of a column tile, or one
16-lane row of a wider tile).  Packed, slo[j] is
t_j = lo_j + hi_(j-1) mod 2^64 for every limb but the first of its piece
(where it is lo_j), and shi[j / PL] holds the high word of the piece's last
limb with bit 56 + q set when limb q's addition carried.  The carry's loader
must recover the same (t_j, generate, propagate) as from the plain sums, and
the top word t_m = hi_(m-1); here both sides are modelled on Python integers,
including the largest high words (2^57 - 1) and additions that carry at
every position of a piece.  The kernel's own write (lanes, DPP row shifts,
which lane writes the piece word) is replayed lane by lane for tiles of 4 to
1024 columns and compared with the format.
"""
import random

import pytest

M64 = (1 << 64) - 1
HI = (1 << 57) - 1


def plain_tgp(lo, hi, j):
    m = len(lo)
    hprev = hi[j - 1] if 1 <= j <= m else 0
    if j < m:
        t = (lo[j] + hprev) & M64
        return t, int(t < lo[j]), int(t == M64)
    return (hprev if j == m else 0), 0, 0


def pack(lo, hi, pl):
    """What col_inv_wide writes: limb q of a piece adds the high word of the
    limb two lanes down (limb q - 1); the last limb writes the piece word."""
    m = len(lo)
    assert m % pl == 0
    slo, shi = [0] * m, [0] * (m // pl)
    for j in range(m):
        q = j % pl
        t = (lo[j] + hi[j - 1]) & M64 if q else lo[j]
        slo[j] = t
        if q == pl - 1:
            w = hi[j]
            assert w <= HI
            for qq in range(1, pl):
                jj = j - (pl - 1) + qq
                g = (lo[jj] + hi[jj - 1]) >> 64
                w |= g << (56 + qq)
            shi[j // pl] = w
    return slo, shi


def packed_tgp(slo, shi, pl, j):
    """limbs_t_packed for one limb (indices clamped as the kernel clamps)."""
    m = len(slo)
    jl = min(max(j, 0), m - 1)
    pc, q = jl // pl, jl % pl
    x = slo[jl]
    ai = pc if (q or pc == 0) else pc - 1
    assert 0 <= ai < len(shi)
    a = shi[ai]
    if 0 <= j < m:
        q = j % pl
        if q:
            t, g = x, (a >> (56 + q)) & 1
        else:
            h = (a & HI) if j else 0
            t = (x + h) & M64
            g = int(t < x)
        return t, g, int(t == M64)
    return ((a & HI) if j == m else 0), 0, 0


def check(lo, hi, pl):
    slo, shi = pack(lo, hi, pl)
    for j in range(len(lo) + 2):
        assert packed_tgp(slo, shi, pl, j) == plain_tgp(lo, hi, j), (pl, j)


@pytest.mark.parametrize("pl", [2, 4, 8])
def test_random_sums(pl):
    rng = random.Random(pl)
    m = 64 * pl
    lo = [rng.getrandbits(64) for _ in range(m)]
    hi = [rng.getrandbits(57) for _ in range(m)]
    check(lo, hi, pl)


@pytest.mark.parametrize("pl", [2, 4, 8])
def test_carries_everywhere(pl):
    """Every addition carries (lo near 2^64, largest high words), every limb
    propagates (t = 2^64 - 1), and the mixes between."""
    m = 16 * pl
    check([M64] * m, [HI] * m, pl)
    check([M64 - HI + 1] * m, [HI] * m, pl)                    # t = 0 with a carry
    check([M64 - 5] * m, [5] * m, pl)                          # t = 2^64 - 1, propagate
    rng = random.Random(9)
    lo = [rng.choice([M64, M64 - 1, 0, 1 << 63]) for _ in range(m)]
    hi = [rng.choice([0, 1, 2, HI]) for _ in range(m)]
    check(lo, hi, pl)


def test_high_word_bound():
    """The high words stay below 2^57: a limb sum is x_even + 2^32 x_odd with
    every coefficient at most 2^25 (2^32 - 1)^2 (the longest operand the NTT
    path accepts has 2^25 base-2^32 digits)."""
    xmax = (1 << 25) * ((1 << 32) - 1) ** 2
    s = xmax + (xmax << 32)
    assert s >> 64 <= HI


def row_shr(vals, s):
    """update_dpp(0, v, row_shr:s): lane l reads lane l - s of its row of 16
    lanes; the first s lanes of a row read 0."""
    return [vals[l - s] if (l & 15) >= s else 0 for l in range(len(vals))]


def kernel_packed_write(lo, hi, lt, l2, nt, seed, p2=None):
    """crt_step's packed write (pg_ntt_impl.h), lane by lane: every tile of
    T = 2^lt columns of every row of a 2^l2-column transform (N1 = 2 rows),
    element e = tid + k nt of a tile on lane tid mod 64 of wave tid / 64;
    limb li = row 2^(l2-1) + C / 2 is held by the even lane of its column
    pair (odd lanes hold other values, which only odd lanes read), pieces of
    P2 = min(T / 2, 8) limbs as sum_pack_log gives them (p2: another size,
    to plant a mistake)."""
    T = 1 << lt
    pl = min(lt, 4) - 1 if p2 is None else p2.bit_length() - 1   # p2: a planted piece size
    p2 = 1 << pl
    n1, n2 = 2, 1 << l2
    ne = n1 * T
    m = n1 * n2 // 2
    assert len(lo) == m and nt % 64 == 0 and nt >= T
    rng = random.Random(seed)
    slo, shi = [None] * m, [None] * (m >> pl)
    for c0 in range(0, n2, T):
        for k in range((ne + nt - 1) // nt):
            for w in range(nt // 64):
                tids = [w * 64 + l for l in range(64)]
                es = [t + k * nt for t in tids]
                C = [c0 + (t & (T - 1)) for t in tids]
                li = [((e >> lt) << (l2 - 1)) + (c >> 1) for e, c in zip(es, C)]
                even = [(t & 1) == 0 for t in tids]
                vlo = [lo[i] if ev and e < ne else rng.getrandbits(64) for i, ev, e in zip(li, even, es)]
                vhi = [hi[i] if ev and e < ne else rng.getrandbits(57) for i, ev, e in zip(li, even, es)]
                q = [(c >> 1) & (p2 - 1) for c in C]
                hp = row_shr(vhi, 2)
                t = [(a + h) & M64 if qq else a for a, h, qq in zip(vlo, hp, q)]
                gb = [(1 << qq) if (qq and tt < a) else 0 for qq, tt, a in zip(q, t, vlo)]
                gb = [g | s for g, s in zip(gb, row_shr(gb, 2))]
                if p2 >= 4:
                    gb = [g | s for g, s in zip(gb, row_shr(gb, 4))]
                if p2 >= 8:
                    gb = [g | s for g, s in zip(gb, row_shr(gb, 8))]
                for l in range(64):
                    if even[l] and es[l] < ne and li[l] < m:
                        slo[li[l]] = t[l]
                        if q[l] == p2 - 1:
                            shi[li[l] >> pl] = vhi[l] | (gb[l] << 56)
    return slo, shi, p2


@pytest.mark.parametrize("lt,l2,nt", [(2, 6, 64), (3, 7, 64), (4, 8, 128), (5, 8, 64), (5, 9, 256),
                                      (6, 8, 128), (6, 9, 1024), (10, 10, 1024)])
def test_kernel_lane_layout(lt, l2, nt):
    """The lanes of crt_step write exactly the packed format of pieces of
    min(T / 2, 8) limbs, for tiles of 4 to 1024 columns (pieces of 8 limbs
    are one 16-lane row from T = 16 on); carries at every position."""
    rng = random.Random(lt * 100 + l2)
    m = 2 * (1 << l2) // 2
    lo = [rng.choice([M64, M64 - 1, 0, rng.getrandbits(64)]) for _ in range(m)]
    hi = [rng.choice([0, HI, rng.getrandbits(57)]) for _ in range(m)]
    slo, shi, p2 = kernel_packed_write(lo, hi, lt, l2, nt, seed=lt)
    want_lo, want_hi = pack(lo, hi, p2)
    assert slo == want_lo and shi == want_hi
    check(lo, hi, p2)


def test_kernel_lane_layout_catches_wide_pieces():
    """The mistake of pieces of T / 2 limbs at T = 32 (16 limbs, two 16-lane
    rows): limb 8 of the piece, at the start of the second row, does not
    receive limb 7's high word, so the write differs from the format."""
    rng = random.Random(5)
    lt, l2 = 5, 8
    m = 1 << l2
    lo = [rng.getrandbits(64) for _ in range(m)]
    hi = [rng.getrandbits(57) | 1 for _ in range(m)]
    slo, shi, p2 = kernel_packed_write(lo, hi, lt, l2, 64, seed=1, p2=16)
    assert p2 == 16
    want_lo, _ = pack(lo, hi, 16)
    assert slo != want_lo
