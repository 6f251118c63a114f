"""Packed limb sums (NttLeaf::pack_log; pg_ntt_impl.h col_inv_wide writes
them, pg_carry.hip limbs_t_packed reads them) against the plain form.

The inverse column kernel holds limb sums S_j = lo_j + 2^64 hi_j in pieces of
PL = T / 2 consecutive limbs (one row of a column tile).  Packed, slo[j] is
t_j = lo_j + hi_(j-1) mod 2^64 for every limb but the first of its piece
(where it is lo_j), and shi[j / PL] holds the high word of the piece's last
limb with bit 56 + q set when limb q's addition carried.  The carry's loader
must recover the same (t_j, generate, propagate) as from the plain sums, and
the top word t_m = hi_(m-1); here both sides are modelled on Python integers,
including the largest high words (2^57 - 1) and additions that carry at
every position of a piece.
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
