"""Model of the GPU carry resolution (pg_carry.hip: k_carry_spec, k_carry_fix)
checked against Python integers on adversarial limb sums.

Limb sums S_j = lo_j + 2^64 hi_j (hi_j < 2^60) are normalized chunk by chunk.
A chunk takes its carry-in from the 64 limbs just below it (the highest one
that does not propagate decides); where all 64 propagate it is written with
carry-in 0 and marked unresolved.  The fix-up walks back over the chunk
aggregates to the nearest chunk whose carry-out is known and adds the +1
ripple where the true carry-in is 1.  The chunk here is 256 limbs (the
kernel's is 2048), so short inputs cross many chunk boundaries.
"""
import random

M64 = (1 << 64) - 1
CHUNK = 256
WINDOW = 64


def normalize(slo, shi, out_len, flag=False):
    """Limbs [0, out_len) of the normalized value; with flag, also the flag
    word (limb m of the value, m = len(slo)), as (limbs, flag word).  The
    fix-up skips a chunk past out_len that does not hold limb m, as
    k_carry_fix does."""
    m = len(slo)
    nch = (max(m + (1 if flag else 0), out_len) + CHUNK - 1) // CHUNK
    fw = [None]

    def tgp(j):
        hprev = shi[j - 1] if 1 <= j <= m else 0
        if j < m:
            t = (slo[j] + hprev) & M64
            return t, int(t < slo[j]), int(t == M64)
        return (hprev if j == m else 0), 0, 0

    out = [0] * out_len
    agg = []
    for c in range(nch):
        base = c * CHUNK
        if c == 0:
            cin, resolved = 0, 1
        else:
            win = [tgp(base - WINDOW + k) for k in range(WINDOW)]
            stops = [k for k in range(WINDOW) if not win[k][2]]
            cin, resolved = (win[max(stops)][1], 1) if stops else (0, 0)
        x, g_all, p_all = cin, 0, 1
        for j in range(base, base + CHUNK):
            t, g, p = tgp(j)
            if j < out_len:
                out[j] = (t + x) & M64
            if j == m and flag:
                fw[0] = (t + x) & M64
            x = g | (p & x)
            g_all, p_all = g | (p & g_all), p & p_all
        agg.append((g_all, p_all, cin, resolved))

    def carry_in(c):
        i = c - 1
        while True:
            g, p, used, res = agg[i]
            if g:
                return 1
            if not p:
                return 0
            if res:
                return used
            i -= 1

    for c in range(nch):
        base = c * CHUNK
        if agg[c][3]:
            continue
        if base >= out_len and not (flag and base <= m < base + CHUNK):
            continue   # stores nothing (k_carry_fix returns at once)
        if not carry_in(c):
            continue
        if agg[c][1]:   # every limb propagates: each becomes 0, stored without reading the limbs
            assert not (flag and base <= m < base + CHUNK)
            for j in range(base, min(base + CHUNK, out_len)):
                out[j] = 0
            continue
        for j in range(base, base + CHUNK):
            t, _, p = tgp(j)
            if j < out_len:
                out[j] = (t + 1) & M64
            if j == m and flag:
                fw[0] = (t + 1) & M64
            if not p:
                break
    return (out, fw[0]) if flag else out


def check(slo, shi, out_len):
    v = sum((slo[j] + (shi[j] << 64)) << (64 * j) for j in range(len(slo)))
    assert normalize(slo, shi, out_len) == [(v >> (64 * j)) & M64 for j in range(out_len)]


def test_uniform_sums():
    rng = random.Random(1)
    for m in (1, 255, 256, 257, 1000):
        slo = [rng.getrandbits(64) for _ in range(m)]
        shi = [rng.getrandbits(50) for _ in range(m)]
        check(slo, shi, m + 1)


def test_long_propagate_runs():
    # limbs that propagate (t = 2^64 - 1) for whole chunks, ended by a
    # generate or a kill at chosen offsets around chunk and window edges
    for end in (0, 1, 63, 64, 65, 191, 192, 255, 256, 257, 700):
        m = 1100
        slo, shi = [M64] * m, [0] * m
        slo[end] = 0
        shi[end] = 1      # limb end + 1 generates
        check(slo, shi, m + 1)
        slo, shi = [M64] * m, [0] * m
        slo[end] = 5      # limb end kills
        check(slo, shi, m)


def test_all_ones_square_pattern():
    # the limb pattern of (2^(64n) - 1)^2 carried from sums of one word each
    n = 600
    v = ((1 << (64 * n)) - 1) ** 2
    slo = [(v >> (64 * j)) & M64 for j in range(2 * n)]
    check(slo, [0] * (2 * n), 2 * n)


def test_random_mixtures():
    rng = random.Random(7)
    for _ in range(20):
        m = rng.choice([300, 513, 900])
        slo = [M64 if rng.random() < 0.97 else rng.getrandbits(64) for _ in range(m)]
        shi = [1 if rng.random() < 0.01 else 0 for _ in range(m)]
        check(slo, shi, m + rng.choice([0, 1]))


def test_difference_with_flag_stored_short():
    """y - p as y + ~p + 1 over m limbs, stored in out_len << m limbs, with
    the flag word (limb m: 1 when y >= p): the top limbs cancel, so every
    chunk above the stored part propagates and the fix-up skips the chunks
    that store nothing; the low limbs and the flag stay exact."""
    rng = random.Random(13)
    for m, out_len in ((1500, 300), (2049, 257), (1024, 1), (700, 699)):
        for delta in (0, 1, 5, -1, -7):
            p = rng.getrandbits(64 * (m - 2)) | 1 << (64 * (m - 2))
            y = p + delta * (1 + rng.getrandbits(64 * (out_len // 2 + 1)) if delta else 0)
            if y < 0:
                continue
            yl = [(y >> (64 * j)) & M64 for j in range(m)]
            pl = [(~(p >> (64 * j))) & M64 for j in range(m)]
            slo = [(a + b) & M64 for a, b in zip(yl, pl)]
            shi = [(a + b) >> 64 for a, b in zip(yl, pl)]
            # + 1 (two's complement) folded into limb 0's sum
            s0 = slo[0] + 1
            slo[0], shi[0] = s0 & M64, shi[0] + (s0 >> 64)
            out, fw = normalize(slo, shi, out_len, flag=True)
            exact = y + ((1 << (64 * m)) - 1 - p) + 1   # y + ~p + 1 over m limbs, carry-out at limb m
            assert out == [(exact >> (64 * j)) & M64 for j in range(out_len)], (m, out_len, delta)
            assert fw == exact >> (64 * m), (m, out_len, delta)
