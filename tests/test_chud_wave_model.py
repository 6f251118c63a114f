"""CPU model of the wave products of the pi binary splitting
(pg_pi.hip: wave_cols, wave_norm, wave_product; k_chud_merge_wave).

wave_cols: lane l of chunk k sums the 128-bit products a_i b_(c-i) of column
c = 64 k + l in three words; wave_norm turns the column sums into limbs 64 at
a time: limb c gathers w0 of column c, w1 of column c - 1 and w2 of column
c - 2 as lo + 2^64 e, adds e of the limb below (x = t + 2^64 g), and the
carries into the chunk's limbs come from the ballot masks G (generate) and P
(t = 2^64 - 1) by one integer addition: carry into bit i = bit i of
((G | P) + G + cin) ^ P, the chunk's carry-out g_63 | (p_63 & c_63).  The
model runs those steps on Python integers and is compared with the exact
product (and the sum of two products, T = lt rq + rt lp), for random
operands, all-ones operands (carries that ripple through every limb and
across chunks) and lengths around the 64-limb chunks; a planted slip in the
carry formula is caught.
"""
import random

import pytest

M64 = (1 << 64) - 1


def to_int(limbs):
    return sum(v << (64 * i) for i, v in enumerate(limbs))


def wave_cols(sums, a, b):
    na, nb = len(a), len(b)
    ncols = na + nb - 1
    for c0 in range(0, ncols, 64):
        ilo = c0 + 1 - nb if c0 + 1 > nb else 0
        ihi = min(na - 1, c0 + 63)
        for lane in range(64):
            c = c0 + lane
            s = 0
            for i in range(ilo, ihi + 1):
                j = c - i
                if i <= c and j < nb:
                    s += a[i] * b[j]
            assert s < 1 << 192
            if c < ncols:
                w = sums[c] + s
                sums[c] = w & ((1 << 192) - 1)   # add192 (the sums stay below 2^192)
                assert w < 1 << 192


def wave_norm(sums, ncols, n, formula=None):
    formula = formula or (lambda G, P, cin: ((((G | P) + G) & M64) + cin) & M64 ^ P)
    out = [0] * n
    e_in, cin = 0, 0
    for c0 in range(0, n, 64):
        los, es = [], []
        for lane in range(64):
            c = c0 + lane
            lo = e = 0
            if c < n:
                a = sums[c] & M64 if c < ncols else 0
                b = (sums[c - 1] >> 64) & M64 if c >= 1 and c - 1 < ncols else 0
                d = sums[c - 2] >> 128 if c >= 2 and c - 2 < ncols else 0
                s = a + b + d
                lo, e = s & M64, s >> 64
                assert e <= 2
            los.append(lo)
            es.append(e)
        ts, G, Pm = [], 0, 0
        for lane in range(64):
            eb = es[lane - 1] if lane else e_in
            x = los[lane] + eb
            t = x & M64
            ts.append(t)
            G |= (x >> 64) << lane
            Pm |= (t == M64) << lane
        assert G & Pm == 0
        Cv = formula(G, Pm, cin)
        for lane in range(64):
            c = c0 + lane
            if c < n:
                out[c] = (ts[lane] + ((Cv >> lane) & 1)) & M64
        cin = ((G >> 63) & 1) | ((Pm >> 63) & (Cv >> 63) & 1)
        e_in = es[63]
    assert cin == 0 and e_in == 0   # the value fits n limbs
    return out


def wave_product(n, x, y, u=None, v=None, formula=None):
    sums = [0] * n
    if x and y:
        wave_cols(sums, x, y)
    if u and v:
        wave_cols(sums, u, v)
    return wave_norm(sums, n, n, formula)


def rand_limbs(rng, k, kind):
    if kind == "ones":
        return [M64] * k
    if kind == "top":
        return [rng.getrandbits(64) for _ in range(k - 1)] + [M64]
    return [rng.getrandbits(64) | (1 if i == k - 1 else 0) for i in range(k)]


@pytest.mark.parametrize("na,nb", [(1, 1), (1, 70), (63, 1), (48, 48), (64, 64), (65, 63), (100, 37), (130, 129)])
@pytest.mark.parametrize("kind", ["rand", "ones", "top"])
def test_wave_product(na, nb, kind):
    rng = random.Random(na * 1000 + nb)
    a, b = rand_limbs(rng, na, kind), rand_limbs(rng, nb, kind)
    got = wave_product(na + nb, a, b)
    assert to_int(got) == to_int(a) * to_int(b)


@pytest.mark.parametrize("kind", ["rand", "ones"])
def test_wave_t_sum(kind):
    """T = lt rq + rt lp in max(lt + rq, rt + lp) + 1 limbs."""
    rng = random.Random(11)
    lt, rq, rt, lp = (rand_limbs(rng, k, kind) for k in (90, 70, 80, 85))
    n = max(90 + 70, 80 + 85) + 1
    got = wave_product(n, lt, rq, rt, lp)
    assert to_int(got) == to_int(lt) * to_int(rq) + to_int(rt) * to_int(lp)


def test_carry_formula_slip_caught():
    """Carries taken as the generate mask shifted up one bit (ignoring the
    propagate runs) fail on all-ones operands."""
    rng = random.Random(3)
    a, b = rand_limbs(rng, 70, "ones"), rand_limbs(rng, 66, "ones")
    slip = lambda G, P, cin: ((G << 1) | cin) & M64
    with pytest.raises(AssertionError):
        got = wave_product(136, a, b, formula=slip)
        assert to_int(got) == to_int(a) * to_int(b)
