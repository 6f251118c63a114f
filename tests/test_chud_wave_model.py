"""CPU model of the wave products of the pi binary splitting
(pg_pi.hip: col_sum, wave_product; k_chud_merge_wave).

col_sum: lane l of chunk k sums the 128-bit products a_i b_(c-i) of column
c = 64 k + l in three words (s0, s1, s2); wave_product then makes the
chunk's 64 limbs: limb c gathers s0 of column c, s1 of column c - 1 and s2 of
column c - 2 (lanes below, or the previous chunk's lanes 62 and 63) as
lo + 2^64 e, adds e of the limb below (x = t + 2^64 g), and the carries into
the chunk's limbs come from the ballot masks G (generate) and P
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


def col_sum(a, b, c0, lane):
    """col_sum: the column sum of column c0 + lane of a * b (three words)."""
    na, nb = len(a), len(b)
    c = c0 + lane
    ilo = c0 + 1 - nb if c0 + 1 > nb else 0
    ihi = min(na - 1, c0 + 63)
    s = 0
    for i in range(ilo, ihi + 1):
        j = c - i
        if i <= c and j < nb:
            s += a[i] * b[j]
    assert s < 1 << 192
    return s


def wave_product(n, x, y, u=None, v=None, formula=None):
    """wave_product: 64 limbs at a time from the chunk's column sums."""
    formula = formula or (lambda G, P, cin: ((((G | P) + G) & M64) + cin) & M64 ^ P)
    out = [0] * n
    s1_63 = s2_62 = s2_63 = 0
    e_in = cin = 0
    for c0 in range(0, n, 64):
        S = []
        for lane in range(64):
            s = 0
            if x and y:
                s += col_sum(x, y, c0, lane)
            if u and v:
                s += col_sum(u, v, c0, lane)
            assert s < 1 << 192
            S.append((s & M64, (s >> 64) & M64, s >> 128))
        los, es = [], []
        for lane in range(64):
            w1 = S[lane - 1][1] if lane else s1_63
            w2 = S[lane - 2][2] if lane >= 2 else (s2_63 if lane else s2_62)
            t = S[lane][0] + w1 + w2
            los.append(t & M64)
            es.append(t >> 64)
            assert es[-1] <= 2
        ts, G, Pm = [], 0, 0
        for lane in range(64):
            eb = es[lane - 1] if lane else e_in
            xx = los[lane] + eb
            t = xx & M64
            ts.append(t)
            G |= (xx >> 64) << lane
            Pm |= (t == M64) << lane
        assert G & Pm == 0
        Cv = formula(G, Pm, cin)
        for lane in range(64):
            c = c0 + lane
            if c < n:
                out[c] = (ts[lane] + ((Cv >> lane) & 1)) & M64
        cin = ((G >> 63) & 1) | ((Pm >> 63) & (Cv >> 63) & 1)
        e_in = es[63]
        s1_63, s2_62, s2_63 = S[63][1], S[62][2], S[63][2]
    assert cin == 0 and e_in == 0 and s1_63 == 0 and s2_62 == 0 and s2_63 == 0   # the value fits n limbs
    return out


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
