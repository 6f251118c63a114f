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


# ---- the decimal conversion's split, one wave per division (k_dec_split_wave) ----

def wave_add(a, b, neg, cin, m):
    """wave_add: (limbs of a + (neg ? ~b : b) + cin over m limbs, carry out of
    limb m - 1), 64 limbs at a time with the ballot carries."""
    out = []
    c = cin
    for c0 in range(0, m, 64):
        s_l, G, P = [], 0, 0
        for lane in range(64):
            k = c0 + lane
            inn = k < m
            x = a[k] if inn and k < len(a) else 0
            y = b[k] if inn and k < len(b) else 0
            if neg and inn:
                y = ~y & M64
            s = x + y
            s_l.append(s & M64)
            G |= (inn and s >> 64) << lane
            P |= (inn and (s & M64) == M64) << lane
        Cv = ((((G | P) + G) & M64) + c) & M64 ^ P
        for lane in range(64):
            if c0 + lane < m:
                out.append((s_l[lane] + ((Cv >> lane) & 1)) & M64)
        top = min(m - c0, 64)
        c = (((G >> 63) & 1) | ((P >> 63) & (Cv >> 63) & 1)) if top == 64 else (Cv >> top) & 1
    return out, c


def limbs(v, n):
    return [(v >> (64 * k)) & M64 for k in range(n)]


def dec_split(y, ybits, W, V, wb):
    """k_dec_split_wave's steps on limbs: returns (r2, q2, flags)."""
    nw = (W.bit_length() + 63) // 64
    vb = (V.bit_length() + 63) // 64
    n = max(1, (ybits + 63) // 64)
    qlen = (ybits + 2 - wb + 63) // 64 + 1
    sb = 2 * wb
    yl = limbs(y, n)
    ax = to_int(wave_product(n + vb, yl, limbs(V, vb)))
    q = (ax >> sb) & ((1 << (64 * qlen)) - 1)
    p = wave_product(qlen + nw, limbs(q, qlen), limbs(W, nw))
    m = max(n, qlen + nw) + 1
    r, f0 = wave_add(yl, p, True, 1, m)
    r = r[:nw + 1]
    t, f = wave_add(r, limbs(W, nw), True, 1, nw + 1)
    r2 = t if f else r
    q2, _ = wave_add(limbs(q, qlen), [], False, f, qlen)
    _, ge = wave_add(r2, limbs(W, nw), True, 1, nw + 1)
    return to_int(r2[:nw]), to_int(q2), (f0, f, ge)


@pytest.mark.parametrize("i", [1, 2, 3, 5])
def test_dec_split_vs_divmod(i):
    """Every division of a level: q2 = y div W, r2 = y mod W, flags (1, f, 0),
    with V = floor(2^(2 bits W) / W), y below W^2 (the level's bound) and
    bit lengths from the level's bound (as the host plans them)."""
    W = 10 ** (19 * 2 ** i)
    wb = W.bit_length()
    V = (1 << (2 * wb)) // W
    rng = random.Random(i)
    ys = [0, 1, W - 1, W, W + 1, W * W - 1, W * (W - 1), (1 << (2 * wb - 1)) % (W * W)]
    ys += [rng.randrange(W * W) for _ in range(12)]
    for y in ys:
        ybits = 2 * wb   # the bound the level plans with (y < W^2 <= 2^(2 wb))
        r2, q2, fl = dec_split(y, ybits, W, V, wb)
        assert (q2, r2) == divmod(y, W), y
        assert fl[0] == 1 and fl[2] == 0, (y, fl)


def test_dec_split_catches_a_short_reciprocal():
    """A reciprocal one unit short for a large quotient leaves q two below:
    the check flag (r2 >= W) reports it."""
    i = 3
    W = 10 ** (19 * 2 ** i)
    wb = W.bit_length()
    V = (1 << (2 * wb)) // W - 2 ** 40
    y = W * W - 1
    r2, q2, fl = dec_split(y, 2 * wb, W, V, wb)
    assert fl[2] == 1 and (q2, r2) != divmod(y, W)
