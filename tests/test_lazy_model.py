"""Model of the register DFTs of pg_modarith.h (dft_dif, dft_dit_inv) in
exact 32-bit arithmetic, in both residue forms: canonical [0, p) and lazy
[0, 2p) (Harvey's butterflies, used for the prime below 2^30).

Every 32-bit operation is checked not to wrap where the kernel relies on
that, and the outputs are compared with the DFT evaluated from its
definition.  The inputs include the extremes of each range (0, p - 1, 2p - 1).
"""
import random

import pytest

M32 = (1 << 32) - 1
PRIMES = {469762049: 3, 1811939329: 13, 2013265921: 31}   # prime: generator
LAZY_LIMIT = 1 << 30


class Arith:
    """The kernel's modular primitives on 32-bit words, with overflow checks."""

    def __init__(self, p, lazy):
        assert not lazy or p < LAZY_LIMIT
        self.p, self.lazy = p, lazy
        self.pr = 2 * p if lazy else p          # reduction bound
        self.top = self.pr                      # residues lie in [0, top)

    def add(self, u, v):
        s = u + v
        assert s <= M32, "sum wraps"
        return min(s, (s - self.pr) & M32)

    def sub(self, u, v):
        d = (u - v) & M32
        return min(d, (d + self.pr) & M32)

    def shoup(self, x, w):
        assert 0 <= x <= M32 and 0 <= w < self.p
        wq = (w << 32) // self.p
        q = (x * wq) >> 32
        r = (x * w - q * self.p) & M32
        assert r < 2 * self.p, "Shoup bound"
        assert r % self.p == (x * w) % self.p
        return r if self.lazy else min(r, (r - self.p) & M32)

    def diff_in(self, u, v):
        # u - v + pr, taken by the product unreduced
        t = u - v + self.pr
        assert 0 < t <= M32, "difference leaves 32 bits"
        return t


def bitrev(q, r):
    out, b = 0, 1
    while b < r:
        out = (out << 1) | ((q // b) & 1)
        b <<= 1
    return out


def omega32(p, g, inverse=False):
    w = pow(g, (p - 1) // 32, p)
    return pow(w, p - 2, p) if inverse else w


def dft_dif(x, A, w32, raw=False):
    """raw: the last stage leaves outputs 1 .. r-1 unreduced (u + v and
    u - v + pr), for a pass whose outputs all but the first go straight into
    a twiddle product."""
    r = len(x)
    x = list(x)
    h = r // 2
    while h >= 1:
        for b in range(0, r, 2 * h):
            for j in range(h):
                u, v = x[b + j], x[b + j + h]
                if raw and h == 1:
                    if b == 0:
                        x[b] = A.add(u, v)
                    else:
                        x[b] = u + v
                        assert x[b] <= M32, "unreduced sum wraps"
                    x[b + 1] = A.diff_in(u, v)
                    continue
                x[b + j] = A.add(u, v)
                if j == 0:
                    x[b + j + h] = A.sub(u, v)
                else:
                    x[b + j + h] = A.shoup(A.diff_in(u, v), w32[j * (16 // h)])
        h //= 2
    return x


def dft_dit_inv(x, A, iw32):
    r = len(x)
    x = list(x)
    h = 1
    while h <= r // 2:
        for b in range(0, r, 2 * h):
            for j in range(h):
                u, v = x[b + j], x[b + j + h]
                if j:
                    v = A.shoup(v, iw32[j * (16 // h)])
                x[b + j] = A.add(u, v)
                x[b + j + h] = A.sub(u, v)
        h *= 2
    return x


def naive_dft(x, p, w):
    r = len(x)
    return [sum(x[n] * pow(w, n * k, p) for n in range(r)) % p for k in range(r)]


def cases(p, top, r, rng):
    yield [0] * r
    yield [top - 1] * r
    yield [p - 1] * r
    for _ in range(20):
        yield [rng.choice([v for v in (0, 1, p - 1, p, top - 1) if v < top] + [rng.randrange(top)])
               for _ in range(r)]


@pytest.mark.parametrize("p,lazy", [(469762049, True), (469762049, False),
                                    (1811939329, False), (2013265921, False)])
@pytest.mark.parametrize("r", [2, 4, 8, 16])
def test_register_dfts(p, lazy, r):
    g = PRIMES[p]
    w, iw = omega32(p, g), omega32(p, g, True)
    w32 = [pow(w, i, p) for i in range(16)]
    iw32 = [pow(iw, i, p) for i in range(16)]
    wr, iwr = pow(w, 32 // r, p), pow(iw, 32 // r, p)
    A = Arith(p, lazy)
    rng = random.Random(p + r)
    for x in cases(p, A.top, r, rng):
        y = dft_dif(x, A, w32)
        assert all(0 <= v < A.top for v in y)
        want = naive_dft([v % p for v in x], p, wr)
        assert [y[bitrev(q, r)] % p for q in range(r)] == want
        # inverse from the bit-reversed spectrum: r times the input
        z = dft_dit_inv(y, A, iw32)
        assert all(0 <= v < A.top for v in z)
        assert [v % p for v in z] == [(v * r) % p for v in x]
        # and from its definition
        yb = [y[bitrev(q, r)] % p for q in range(r)]
        assert [v % p for v in z] == naive_dft(yb, p, iwr)


@pytest.mark.parametrize("p,lazy", [(469762049, True), (469762049, False),
                                    (1811939329, False), (2013265921, False)])
@pytest.mark.parametrize("r", [2, 4, 8, 16])
def test_raw_outputs_into_twiddles(p, lazy, r):
    """A forward pass with inter-pass twiddles: the register DFT's last stage
    leaves outputs 1 .. r-1 unreduced and Shoup's product takes them (any
    x < 2^32); the twiddled outputs equal those of the reduced form."""
    g = PRIMES[p]
    w = omega32(p, g)
    w32 = [pow(w, i, p) for i in range(16)]
    A = Arith(p, lazy)
    rng = random.Random(3 * p + r)
    tw = [rng.randrange(p) for _ in range(r)]
    for x in cases(p, A.top, r, rng):
        y0 = dft_dif(x, A, w32)
        y1 = dft_dif(x, A, w32, raw=True)
        assert y1[0] == y0[0] and 0 <= y1[0] < A.top
        assert all(y1[q] < 2 * A.pr for q in range(1, r))
        assert [v % p for v in y1] == [v % p for v in y0]
        t0 = [y0[0]] + [A.shoup(y0[q], tw[q]) for q in range(1, r)]
        t1 = [y1[0]] + [A.shoup(y1[q], tw[q]) for q in range(1, r)]
        # lazy residues may differ by p as representatives; canonical ones agree exactly
        assert [v % p for v in t0] == [v % p for v in t1] and all(0 <= v < A.top for v in t1)
        assert lazy or t0 == t1


def test_lazy_needs_the_small_prime():
    # the two larger primes cannot take lazy residues: a sum of two values
    # below 2p leaves 32 bits
    for p in (1811939329, 2013265921):
        assert 4 * p > M32
        with pytest.raises(AssertionError):
            Arith(p, True)
    assert 4 * 469762049 <= M32


def test_pointwise_product_takes_lazy_operands():
    # mont_mul(a, b) needs a * b < p 2^32; lazy operands give a * b < 4 p^2
    p = 469762049
    assert (2 * p - 1) ** 2 < p << 32


def mont_tw(x, wm, p, lazy):
    """mul_mont_tw of pg_ntt_impl.h: one Montgomery product of x by a twiddle
    held as wm = w 2^32 mod p, in exact 32/64-bit steps."""
    ninv = (-pow(p, -1, 1 << 32)) % (1 << 32)
    t = x * wm
    assert t < 1 << 64
    m = (t & M32) * ninv & M32
    u = t + m * p
    assert u < 1 << 64, "64-bit sum wraps"
    assert u & M32 == 0
    r = u >> 32
    assert r < 2 * p, "Montgomery bound"
    return r if lazy else min(r, (r - p) & M32)


@pytest.mark.parametrize("p", sorted(PRIMES))
def test_inverse_column_lds_twiddles(p):
    # The inverse column kernel holds its pass twiddles in LDS as
    # wm = wq (2^32 - p) mod 2^32, from the Shoup pair (w, wq = floor(w 2^32 / p)):
    # that is w 2^32 mod p, and the Montgomery product by it is x w mod p,
    # within [0, 2p) unreduced for the lazy prime's inputs x < 2p and within
    # [0, p) after the correction for any x < 2^32.
    rng = random.Random(p)
    ws = [0, 1, p - 1, 2, (p - 1) // 2] + [rng.randrange(p) for _ in range(300)]
    xs_lazy = [0, 1, p - 1, p, 2 * p - 1] + [rng.randrange(2 * p) for _ in range(20)]
    xs_any = [0, 1, p - 1, p, M32] + [rng.randrange(1 << 32) for _ in range(20)]
    for w in ws:
        wq = (w << 32) // p
        wm = wq * ((1 << 32) - p) & M32
        assert wm == (w << 32) % p
        assert wm < p
        for x in xs_any:
            assert mont_tw(x, wm, p, False) == x * w % p
        if p < LAZY_LIMIT:
            for x in xs_lazy:
                r = mont_tw(x, wm, p, True)
                assert r < 2 * p and r % p == x * w % p


def test_inverse_column_lds_twiddles_catch_a_wrong_form():
    # the check above rejects a plausible slip: the Shoup quotient itself
    # (or w) taken as the Montgomery form
    p = 1811939329
    w = 123456789
    wq = (w << 32) // p
    assert mont_tw(5, wq % p, p, False) != 5 * w % p
    assert mont_tw(5, w, p, False) != 5 * w % p
