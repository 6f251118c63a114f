"""ctypes binding of the oracle (oracle/liboracle.so) for the tests.

Test infrastructure only: the oracle is the checker, never the thing tested.
"""
import ctypes
import hashlib
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
P64 = ctypes.POINTER(ctypes.c_uint64)
P32 = ctypes.POINTER(ctypes.c_uint32)
SZ = ctypes.c_size_t


def _p64(a):
    return a.ctypes.data_as(P64)


def _p32(a):
    return a.ctypes.data_as(P32)


def trim(a):
    a = np.asarray(a, dtype=np.uint64)
    n = len(a)
    while n and a[n - 1] == 0:
        n -= 1
    return a[:n]


def digest(limbs):
    t = trim(limbs)
    return hashlib.sha256(t.astype("<u8").tobytes()).hexdigest()


class Oracle:
    def __init__(self, path=None):
        self.lib = L = ctypes.CDLL(path or os.path.join(ROOT, "oracle", "liboracle.so"))
        for f in (L.or_mul_ntt, L.or_mul_basic):
            f.restype = SZ
            f.argtypes = [P64, SZ, P64, SZ, P64, SZ]
        for f in (L.or_mul_fft,):
            f.restype = SZ
            f.argtypes = [P64, SZ, P64, SZ, P64, SZ, ctypes.POINTER(ctypes.c_double)]
        L.or_mul_fft_bits.restype = SZ
        L.or_mul_fft_bits.argtypes = [P64, SZ, P64, SZ, P64, SZ, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_double)]
        L.or_mul_into.restype = SZ
        L.or_mul_into.argtypes = [P64, SZ, P64, SZ, P64, SZ, ctypes.c_int]
        L.or_sqr.restype = SZ
        L.or_sqr.argtypes = [P64, SZ, P64, SZ]
        L.or_ntt_prime.restype = ctypes.c_uint32
        L.or_ntt_prime.argtypes = [ctypes.c_int]
        L.or_ntt_ctx.restype = ctypes.c_int
        L.or_ntt_ctx.argtypes = [ctypes.c_int, P32]
        L.or_mont_mul.restype = ctypes.c_uint32
        L.or_mont_mul.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32]
        L.or_ntt_forward.argtypes = [ctypes.c_int, P32, SZ]
        L.or_ntt_inverse.argtypes = [ctypes.c_int, P32, SZ]
        L.or_ntt_convolve.argtypes = [ctypes.c_int, P32, SZ, P32, SZ, SZ, P32]
        L.or_ntt_crt2.restype = ctypes.c_uint64
        L.or_ntt_crt2.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.or_ntt_digits.restype = SZ
        L.or_ntt_digits.argtypes = [P64, SZ, P32]
        L.or_ntt_plan.restype = SZ
        L.or_ntt_plan.argtypes = [SZ, P64, SZ, P64, ctypes.POINTER(ctypes.c_int)]
        L.or_fill.argtypes = [P64, SZ, ctypes.c_uint64, ctypes.c_int]
        L.or_divmod.restype = SZ
        L.or_divmod.argtypes = [P64, SZ, ctypes.POINTER(SZ), P64, SZ, ctypes.POINTER(SZ), P64, SZ, P64, SZ]
        L.or_recip.restype = SZ
        L.or_recip.argtypes = [P64, SZ, P64, SZ, SZ]
        L.or_sqrt.restype = SZ
        L.or_sqrt.argtypes = [P64, SZ, P64, SZ]
        L.or_mod_p.restype = ctypes.c_uint64
        L.or_mod_p.argtypes = [P64, SZ, ctypes.c_int, ctypes.c_uint64]
        L.or_mod_p3.argtypes = [P64, SZ, P64]
        L.or_dec_mod_p3.argtypes = [ctypes.c_char_p, SZ, P64]

    def _out(self, an, bn):
        return np.zeros(max(an + bn, 1), np.uint64)

    def mul_ntt(self, a, b):
        r = self._out(len(a), len(b))
        n = self.lib.or_mul_ntt(_p64(r), len(r), _p64(a), len(a), _p64(b), len(b))
        return r[:n]

    def mul_fft(self, a, b):
        r = self._out(len(a), len(b))
        e = ctypes.c_double(0)
        n = self.lib.or_mul_fft(_p64(r), len(r), _p64(a), len(a), _p64(b), len(b), ctypes.byref(e))
        return r[:n], e.value

    def mul_fft_bits(self, a, b, bits):
        r = self._out(len(a), len(b))
        e = ctypes.c_double(0)
        n = self.lib.or_mul_fft_bits(_p64(r), len(r), _p64(a), len(a), _p64(b), len(b), bits,
                                     ctypes.byref(e))
        return r[:n], e.value

    def mul_basic(self, a, b):
        r = self._out(len(a), len(b))
        n = self.lib.or_mul_basic(_p64(r), len(r), _p64(a), len(a), _p64(b), len(b))
        return r[:n]

    def mul_into(self, a, b, sqr=False):
        r = self._out(len(a), len(b))
        n = self.lib.or_mul_into(_p64(r), len(r), _p64(a), len(a), _p64(b), len(b), int(sqr))
        return r[:n]

    def prime(self, k):
        return self.lib.or_ntt_prime(k)

    def ctx(self, k):
        out = np.zeros(5, np.uint32)
        rc = self.lib.or_ntt_ctx(k, _p32(out))
        return rc, [int(x) for x in out]

    def mont_mul(self, k, a, b):
        return self.lib.or_mont_mul(k, a, b)

    def ntt_forward(self, k, a):
        a = np.ascontiguousarray(a, dtype=np.uint32).copy()
        self.lib.or_ntt_forward(k, _p32(a), len(a))
        return a

    def ntt_inverse(self, k, a):
        a = np.ascontiguousarray(a, dtype=np.uint32).copy()
        self.lib.or_ntt_inverse(k, _p32(a), len(a))
        return a

    def convolve(self, k, a, b, n):
        a = np.ascontiguousarray(a, dtype=np.uint32)
        b = np.ascontiguousarray(b, dtype=np.uint32)
        out = np.zeros(n, np.uint32)
        self.lib.or_ntt_convolve(k, _p32(a), len(a), _p32(b), len(b), n, _p32(out))
        return out

    def crt2(self, r1, r2):
        return self.lib.or_ntt_crt2(r1, r2)

    def digits(self, a):
        out = np.zeros(4 * max(len(a), 1), np.uint32)
        n = self.lib.or_ntt_digits(_p64(a), len(a), _p32(out))
        return out[:n]

    def ntt_plan(self, a, b):
        p = ctypes.c_int(0)
        n = self.lib.or_ntt_plan(len(a), _p64(a), len(b), _p64(b), ctypes.byref(p))
        return n, p.value

    def divmod(self, a, b):
        """(floor(a / b), a mod b), trimmed."""
        a, b = trim(a), trim(b)
        q = np.zeros(max(len(a) - len(b) + 1, 1), np.uint64)
        r = np.zeros(max(len(b), len(a), 1), np.uint64)
        ql, rl = SZ(0), SZ(0)
        st = self.lib.or_divmod(_p64(q), len(q), ctypes.byref(ql), _p64(r), len(r), ctypes.byref(rl),
                                _p64(a), len(a), _p64(b), len(b))
        assert st == 0
        return q[:ql.value].copy(), r[:rl.value].copy()

    def recip(self, b, bits):
        b = trim(b)
        x = np.zeros(bits // 64 + 2, np.uint64)
        n = self.lib.or_recip(_p64(x), len(x), _p64(b), len(b), bits)
        assert n != 2**64 - 1
        return x[:n].copy()

    def sqrt(self, a):
        a = trim(a)
        x = np.zeros(len(a) // 2 + 3, np.uint64)
        n = self.lib.or_sqrt(_p64(x), len(x), _p64(a), len(a))
        assert n != 2**64 - 1
        return x[:n].copy()

    def mod_p(self, a, k, c):
        """a mod (2^k - c) of a little-endian limb array."""
        a = np.ascontiguousarray(a, dtype=np.uint64)
        return int(self.lib.or_mod_p(_p64(a), len(a), k, c))

    MODS3 = ((1 << 61) - 1, (1 << 63) - 25, (1 << 64) - 59)

    def mod_p3(self, a):
        """a mod (2^61 - 1, 2^63 - 25, 2^64 - 59) in one pass."""
        a = np.ascontiguousarray(a, dtype=np.uint64)
        out = np.zeros(3, np.uint64)
        self.lib.or_mod_p3(_p64(a), len(a), _p64(out))
        return tuple(int(v) for v in out)

    def dec_mod_p3(self, s):
        """A decimal digit string's value mod the MODS3 moduli."""
        b = s.encode() if isinstance(s, str) else bytes(s)
        out = np.zeros(3, np.uint64)
        self.lib.or_dec_mod_p3(b, len(b), _p64(out))
        return tuple(int(v) for v in out)

    def fill(self, n, seed, dist):
        a = np.zeros(n, np.uint64)
        self.lib.or_fill(_p64(a), n, seed, dist)
        return a
