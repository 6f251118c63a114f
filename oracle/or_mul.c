/*
 * or_mul.c -- oracle restatement of preclib's dispatcher, schoolbook and
 * blocked multiply (src/mul_basic.cpp).  TEST INFRASTRUCTURE ONLY.
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>

#define KARATSUBA_THRESHOLD 24   /* mul_basic.cpp:3 */
#define FFT_THRESHOLD 192        /* mul_basic.cpp:4 */
#define NTT_THRESHOLD 6144       /* mul_basic.cpp:5 */

typedef unsigned __int128 u128;

static size_t trim(const uint64_t *r, size_t n) { while (n && r[n - 1] == 0) --n; return n; }

/* Schoolbook into a zeroed buffer of an+bn limbs (mul_basic.cpp:131-185). */
static void school(uint64_t *t, const uint64_t *a, size_t an, const uint64_t *b, size_t bn) {
    memset(t, 0, (an + bn) * sizeof(uint64_t));
    for (size_t i = 0; i < an; ++i) {
        uint64_t carry = 0;
        for (size_t j = 0; j < bn; ++j) {
            u128 p = (u128)a[i] * b[j] + t[i + j] + carry;
            t[i + j] = (uint64_t)p;
            carry = (uint64_t)(p >> 64);
        }
        t[i + bn] = carry;
    }
}

size_t or_mul_basic(uint64_t *r, size_t rcap, const uint64_t *a, size_t an,
                    const uint64_t *b, size_t bn) {
    an = trim(a, an); bn = trim(b, bn);
    if (!an || !bn) return 0;
    uint64_t *t = (uint64_t *)malloc((an + bn) * sizeof(uint64_t));
    school(t, a, an, b, bn);
    size_t len = trim(t, an + bn);
    if (len > rcap) { free(t); return (size_t)-1; }
    memcpy(r, t, len * sizeof(uint64_t));
    free(t);
    return len;
}

/* precn_sqr (mul_basic.cpp:193-205).  The <=64 square basecase and the
 * Karatsuba range produce the schoolbook integer, so schoolbook stands in for
 * both; sizes above 192 go to mul_fft(a, a) at any size, as the reference
 * does. */
size_t or_sqr(uint64_t *r, size_t rcap, const uint64_t *a, size_t an) {
    an = trim(a, an);
    if (!an) return 0;
    if (an > FFT_THRESHOLD) return or_mul_fft(r, rcap, a, an, a, an, NULL);
    return or_mul_basic(r, rcap, a, an, a, an);
}

static size_t mul_into_rec(uint64_t *r, size_t rcap, const uint64_t *a, size_t an,
                           const uint64_t *b, size_t bn);

/* mul_unbalanced_into (mul_basic.cpp:207-236) with mul_add_shift (51-74):
 * blocks of 2*lo limbs of the longer operand, each multiplied through the
 * dispatcher and added at its limb offset. */
static size_t unbalanced(uint64_t *r, size_t rcap, const uint64_t *a, size_t an,
                         const uint64_t *b, size_t bn) {
    const uint64_t *x = a, *y = b;
    size_t xn = an, yn = bn;
    if (xn < yn) { x = b; xn = bn; y = a; yn = an; }
    size_t total = xn + yn + 1;
    uint64_t *acc = (uint64_t *)calloc(total, sizeof(uint64_t));
    uint64_t *prod = (uint64_t *)malloc((2 * yn + yn + 2) * sizeof(uint64_t));
    size_t block = yn * 2;
    for (size_t off = 0; off < xn; off += block) {
        size_t n = block < xn - off ? block : xn - off;
        size_t pn = trim(x + off, n);
        if (!pn) continue;
        size_t plen = mul_into_rec(prod, pn + yn, x + off, pn, y, yn);
        uint64_t carry = 0;
        size_t i = 0;
        for (; i < plen; ++i) {
            u128 s = (u128)acc[off + i] + prod[i] + carry;
            acc[off + i] = (uint64_t)s;
            carry = (uint64_t)(s >> 64);
        }
        for (; carry; ++i) {
            u128 s = (u128)acc[off + i] + carry;
            acc[off + i] = (uint64_t)s;
            carry = (uint64_t)(s >> 64);
        }
    }
    free(prod);
    size_t len = trim(acc, total);
    if (len > rcap) { free(acc); return (size_t)-1; }
    memcpy(r, acc, len * sizeof(uint64_t));
    free(acc);
    return len;
}

/* mul_into (mul_basic.cpp:238-276) without the alias branch.  The Karatsuba
 * branch (271-273) yields the schoolbook integer and is answered by it. */
static size_t mul_into_rec(uint64_t *r, size_t rcap, const uint64_t *a, size_t an,
                           const uint64_t *b, size_t bn) {
    an = trim(a, an); bn = trim(b, bn);
    if (!an || !bn) return 0;
    size_t lo = an < bn ? an : bn, hi = an > bn ? an : bn;
    if (lo <= KARATSUBA_THRESHOLD) return or_mul_basic(r, rcap, a, an, b, bn);
    if (hi > lo * 2) return unbalanced(r, rcap, a, an, b, bn);
    if (hi > NTT_THRESHOLD) return or_mul_ntt(r, rcap, a, an, b, bn);
    if (hi > FFT_THRESHOLD) return or_mul_fft(r, rcap, a, an, b, bn, NULL);
    return or_mul_basic(r, rcap, a, an, b, bn);
}

size_t or_mul_into(uint64_t *r, size_t rcap, const uint64_t *a, size_t an,
                   const uint64_t *b, size_t bn, int sqr) {
    if (sqr) return or_sqr(r, rcap, a, an);     /* 244-247 */
    return mul_into_rec(r, rcap, a, an, b, bn);
}

/* splitmix64 (Vigna): the (k+1)-th output of the stream whose state starts at
 * seed. */
static inline uint64_t sm64(uint64_t seed, uint64_t k) {
    uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void or_fill(uint64_t *a, size_t n, uint64_t seed, int dist) {
    if (!n) return;
    for (size_t i = 0; i < n; ++i) {
        if (dist == 1) a[i] = ~(uint64_t)0;
        else if (dist == 2) {
            uint64_t u = sm64(seed, 2 * i), v = sm64(seed, 2 * i + 1);
            a[i] = (u >> 32) < 42949673u ? (v ? v : 1) : 0;
        } else a[i] = sm64(seed, i);
    }
    if (a[n - 1] == 0) {
        uint64_t v = sm64(seed, 2 * (n - 1) + 1);
        a[n - 1] = v ? v : 1;
    }
}

/* t mod (2^k - c) for t < 2^80: fold the bits above k back in as hi * c. */
static uint64_t fold(u128 t, int k, uint64_t c) {
    const u128 p = ((u128)1 << k) - c, mask = ((u128)1 << k) - 1;
    while (t >> k) t = (t >> k) * c + (t & mask);
    while (t >= p) t -= p;
    return (uint64_t)t;
}

uint64_t or_mod_p(const uint64_t *a, size_t n, int k, uint64_t c) {
    const uint64_t m = c << (64 - k);   /* 2^64 mod p */
    uint64_t r = 0;
    for (size_t i = n; i-- > 0;) r = fold((u128)r * m + a[i], k, c);
    return r;
}

void or_mod_p3(const uint64_t *a, size_t n, uint64_t out[3]) {
    uint64_t r0 = 0, r1 = 0, r2 = 0;
    for (size_t i = n; i-- > 0;) {
        r0 = fold((u128)r0 * 8 + a[i], 61, 1);     /* 2^64 = 2^3 mod 2^61 - 1 */
        r1 = fold((u128)r1 * 50 + a[i], 63, 25);   /* 2^64 = 2 * 25 mod 2^63 - 25 */
        r2 = fold((u128)r2 * 59 + a[i], 64, 59);   /* 2^64 = 59 mod 2^64 - 59 */
    }
    out[0] = r0; out[1] = r1; out[2] = r2;
}

void or_dec_mod_p3(const char *s, size_t n, uint64_t out[3]) {
    uint64_t r0 = 0, r1 = 0, r2 = 0;
    for (size_t i = 0; i < n; ++i) {
        const uint64_t d = (uint64_t)(s[i] - '0');
        r0 = fold((u128)r0 * 10 + d, 61, 1);
        r1 = fold((u128)r1 * 10 + d, 63, 25);
        r2 = fold((u128)r2 * 10 + d, 64, 59);
    }
    out[0] = r0; out[1] = r1; out[2] = r2;
}
