/*
 * or_ntt.c -- oracle restatement of preclib's mul_ntt (src/mul_ntt.cpp).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Scalar C; the reference's AVX2
 * lanes compute the same modular values, so they are not restated.
 */
#include "oracle.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* mul_ntt.cpp:27-35 */
#define NTT_MAX_LIMBS ((size_t)1 << 24)
#define NTT_MAX_TRANSFORM ((size_t)1 << 26)
static const uint32_t PRIMES[3] = {469762049u, 1811939329u, 2013265921u};
static const uint32_t ROOTS[3] = {3u, 13u, 31u};
/* The frozen table of mul_ntt.cpp:37-41 ({mod, root, ninv, r2, one}). */
static const uint32_t FROZEN[3][5] = {
    {469762049u, 3u, 469762047u, 460175152u, 67108855u},
    {1811939329u, 13u, 1811939327u, 959408210u, 671088638u},
    {2013265921u, 31u, 2013265919u, 1172168163u, 268435454u},
};

typedef struct { uint32_t mod, root, ninv, r2, one; } mctx;

/* -p^-1 mod 2^32 by Newton lifting (tools/ntt_montgomery_constants.py:9-13
 * computes the same quantity). */
static uint32_t neg_inv32(uint32_t p) {
    uint32_t x = 1;
    for (int i = 0; i < 5; ++i) x *= 2u - p * x;
    return (uint32_t)(0u - x);
}

static mctx make_ctx(int idx) {
    mctx c;
    c.mod = PRIMES[idx];
    c.root = ROOTS[idx];
    c.ninv = neg_inv32(c.mod);
    unsigned __int128 R2 = ((unsigned __int128)1 << 64) % c.mod;
    c.r2 = (uint32_t)R2;
    c.one = (uint32_t)(((uint64_t)1 << 32) % c.mod);
    return c;
}

int or_ntt_ctx(int idx, uint32_t out[5]) {
    mctx c = make_ctx(idx);
    out[0] = c.mod; out[1] = c.root; out[2] = c.ninv; out[3] = c.r2; out[4] = c.one;
    for (int k = 0; k < 5; ++k)
        if (out[k] != FROZEN[idx][k]) return -1;
    return 0;
}

uint32_t or_ntt_prime(int idx) { return PRIMES[idx]; }

/* mul_ntt.cpp:43-48: (x + ((x*ninv) mod 2^32) * p) / 2^32, one conditional
 * subtraction; valid for x < p * 2^32. */
static inline uint32_t mont_reduce(const mctx *c, uint64_t x) {
    uint32_t m = (uint32_t)x * c->ninv;
    uint64_t t = (x + (uint64_t)m * c->mod) >> 32;
    if (t >= c->mod) t -= c->mod;
    return (uint32_t)t;
}
static inline uint32_t mont_in(const mctx *c, uint32_t x) {   /* 50-55 */
    return mont_reduce(c, (uint64_t)x * c->r2);
}
static inline uint32_t mont_mul(const mctx *c, uint32_t a, uint32_t b) {  /* 57-59 */
    return mont_reduce(c, (uint64_t)a * b);
}
uint32_t or_mont_mul(int idx, uint32_t a, uint32_t b) {
    mctx c = make_ctx(idx);
    return mont_mul(&c, a, b);
}

/* mul_ntt.cpp:139-148: result in Montgomery form. */
static uint32_t mont_pow(const mctx *c, uint32_t a, uint64_t e) {
    uint32_t r = c->one, x = mont_in(c, a);
    while (e) {
        if (e & 1) r = mont_mul(c, r, x);
        e >>= 1;
        if (e) x = mont_mul(c, x, x);
    }
    return r;
}

typedef struct {
    mctx c;
    uint32_t inv_n;
    size_t n;
    uint32_t *rf, *ri;   /* per-level tables, entry [half + j] (150-172) */
} plan_t;

static void plan_make(plan_t *p, int idx, size_t n) {
    p->c = make_ctx(idx);
    p->n = n;
    p->inv_n = mont_pow(&p->c, (uint32_t)(n % p->c.mod), p->c.mod - 2);   /* 177 */
    p->rf = (uint32_t *)malloc(n * sizeof(uint32_t) + 4);
    p->ri = (uint32_t *)malloc(n * sizeof(uint32_t) + 4);
    p->rf[0] = p->ri[0] = p->c.one;
    for (size_t len = 2; len <= n; len <<= 1) {
        size_t half = len >> 1;
        uint32_t wlen = mont_pow(&p->c, p->c.root, (p->c.mod - 1) / len);
        uint32_t w = p->c.one;
        for (size_t i = 0; i < half; ++i) {
            p->rf[half + i] = w;
            w = mont_mul(&p->c, w, wlen);
        }
        p->ri[half] = p->c.one;
        for (size_t i = 1; i < half; ++i) p->ri[half + i] = p->c.mod - p->rf[half + half - i];
    }
}
static void plan_free(plan_t *p) { free(p->rf); free(p->ri); }

/* Gentleman-Sande DIF, natural in, bit-reversed out (mul_ntt.cpp:193-255). */
static void fwd(const plan_t *p, uint32_t *a, size_t n) {
    const uint32_t mod = p->c.mod;
    for (size_t len = n; len >= 2; len >>= 1) {
        size_t half = len >> 1;
        for (size_t i = 0; i < n; i += len)
            for (size_t j = 0; j < half; ++j) {
                uint32_t u = a[i + j], v = a[i + j + half];
                uint32_t s = u + v;
                if (s >= mod) s -= mod;
                uint32_t d = u >= v ? u - v : u + mod - v;
                a[i + j] = s;
                a[i + j + half] = mont_mul(&p->c, d, p->rf[half + j]);
            }
    }
}

/* Cooley-Tukey DIT, bit-reversed in, natural out, then 1/n (257-329). */
static void inv(const plan_t *p, uint32_t *a, size_t n) {
    const uint32_t mod = p->c.mod;
    for (size_t len = 2; len <= n; len <<= 1) {
        size_t half = len >> 1;
        for (size_t i = 0; i < n; i += len)
            for (size_t j = 0; j < half; ++j) {
                uint32_t u = a[i + j];
                uint32_t v = mont_mul(&p->c, a[i + j + half], p->ri[half + j]);
                uint32_t s = u + v;
                if (s >= mod) s -= mod;
                uint32_t d = u >= v ? u - v : u + mod - v;
                a[i + j] = s;
                a[i + j + half] = d;
            }
    }
    for (size_t i = 0; i < n; ++i) a[i] = mont_mul(&p->c, a[i], p->inv_n);
}

void or_ntt_forward(int idx, uint32_t *a, size_t n) {
    plan_t p; plan_make(&p, idx, n); fwd(&p, a, n); plan_free(&p);
}
void or_ntt_inverse(int idx, uint32_t *a, size_t n) {
    plan_t p; plan_make(&p, idx, n); inv(&p, a, n); plan_free(&p);
}

/* mul_ntt.cpp:349-387 (plan passed in so the caller can reuse it). */
static void convolve(const plan_t *p, const uint32_t *a, size_t na, const uint32_t *b,
                     size_t nb, size_t n, uint32_t *out, uint32_t *scratch) {
    memset(out, 0, n * sizeof(uint32_t));
    memset(scratch, 0, n * sizeof(uint32_t));
    for (size_t i = 0; i < na; ++i) out[i] = mont_in(&p->c, a[i]);
    for (size_t i = 0; i < nb; ++i) scratch[i] = mont_in(&p->c, b[i]);
    fwd(p, out, n);
    fwd(p, scratch, n);
    for (size_t i = 0; i < n; ++i) out[i] = mont_mul(&p->c, out[i], scratch[i]);
    inv(p, out, n);
    for (size_t i = 0; i < n; ++i) out[i] = mont_reduce(&p->c, out[i]);
}

void or_ntt_convolve(int idx, const uint32_t *a, size_t na, const uint32_t *b,
                     size_t nb, size_t n, uint32_t *out) {
    plan_t p; plan_make(&p, idx, n);
    uint32_t *scratch = (uint32_t *)malloc(n * sizeof(uint32_t));
    convolve(&p, a, na, b, nb, n, out, scratch);
    free(scratch);
    plan_free(&p);
}

/* mul_ntt.cpp:389-403, extended Euclid. */
static uint32_t mod_inv_u32(uint64_t a, uint32_t mod) {
    long long t = 0, nt = 1, r = mod, nr = (long long)(a % mod);
    while (nr) {
        long long q = r / nr, ot = t, orr = r;
        t = nt; nt = ot - q * nt;
        r = nr; nr = orr - q * nr;
    }
    if (t < 0) t += mod;
    return (uint32_t)t;
}

/* Garner, two primes (mul_ntt.cpp:405-412). */
uint64_t or_ntt_crt2(uint32_t r1, uint32_t r2) {
    uint32_t inv12 = mod_inv_u32(PRIMES[0], PRIMES[1]);
    uint64_t x1 = r1;
    uint64_t d2 = r2 >= x1 ? r2 - x1 : r2 + (uint64_t)PRIMES[1] - x1;
    uint64_t t2 = d2 * inv12 % PRIMES[1];
    return x1 + (uint64_t)PRIMES[0] * t2;
}

/* Garner, three primes (mul_ntt.cpp:414-424); the result can exceed 64 bits,
 * so it is returned as 128 bits.  The reference's uint64 version is only
 * reached when the two-prime bound fails, which its own size caps prevent;
 * here the wider type keeps the restatement exact in that branch too. */
static unsigned __int128 crt3(uint32_t r1, uint32_t r2, uint32_t r3) {
    const uint64_t m12 = (uint64_t)PRIMES[0] * PRIMES[1];
    uint32_t inv123 = mod_inv_u32(m12 % PRIMES[2], PRIMES[2]);
    uint64_t x12 = or_ntt_crt2(r1, r2);
    uint64_t x12_m3 = x12 % PRIMES[2];
    uint64_t d3 = r3 >= x12_m3 ? r3 - x12_m3 : r3 + (uint64_t)PRIMES[2] - x12_m3;
    uint64_t t3 = d3 * inv123 % PRIMES[2];
    return (unsigned __int128)x12 + (unsigned __int128)m12 * t3;
}

size_t or_ntt_digits(const uint64_t *a, size_t an, uint32_t *out) {   /* 331-342 */
    size_t k = 0;
    for (size_t i = 0; i < an; ++i) {
        out[k++] = (uint32_t)(a[i] & 0xFFFFu);
        out[k++] = (uint32_t)((a[i] >> 16) & 0xFFFFu);
        out[k++] = (uint32_t)((a[i] >> 32) & 0xFFFFu);
        out[k++] = (uint32_t)(a[i] >> 48);
    }
    while (k && out[k - 1] == 0) --k;
    return k;
}

/* mul_ntt.cpp:489-492 */
static int two_mod_ok(size_t terms) {
    uint64_t m12 = (uint64_t)PRIMES[0] * PRIMES[1];
    return terms <= (m12 - 1) / (0xFFFFull * 0xFFFFull);
}

static size_t trim(const uint64_t *r, size_t n) {
    while (n && r[n - 1] == 0) --n;
    return n;
}

size_t or_ntt_plan(size_t an, const uint64_t *a, size_t bn, const uint64_t *b, int *primes) {
    if (!an || !bn) return 0;
    size_t hi = an > bn ? an : bn;
    if (hi <= 192 || hi > NTT_MAX_LIMBS) return 0;
    size_t da = 4 * (an - 1), db = 4 * (bn - 1);
    for (uint64_t t = a[an - 1]; t; t >>= 16) ++da;
    for (uint64_t t = b[bn - 1]; t; t >>= 16) ++db;
    size_t n = 1;
    while (n < da + db) n <<= 1;
    if (n > NTT_MAX_TRANSFORM) return 0;
    if (primes) *primes = two_mod_ok(da < db ? da : db) ? 2 : 3;
    return n;
}

size_t or_mul_ntt(uint64_t *r, size_t rcap, const uint64_t *a, size_t an,
                  const uint64_t *b, size_t bn) {
    an = trim(a, an); bn = trim(b, bn);
    if (!an || !bn) return 0;
    size_t hi = an > bn ? an : bn;
    if (hi <= 192) return or_mul_basic(r, rcap, a, an, b, bn);          /* 496 */
    if (hi > NTT_MAX_LIMBS) return or_mul_fft(r, rcap, a, an, b, bn, NULL);  /* 498 */
    uint32_t *da = (uint32_t *)malloc(4 * an * sizeof(uint32_t));
    uint32_t *db = (uint32_t *)malloc(4 * bn * sizeof(uint32_t));
    size_t na = or_ntt_digits(a, an, da), nb = or_ntt_digits(b, bn, db);
    size_t n = 1;
    while (n < na + nb) n <<= 1;
    if (n > NTT_MAX_TRANSFORM) {                                          /* 506 */
        free(da); free(db);
        return or_mul_fft(r, rcap, a, an, b, bn, NULL);
    }
    int three = !two_mod_ok(na < nb ? na : nb);
    uint32_t *res[3] = {0, 0, 0};
    uint32_t *scratch = (uint32_t *)malloc(n * sizeof(uint32_t));
    for (int k = 0; k < 2 + three; ++k) {
        plan_t p; plan_make(&p, k, n);
        res[k] = (uint32_t *)malloc(n * sizeof(uint32_t));
        convolve(&p, da, na, db, nb, n, res[k], scratch);
        plan_free(&p);
    }
    free(scratch); free(da); free(db);

    /* ntt_from_residues2/3 (438-487): serial base-2^16 carry, digits OR-ed
     * into limbs. */
    size_t cap = n / 4 + 8;
    uint64_t *t = (uint64_t *)calloc(cap, sizeof(uint64_t));
    unsigned __int128 carry = 0;
    size_t id = 0;
    for (size_t i = 0; i < n; ++i, ++id) {
        unsigned __int128 cur = three ? crt3(res[0][i], res[1][i], res[2][i])
                                      : (unsigned __int128)or_ntt_crt2(res[0][i], res[1][i]);
        cur += carry;
        t[id >> 2] |= (uint64_t)(cur & 0xFFFFu) << ((id & 3) * 16);
        carry = cur >> 16;
    }
    while (carry) {
        if ((id >> 2) >= cap) { size_t nc = cap * 2; t = (uint64_t *)realloc(t, nc * 8); memset(t + cap, 0, (nc - cap) * 8); cap = nc; }
        t[id >> 2] |= (uint64_t)(carry & 0xFFFFu) << ((id & 3) * 16);
        carry >>= 16;
        ++id;
    }
    for (int k = 0; k < 3; ++k) free(res[k]);
    size_t len = trim(t, cap);
    if (len > rcap) { free(t); return (size_t)-1; }
    memcpy(r, t, len * sizeof(uint64_t));
    free(t);
    return len;
}
