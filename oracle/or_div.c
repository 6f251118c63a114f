/*
 * or_div.c -- CPU restatement of preclib's division results.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The quotient floor(a / b) and the remainder are unique, so the oracle uses
 * the textbook long division the reference keeps as its schoolbook path
 * (div_schoolbook_impl, src/div_basic.cpp:200-306: Knuth's Algorithm D with
 * a normalising shift, two-limb quotient estimate and add-back).  The GPU's
 * Newton division (div_newton.cpp) must agree with it exactly.
 */
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

typedef unsigned __int128 u128;

static size_t trim(const uint64_t *a, size_t n) {
    while (n && !a[n - 1]) --n;
    return n;
}

static int cmp(const uint64_t *a, size_t an, const uint64_t *b, size_t bn) {
    an = trim(a, an);
    bn = trim(b, bn);
    if (an != bn) return an < bn ? -1 : 1;
    for (size_t i = an; i-- > 0;)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}

size_t or_divmod(uint64_t *q, size_t qcap, size_t *qlen, uint64_t *r, size_t rcap, size_t *rlen,
                 const uint64_t *a, size_t an, const uint64_t *b, size_t bn) {
    an = trim(a, an);
    bn = trim(b, bn);
    *qlen = 0;
    if (rlen) *rlen = 0;
    if (!an || !bn) return 0;                       /* div_basic.cpp:202-205 */
    if (cmp(a, an, b, bn) < 0) {                    /* 206-209 */
        if (r) {
            if (rcap < an) return (size_t)-1;
            memcpy(r, a, an * 8);
            *rlen = an;
        }
        return 0;
    }
    if (bn == 1) {                                  /* 210-213: one-limb divisor */
        if (qcap < an) return (size_t)-1;
        u128 rem = 0;
        for (size_t i = an; i-- > 0;) {
            const u128 cur = (rem << 64) | a[i];
            q[i] = (uint64_t)(cur / b[0]);
            rem = cur % b[0];
        }
        *qlen = trim(q, an);
        if (r) {
            if (rcap < 1) return (size_t)-1;
            r[0] = (uint64_t)rem;
            *rlen = rem ? 1 : 0;
        }
        return 0;
    }
    const size_t n = bn, m = an - bn;
    if (qcap < m + 1) return (size_t)-1;
    const int s = __builtin_clzll(b[n - 1]);
    uint64_t *vn = malloc(n * 8), *un = calloc(an + 1, 8);
    for (size_t i = n; i-- > 0;) vn[i] = (b[i] << s) | (s && i ? b[i - 1] >> (64 - s) : 0);
    un[an] = s ? a[an - 1] >> (64 - s) : 0;
    for (size_t i = an; i-- > 0;) un[i] = (a[i] << s) | (s && i ? a[i - 1] >> (64 - s) : 0);
    for (size_t j = m + 1; j-- > 0;) {
        /* estimate from the top two limbs, refined with the next one */
        const u128 top = ((u128)un[j + n] << 64) | un[j + n - 1];
        u128 qhat = top / vn[n - 1], rhat = top % vn[n - 1];
        while (qhat >> 64) {   /* qhat <= 2^64 + 1 since un[j + n] <= vn[n - 1] */
            --qhat;
            rhat += vn[n - 1];
        }
        while (!(rhat >> 64) && qhat * vn[n - 2] > ((rhat << 64) | un[j + n - 2])) {
            --qhat;
            rhat += vn[n - 1];
        }
        /* un[j .. j+n] -= qhat * vn */
        uint64_t borrow = 0, carry = 0;
        for (size_t i = 0; i < n; ++i) {
            const u128 p = (u128)(uint64_t)qhat * vn[i] + carry;
            carry = (uint64_t)(p >> 64);
            const uint64_t lo = (uint64_t)p;
            const uint64_t t = un[i + j] - lo - borrow;
            borrow = (un[i + j] < lo) || (un[i + j] - lo < borrow);
            un[i + j] = t;
        }
        const uint64_t t = un[j + n] - carry - borrow;
        const int neg = (un[j + n] < carry) || (un[j + n] - carry < borrow);
        un[j + n] = t;
        if (neg) {   /* add back */
            --qhat;
            uint64_t c = 0;
            for (size_t i = 0; i < n; ++i) {
                const u128 sum = (u128)un[i + j] + vn[i] + c;
                un[i + j] = (uint64_t)sum;
                c = (uint64_t)(sum >> 64);
            }
            un[j + n] += c;
        }
        q[j] = (uint64_t)qhat;
    }
    *qlen = trim(q, m + 1);
    if (r) {
        if (rcap < n) {
            free(vn);
            free(un);
            return (size_t)-1;
        }
        for (size_t i = 0; i < n; ++i) r[i] = (un[i] >> s) | (s ? un[i + 1] << (64 - s) : 0);
        *rlen = trim(r, n);
    }
    free(vn);
    free(un);
    return 0;
}

/* precn_reciprocal_newton's result (div_newton.cpp:70-87): floor(2^bits / b),
 * 0 when b == 0 or bits < bit_length(b). */
size_t or_recip(uint64_t *x, size_t xcap, const uint64_t *b, size_t bn, size_t bits) {
    bn = trim(b, bn);
    if (!bn) return 0;
    const size_t nb = (bn - 1) * 64 + 64 - (size_t)__builtin_clzll(b[bn - 1]);
    if (bits < nb) return 0;
    const size_t pn = bits / 64 + 1;
    uint64_t *p = calloc(pn, 8);
    p[bits / 64] = 1ull << (bits % 64);
    size_t ql = 0;
    const size_t st = or_divmod(x, xcap, &ql, NULL, 0, NULL, p, pn, b, bn);
    free(p);
    return st == (size_t)-1 ? st : ql;
}

/* precn_sqrt's result (src/sqrt.cpp:71-86): floor(sqrt(a)), by Newton's
 * iteration x <- (x + a / x) / 2 from an upper bound (sqrt.cpp:41-60). */
size_t or_sqrt(uint64_t *x, size_t xcap, const uint64_t *a, size_t an) {
    an = trim(a, an);
    if (!an) return 0;
    const size_t nbits = (an - 1) * 64 + 64 - (size_t)__builtin_clzll(a[an - 1]);
    const size_t xn = (nbits + 1) / 2 / 64 + 2;
    if (xcap < xn) return (size_t)-1;
    uint64_t *cur = calloc(xn + 1, 8), *qq = calloc(an + 1, 8), *nx = calloc(xn + 2, 8);
    const size_t e = (nbits + 1) / 2 + 1;   /* 2^e > sqrt(a) */
    cur[e / 64] = 1ull << (e % 64);
    size_t cl = e / 64 + 1;
    for (;;) {
        size_t ql = 0;
        or_divmod(qq, an + 1, &ql, NULL, 0, NULL, a, an, cur, cl);
        /* nx = (cur + qq) >> 1 */
        const size_t L = (cl > ql ? cl : ql) + 1;
        uint64_t c = 0;
        for (size_t i = 0; i < L; ++i) {
            const u128 sum = (u128)(i < cl ? cur[i] : 0) + (i < ql ? qq[i] : 0) + c;
            nx[i] = (uint64_t)sum;
            c = (uint64_t)(sum >> 64);
        }
        for (size_t i = 0; i < L; ++i) nx[i] = (nx[i] >> 1) | (i + 1 < L ? nx[i + 1] << 63 : 0);
        const size_t nl = trim(nx, L);
        if (cmp(nx, nl, cur, cl) >= 0) break;
        memset(cur, 0, (xn + 1) * 8);
        memcpy(cur, nx, nl * 8);
        cl = nl;
    }
    memcpy(x, cur, cl * 8);
    free(cur);
    free(qq);
    free(nx);
    return trim(x, cl);
}
