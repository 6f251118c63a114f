/*
 * or_fft.c -- oracle restatement of preclib's mul_fft (src/mul_fft.cpp).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Built with -ffp-contract=off: the reference multiplies and adds separately
 * (its AVX2/SSE2 lanes use mul then add/sub, README.md:177-178), so the same
 * scalar operation order gives the same doubles, and therefore the same
 * rounding-distance statistics, as the reference's COUNT_FFTS build.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

static or_fft_stats g_stats;   /* mul_fft.cpp:32-38 (COUNT_FFTS globals) */
void or_fft_stats_reset(void) { memset(&g_stats, 0, sizeof g_stats); }
void or_fft_stats_get(or_fft_stats *s) { *s = g_stats; }

/* mul_fft.cpp:57-76: recurrence w_i = w_{i-1} * w_1 with an exact cos/sin
 * refresh every 1024 entries. */
static void fill_roots(double *wr, double *wi, size_t half, double ang) {
    if (!half) return;
    double wlr = cos(ang), wli = sin(ang);
    wr[0] = 1.0; wi[0] = 0.0;
    for (size_t i = 1; i < half; ++i) {
        if ((i & 1023) == 0) {
            wr[i] = cos(ang * (double)i);
            wi[i] = sin(ang * (double)i);
        } else {
            double r = wr[i - 1] * wlr - wi[i - 1] * wli;
            double m = wr[i - 1] * wli + wi[i - 1] * wlr;
            wr[i] = r; wi[i] = m;
        }
    }
}

/* One transform: bit reversal then radix-2 DIT levels (mul_fft.cpp:200-220),
 * butterflies of fft_stage (102-172), 1/n scale on the inverse (174-198).
 * Twiddles are regenerated per call; the reference caches them per n
 * (78-100) but computes the same values. */
static void fft(double *re, double *im, size_t n, int inverse) {
    size_t bits = 0;
    while (((size_t)1 << bits) < n) ++bits;
    size_t *rev = (size_t *)malloc(n * sizeof(size_t));
    rev[0] = 0;
    for (size_t i = 1; i < n; ++i) rev[i] = (rev[i >> 1] >> 1) | ((i & 1) << (bits - 1));
    for (size_t i = 1; i < n; ++i) {
        size_t j = rev[i];
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    free(rev);
    const double PI = acos(-1.0);
    double *wr = (double *)malloc((n / 2 + 1) * sizeof(double));
    double *wi = (double *)malloc((n / 2 + 1) * sizeof(double));
    for (size_t len = 2; len <= n; len <<= 1) {
        size_t half = len >> 1;
        fill_roots(wr, wi, half, (inverse ? -2.0 : 2.0) * PI / (double)len);
        for (size_t i = 0; i < n; i += len)
            for (size_t j = 0; j < half; ++j) {
                size_t p = i + j, q = p + half;
                double br = re[q], bi = im[q];
                double vr = br * wr[j] - bi * wi[j];
                double vi = br * wi[j] + bi * wr[j];
                double ur = re[p], ui = im[p];
                re[p] = ur + vr; im[p] = ui + vi;
                re[q] = ur - vr; im[q] = ui - vi;
            }
    }
    free(wr); free(wi);
    if (inverse) {
        double s = 1.0 / (double)n;
        for (size_t i = 0; i < n; ++i) { re[i] *= s; im[i] *= s; }
    }
}

/* mul_fft.cpp:222-232 */
static size_t digit_count(const uint64_t *a, size_t an, int bits) {
    if (!an) return 0;
    size_t per = 64 / bits, n = (an - 1) * per;
    uint64_t top = a[an - 1];
    do { ++n; top >>= bits; } while (top);
    return n;
}

/* mul_fft.cpp:234-246 */
static void load_digits(double *dst, const uint64_t *a, size_t an, int bits) {
    uint64_t mask = ((uint64_t)1 << bits) - 1;
    size_t per = 64 / bits, k = 0;
    for (size_t i = 0; i < an; ++i) {
        uint64_t x = a[i];
        for (size_t j = 0; j < per; ++j) { dst[k++] = (double)(x & mask); x >>= bits; }
    }
}

static long long fft_round(double x) { return (long long)(x >= 0.0 ? x + 0.5 : x - 0.5); }  /* 307-309 */

static size_t trim(const uint64_t *r, size_t n) { while (n && r[n - 1] == 0) --n; return n; }

size_t or_mul_fft_bits(uint64_t *r, size_t rcap, const uint64_t *a, size_t an,
                       const uint64_t *b, size_t bn, int bits, double *max_round) {
    if (max_round) *max_round = 0.0;
    size_t da = digit_count(a, an, bits), db = digit_count(b, bn, bits);
    if (!da || !db) return 0;
    size_t n = 1;
    while (n < da + db) n <<= 1;
    size_t per = 64 / bits;
    size_t alloc = n > an * per ? n : an * per;
    if (alloc < bn * per) alloc = bn * per;
    double *fr = (double *)calloc(alloc, sizeof(double));
    double *fi = (double *)calloc(alloc, sizeof(double));
    load_digits(fr, a, an, bits);           /* a -> real part (350) */
    load_digits(fi, b, bn, bits);           /* b -> imaginary part (351) */
    fft(fr, fi, n, 0);

    /* fft_pair_convolution (248-305): A_k = (X_k + conj X_{n-k})/2,
     * B_k = (X_k - conj X_{n-k})/(2i), out_k = A_k * B_k. */
    double *pr = (double *)malloc(n * sizeof(double));
    double *pi = (double *)malloc(n * sizeof(double));
    for (size_t k = 0; k < n; ++k) {
        size_t j = (n - k) & (n - 1);
        double xr = fr[k], xi = fi[k], yr = fr[j], yi = fi[j];
        double ar = (xr + yr) * 0.5, ai = (xi - yi) * 0.5;
        double br = (xi + yi) * 0.5, bi = (yr - xr) * 0.5;
        pr[k] = ar * br - ai * bi;
        pi[k] = ar * bi + ai * br;
    }
    free(fr); free(fi);
    fft(pr, pi, n, 1);

    /* rounding and COUNT_FFTS statistics (359-382) */
    long long *d = (long long *)malloc((n + 8) * sizeof(long long));
    double mx = 0.0;
    int z8 = 0, z4 = 0, z38 = 0;
    for (size_t i = 0; i < n; ++i) {
        double v = pr[i];
        double err = fabs(v - nearbyint(v));
        if (!isfinite(v)) err = INFINITY;
        if (err > mx) mx = err;
        if (err > 0.125) z8 = 1;
        if (err > 0.25) z4 = 1;
        if (err > 0.375) z38 = 1;
        d[i] = fft_round(v);
    }
    free(pr); free(pi);
    g_stats.total++;
    if (mx > g_stats.max_round) g_stats.max_round = mx;
    g_stats.danger_1_8 += z8; g_stats.danger_1_4 += z4; g_stats.danger_3_8 += z38;
    if (max_round) *max_round = mx;

    /* fft_from_digits (311-338): signed carry in base 2^bits, then pack. */
    uint64_t mask = ((uint64_t)1 << bits) - 1;
    size_t nd = n;
    long long carry = 0;
    for (size_t i = 0; i < nd; ++i) {
        long long cur = d[i] + carry;
        d[i] = cur & (long long)mask;
        carry = cur >> bits;
    }
    while (carry) {
        d = (long long *)realloc(d, (nd + 1) * sizeof(long long));
        d[nd++] = carry & (long long)mask;
        carry >>= bits;
    }
    while (nd && d[nd - 1] == 0) --nd;
    size_t limbs = (nd + per - 1) / per;
    uint64_t *t = (uint64_t *)calloc(limbs + 1, sizeof(uint64_t));
    for (size_t i = 0; i < nd; ++i) t[i / per] |= (uint64_t)d[i] << ((i % per) * bits);
    free(d);
    size_t len = trim(t, limbs);
    if (len > rcap) { free(t); return (size_t)-1; }
    memcpy(r, t, len * sizeof(uint64_t));
    free(t);
    return len;
}

size_t or_mul_fft(uint64_t *r, size_t rcap, const uint64_t *a, size_t an,
                  const uint64_t *b, size_t bn, double *max_round) {   /* 386-394 */
    while (an && a[an - 1] == 0) --an;
    while (bn && b[bn - 1] == 0) --bn;
    if (max_round) *max_round = 0.0;
    if (!an || !bn) return 0;
    size_t hi = an > bn ? an : bn;
    if (hi <= 192) return or_mul_basic(r, rcap, a, an, b, bn);
    return or_mul_fft_bits(r, rcap, a, an, b, bn, hi <= 1024 ? 16 : 8, max_round);
}
