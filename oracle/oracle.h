/*
 * oracle.h -- CPU restatement of preclib's large-integer product path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (preclib_amd/, include/)
 * may link, load or call this code.  Only tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py use it, and only as the checker.
 *
 * Every routine restates one reference function (file:line given beside it).
 * Limb arrays are little-endian uint64, trimmed (top limb nonzero) unless a
 * routine says otherwise.  Output routines write at most rcap limbs and return
 * the trimmed product length, or (size_t)-1 if rcap is too small.
 */
#ifndef PRECLIB_ORACLE_H
#define PRECLIB_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- NTT (src/mul_ntt.cpp) ------------------------------------------- */
/* prime index 0,1,2 -> 469762049, 1811939329, 2013265921 (mul_ntt.cpp:29-41) */
uint32_t or_ntt_prime(int idx);
/* Montgomery constants {mod, root, ninv, r2, one} (mul_ntt.cpp:37-41), derived
 * from the primes at run time and compared against the frozen table. */
int or_ntt_ctx(int idx, uint32_t out[5]);
uint32_t or_mont_mul(int idx, uint32_t a, uint32_t b);           /* mul_ntt.cpp:57-59 */
/* In-place forward DIF / inverse DIT on Montgomery-form data, n a power of
 * two (mul_ntt.cpp:193-255, 257-329).  The inverse includes the 1/n scale. */
void or_ntt_forward(int idx, uint32_t *a, size_t n);
void or_ntt_inverse(int idx, uint32_t *a, size_t n);
/* Cyclic convolution of plain digit vectors mod prime idx, length n
 * (mul_ntt.cpp:349-387): out[i] = sum_{j+k=i mod n} a[j] b[k] mod p. */
void or_ntt_convolve(int idx, const uint32_t *a, size_t na, const uint32_t *b,
                     size_t nb, size_t n, uint32_t *out);
uint64_t or_ntt_crt2(uint32_t r1, uint32_t r2);                  /* mul_ntt.cpp:405-412 */
/* Base-2^16 digits of a limb array with trailing zero digits trimmed
 * (mul_ntt.cpp:331-342).  Returns the digit count; out needs 4*an slots. */
size_t or_ntt_digits(const uint64_t *a, size_t an, uint32_t *out);
/* mul_ntt (mul_ntt.cpp:494-518), including its <=192 and size-cap fallbacks. */
size_t or_mul_ntt(uint64_t *r, size_t rcap, const uint64_t *a, size_t an,
                  const uint64_t *b, size_t bn);
/* Transform length and prime count mul_ntt would use (0 if it would not
 * transform). */
size_t or_ntt_plan(size_t an, const uint64_t *a, size_t bn, const uint64_t *b,
                   int *primes);

/* ---- FFT (src/mul_fft.cpp) ------------------------------------------- */
typedef struct {
    double max_round;    /* max |v - nearbyint(v)| over every coefficient    */
    uint64_t total;      /* number of transform products (COUNT_FFTS totals) */
    uint64_t danger_1_8, danger_1_4, danger_3_8;
} or_fft_stats;
void or_fft_stats_reset(void);
void or_fft_stats_get(or_fft_stats *s);
/* mul_fft (mul_fft.cpp:386-394); *max_round receives this call's maximum
 * rounding distance (0 if no transform ran). */
size_t or_mul_fft(uint64_t *r, size_t rcap, const uint64_t *a, size_t an,
                  const uint64_t *b, size_t bn, double *max_round);
/* mul_fft_bits with an explicit digit width (mul_fft.cpp:340-384). */
size_t or_mul_fft_bits(uint64_t *r, size_t rcap, const uint64_t *a, size_t an,
                       const uint64_t *b, size_t bn, int bits, double *max_round);

/* ---- schoolbook, blocked, dispatcher (src/mul_basic.cpp) ------------- */
size_t or_mul_basic(uint64_t *r, size_t rcap, const uint64_t *a, size_t an,
                    const uint64_t *b, size_t bn);                  /* 131-185 */
/* mul_into (238-276); sqr != 0 models the &a == &b alias that routes to
 * precn_sqr (193-205, 244-246). */
size_t or_mul_into(uint64_t *r, size_t rcap, const uint64_t *a, size_t an,
                   const uint64_t *b, size_t bn, int sqr);
size_t or_sqr(uint64_t *r, size_t rcap, const uint64_t *a, size_t an);

/* ---- division (src/div_basic.cpp, src/div_newton.cpp, src/sqrt.cpp) --- */
/* floor(a / b) and a mod b, both trimmed (lengths in *qlen / *rlen; r may be
 * NULL).  b == 0 or a == 0 gives empty results (div_newton.cpp:90-93).
 * Returns 0, or (size_t)-1 if a capacity is too small. */
size_t or_divmod(uint64_t *q, size_t qcap, size_t *qlen, uint64_t *r, size_t rcap, size_t *rlen,
                 const uint64_t *a, size_t an, const uint64_t *b, size_t bn);
/* floor(2^bits / b); 0 when b == 0 or bits < bit_length(b) (div_newton.cpp:70-87). */
size_t or_recip(uint64_t *x, size_t xcap, const uint64_t *b, size_t bn, size_t bits);
/* floor(sqrt(a)) (sqrt.cpp:71-86). */
size_t or_sqrt(uint64_t *x, size_t xcap, const uint64_t *a, size_t an);

/* ---- checks ------------------------------------------------------------ */
/* (sum a[i] 2^(64 i)) mod p for a pseudo-Mersenne modulus p = 2^k - c,
 * 32 <= k <= 64, 1 <= c < 2^16 (Horner from the top limb; 2^64 = c 2^(64-k)
 * mod p, so every step is a small product and a fold, no division).
 * No reference counterpart: the result checks of the serving tests
 * (C == A B and A == Q B + R modulo several primes). */
uint64_t or_mod_p(const uint64_t *a, size_t n, int k, uint64_t c);
/* The same for the three moduli 2^61 - 1, 2^63 - 25 and 2^64 - 59 in one
 * pass (independent chains); out[3]. */
void or_mod_p3(const uint64_t *a, size_t n, uint64_t out[3]);
/* The value of the decimal digits s[0, n) ('0'..'9', most significant
 * first) modulo the same three moduli, by Horner's rule on the digits;
 * checks the decimal parse without a second parser. */
void or_dec_mod_p3(const char *s, size_t n, uint64_t out[3]);

/* ---- synthetic operands ------------------------------------------------ */
/* splitmix64 stream (Vigna's published generator). dist: 0 uniform, 1 all
 * ones, 2 sparse (1% nonzero).  The top limb is always nonzero. */
void or_fill(uint64_t *a, size_t n, uint64_t seed, int dist);

#ifdef __cplusplus
}
#endif
#endif
