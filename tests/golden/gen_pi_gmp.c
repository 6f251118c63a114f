/*
 * gen_pi_gmp.c -- the digit string of preclib's pi program, computed with GMP.
 *
 * Golden-vector generator (runs in a container that has GMP; the GPU box does
 * not need it).  It evaluates the same exact integer formula as
 * pi/pi_chudnovsky.cpp:162-193:
 *   w = d + 10 guard digits, terms = w / 14 + 2;
 *   (Q, T) = Chudnovsky binary splitting over [0, terms) (leaf 117-135,
 *            merge 137-147; P, Q, T are split-independent);
 *   s = floor(sqrt(10005 * 10^(2w)));
 *   x = floor(floor(Q * 426880 * s / T) / 10^10);
 *   output = decimal(x) with '.' before the last d digits, then "\n" (229).
 * Usage: gen_pi_gmp DIGITS OUTFILE
 */
#include <gmp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static const unsigned long long C3_OVER_24 = 10939058860032000ULL;
static const unsigned long CHUD_A = 13591409UL, CHUD_B = 545140134UL;

static void set_u64(mpz_t z, unsigned long long v) { mpz_import(z, 1, -1, 8, 0, 0, &v); }

static void bs(mpz_t P, mpz_t Q, mpz_t T, unsigned long a, unsigned long b, int need_p) {
    if (b - a == 1) {
        if (a == 0) {
            mpz_set_ui(P, 1);
            mpz_set_ui(Q, 1);
            mpz_set_ui(T, CHUD_A);
            return;
        }
        mpz_t t;
        mpz_init(t);
        set_u64(P, 6ULL * a - 5);
        set_u64(t, 2ULL * a - 1);
        mpz_mul(P, P, t);
        set_u64(t, 6ULL * a - 1);
        mpz_mul(P, P, t);
        set_u64(Q, a);
        mpz_pow_ui(Q, Q, 3);
        set_u64(t, C3_OVER_24);
        mpz_mul(Q, Q, t);
        set_u64(t, (unsigned long long)CHUD_B * a + CHUD_A);
        mpz_mul(T, P, t);
        if (a & 1) mpz_neg(T, T);
        mpz_clear(t);
        return;
    }
    const unsigned long m = a + (b - a) / 2;
    mpz_t lp, lq, lt, rp, rq, rt;
    mpz_inits(lp, lq, lt, rp, rq, rt, NULL);
    bs(lp, lq, lt, a, m, 1);
    bs(rp, rq, rt, m, b, need_p);
    mpz_mul(Q, lq, rq);
    mpz_mul(T, lt, rq);
    mpz_mul(rt, rt, lp);
    mpz_add(T, T, rt);
    if (need_p) mpz_mul(P, lp, rp);
    mpz_clears(lp, lq, lt, rp, rq, rt, NULL);
}

int main(int argc, char **argv) {
    if (argc < 3) return 2;
    const unsigned long d = strtoul(argv[1], NULL, 10), w = d + 10, terms = w / 14 + 2;
    mpz_t P, Q, T, s, x;
    mpz_inits(P, Q, T, s, x, NULL);
    bs(P, Q, T, 0, terms, 0);
    mpz_ui_pow_ui(s, 10, 2 * w);
    mpz_mul_ui(s, s, 10005);
    mpz_sqrt(s, s);
    mpz_mul_ui(x, Q, 426880);
    mpz_mul(x, x, s);
    mpz_tdiv_q(x, x, T);
    mpz_tdiv_q_ui(x, x, 10000000000UL);
    char *str = mpz_get_str(NULL, 10, x);
    size_t n = strlen(str);
    FILE *f = fopen(argv[2], "wb");
    if (!f) return 1;
    if (d == 0) {
        fputs(str, f);
    } else {
        /* pad to d + 1 characters, then '.' before the last d (190-191) */
        const size_t pad = n <= d ? d + 1 - n : 0, len = n + pad;
        char *s2 = malloc(len + 1);
        memset(s2, '0', pad);
        memcpy(s2 + pad, str, n + 1);
        fwrite(s2, 1, len - d, f);
        fputc('.', f);
        fputs(s2 + len - d, f);
        free(s2);
    }
    fputc('\n', f);
    fclose(f);
    return 0;
}
