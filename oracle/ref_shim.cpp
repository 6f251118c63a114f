// ref_shim.cpp -- extern "C" wrapper over the reference's own product path,
// compiled together with the reference sources into oracle/_ref/.
// TEST INFRASTRUCTURE ONLY: used to pin the oracle restatement and as the
// reference CPU baseline.  This file is this project's code; it includes the
// reference's public header (prec.hpp) from the reference tree.
#include "prec.hpp"
#include <cstring>

static precn_t from_limbs(const uint64_t *a, size_t n) {
    precn_t r;
    if (n) {
        r.a = (uint64_t *)realloc(r.a, n * sizeof(uint64_t));
        r.asiz = n;
        memcpy(r.a, a, n * sizeof(uint64_t));
        r.rsiz = n;
        while (r.rsiz && r.a[r.rsiz - 1] == 0) --r.rsiz;
        if (!r.rsiz) r.a[0] = 0;
    }
    return r;
}

static size_t to_limbs(const precn_t &x, uint64_t *r, size_t rcap) {
    if (x.rsiz > rcap) return (size_t)-1;
    if (x.rsiz) memcpy(r, x.a, x.rsiz * sizeof(uint64_t));
    return x.rsiz;
}

extern "C" {

size_t ref_mul_into(uint64_t *r, size_t rcap, const uint64_t *a, size_t an,
                    const uint64_t *b, size_t bn, int sqr) {
    precn_t x = from_limbs(a, an);
    precn_t out;
    if (sqr) {
        mul_into(out, x, x);
    } else {
        precn_t y = from_limbs(b, bn);
        mul_into(out, x, y);
    }
    return to_limbs(out, r, rcap);
}

size_t ref_mul_ntt(uint64_t *r, size_t rcap, const uint64_t *a, size_t an,
                   const uint64_t *b, size_t bn) {
    precn_t x = from_limbs(a, an), y = from_limbs(b, bn);
    return to_limbs(mul_ntt(x, y), r, rcap);
}

size_t ref_mul_fft(uint64_t *r, size_t rcap, const uint64_t *a, size_t an,
                   const uint64_t *b, size_t bn, double *max_round) {
    precn_t x = from_limbs(a, an), y = from_limbs(b, bn);
    max_fft_rounding_error = 0.0;
    precn_t p = mul_fft(x, y);
    if (max_round) *max_round = max_fft_rounding_error;
    return to_limbs(p, r, rcap);
}

size_t ref_mul_basic(uint64_t *r, size_t rcap, const uint64_t *a, size_t an,
                     const uint64_t *b, size_t bn) {
    precn_t x = from_limbs(a, an), y = from_limbs(b, bn);
    return to_limbs(mul_basic(x, y), r, rcap);
}

size_t ref_div(uint64_t *q, size_t qcap, const uint64_t *a, size_t an,
               const uint64_t *b, size_t bn) {
    precn_t x = from_limbs(a, an), y = from_limbs(b, bn);
    return to_limbs(x / y, q, qcap);
}

size_t ref_div_mulinv(uint64_t *q, size_t qcap, const uint64_t *a, size_t an,
                      const uint64_t *b, size_t bn) {
    precn_t x = from_limbs(a, an), y = from_limbs(b, bn);
    return to_limbs(div_mulinv(x, y), q, qcap);
}

size_t ref_reciprocal_newton(uint64_t *x, size_t xcap, const uint64_t *b, size_t bn,
                             size_t bits) {
    precn_t y = from_limbs(b, bn);
    return to_limbs(precn_reciprocal_newton(y, bits), x, xcap);
}

size_t ref_sqrt(uint64_t *x, size_t xcap, const uint64_t *a, size_t an) {
    precn_t y = from_limbs(a, an);
    return to_limbs(precn_sqrt(y), x, xcap);
}

}  // extern "C"
