// pg_pi.h -- binary splitting, square root, decimal conversion and the pi
// program on device operands (pg_pi.hip).
#pragma once
#include <string>
#include <vector>
#include "pg_dint.h"

namespace pg {

// P, Q, T of Chudnovsky's series over terms [0, nterms) (pi_chudnovsky.cpp:
// 117-148); T >= 0 for every range that starts at term 0.  P is skipped
// when need_p is false.  keep holds the tree's storage until the caller drops it.
int chud_split(size_t nterms, DBuf &P, size_t &lp, DBuf &Q, size_t &lq, DBuf &T, size_t &lt, bool need_p,
               std::vector<DBuf> &keep);
// floor(sqrt(a)), a trimmed (precn_sqrt, src/sqrt.cpp:71-86).
int isqrt(DBuf &x, size_t &xl, const uint64_t *a, size_t an);
// Decimal digits of x, most significant first, no leading zeros; empty for
// zero (precn_base_convert with base 10, src/base_convert.cpp:107-172).
// point > 0: formatted as the pi program prints x / 10^point instead
// (src/pi_chudnovsky.cpp:186-192): at least one integer digit, '.', then
// `point` digits ("0" for zero before the padding).  With dst set the string
// goes there instead of `out` (PG_EINVAL past cap; *dlen its length either way).
int to_decimal(const uint64_t *x, size_t xn, std::string &out, size_t point = 0, char *dst = nullptr,
               size_t cap = 0, size_t *dlen = nullptr);
// The value of a string of n decimal digits ('0'..'9' only, most significant
// first), trimmed limbs (empty for zero): precn_t(std::string),
// src/init_basic.cpp:66-92.
int from_decimal(const char *digits, size_t n, std::vector<uint64_t> &out);
// The output of pi_chudnovsky DIGITS (pi_chudnovsky.cpp:162-193) without the
// trailing newline; phase_ms (if set) gets binary_split, sqrt_scale,
// final_div and to_decimal times.
int pi_digits(size_t digits, std::string &out, double phase_ms[4], char *dst = nullptr, size_t cap = 0,
              size_t *dlen = nullptr);

}  // namespace pg
