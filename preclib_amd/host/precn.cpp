// precn.cpp -- precn_t (include/precn.hpp) over the libpgpu.so C ABI.
//
// Lifecycle follows src/init_basic.cpp:47-135 (malloc / realloc / free,
// moved-from objects hold a == nullptr).  Products, divisions by divisors of
// two or more limbs, square roots and decimal conversion in both directions
// go through pg_mul / pg_div / pg_recip / pg_sqrt / pg_to_decimal /
// pg_from_decimal; the dispatcher rules of mul_into (mul_basic.cpp:238-276)
// are applied by the library (PG_PATH_AUTO): the reference's structure, with
// the NTT where preclib's CPU thresholds pick mul_fft (pgpu.h).  The
// reference's alternative multipliers (mul_basic, mul_karatsuba, mul_toom*)
// compute the same product and all take that path.  Additions,
// subtractions, shifts, comparisons and products or quotients by one
// machine word are the linear host steps of src/add.cpp, src/sub.cpp,
// src/compare_shift.cpp, src/mul_u32.cpp and src/div_basic.cpp:135-198.
#include "precn.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "pgpu.h"

namespace {

std::once_flag g_init;
int g_init_status = PG_OK;

void ensure_init() {
    std::call_once(g_init, [] { g_init_status = pg_init(0, 0); });
    if (g_init_status != PG_OK) throw precn_gpu_error(g_init_status, pg_last_error());
}

void check(int st) {
    if (st != PG_OK) throw precn_gpu_error(st, pg_last_error());
}

// mul_reserve (mul_basic.cpp:14-18): room for n limbs, keeping contents.
void reserve(precn_t &r, size_t n) {
    n = std::max<size_t>(n, 1);
    if (r.a && r.asiz >= n) return;
    r.a = (uint64_t *)realloc(r.a, n * sizeof(uint64_t));
    r.asiz = n;
}

void set_zero(precn_t &r) {
    reserve(r, 1);
    r.rsiz = 0;
    r.a[0] = 0;
}

void trim(precn_t &r) {
    while (r.rsiz && !r.a[r.rsiz - 1]) --r.rsiz;
    if (!r.rsiz) r.a[0] = 0;
}

void gpu_mul(precn_t &r, const precn_t &a, const precn_t &b, int path, double *max_round = nullptr) {
    ensure_init();
    if (max_round) *max_round = 0.0;
    if (!a.rsiz || !b.rsiz) {   // mul_basic.cpp:239-242
        set_zero(r);
        return;
    }
    // r may alias a or b: pg_mul reads both operands before it writes r, and
    // the reserve keeps the aliased limbs in place.
    const size_t need = a.rsiz + b.rsiz;
    std::vector<uint64_t> tmp;
    uint64_t *dst;
    if (&r == &a || &r == &b) {
        tmp.resize(need);
        dst = tmp.data();
    } else {
        reserve(r, need);
        dst = r.a;
    }
    size_t len = 0;
    check(pg_mul(dst, need, &len, a.a, a.rsiz, b.a, b.rsiz, path, max_round));
    if (!tmp.empty()) {
        reserve(r, need);
        std::memcpy(r.a, tmp.data(), len * sizeof(uint64_t));
    }
    r.rsiz = len;
    if (!len) r.a[0] = 0;
}

int cmp(const precn_t &a, const precn_t &b) {
    if (a.rsiz != b.rsiz) return a.rsiz < b.rsiz ? -1 : 1;
    for (size_t i = a.rsiz; i-- > 0;)
        if (a.a[i] != b.a[i]) return a.a[i] < b.a[i] ? -1 : 1;
    return 0;
}

precn_t from_limbs(const uint64_t *p, size_t n) {
    precn_t r;
    reserve(r, n);
    if (n) std::memcpy(r.a, p, n * sizeof(uint64_t));
    r.rsiz = n;
    trim(r);
    return r;
}

}  // namespace

void precn_gpu_init(int device, size_t hbm_budget) {
    std::call_once(g_init, [&] { g_init_status = pg_init(device, hbm_budget); });
    if (g_init_status != PG_OK) throw precn_gpu_error(g_init_status, pg_last_error());
}

// ---- lifecycle (init_basic.cpp:47-135) -------------------------------------------

precn_t::precn_t() {
    a = (uint64_t *)malloc(sizeof(uint64_t));
    asiz = 1;
    rsiz = 0;
    *a = 0;
}

precn_t::precn_t(const precn_t &o) {
    asiz = std::max<size_t>(o.rsiz, 1);
    rsiz = o.rsiz;
    a = (uint64_t *)malloc(asiz * sizeof(uint64_t));
    if (rsiz) std::memcpy(a, o.a, rsiz * sizeof(uint64_t));
    else a[0] = 0;
}

precn_t::precn_t(precn_t &&o) {
    asiz = o.asiz;
    rsiz = o.rsiz;
    a = o.a;
    o.a = nullptr;
    o.asiz = 0;
    o.rsiz = 0;
}

precn_t &precn_t::operator=(const precn_t &o) {
    if (this == &o) return *this;
    asiz = std::max<size_t>(o.rsiz, 1);
    rsiz = o.rsiz;
    a = (uint64_t *)realloc(a, asiz * sizeof(uint64_t));
    if (rsiz) std::memcpy(a, o.a, rsiz * sizeof(uint64_t));
    else a[0] = 0;
    return *this;
}

precn_t &precn_t::operator=(precn_t &&o) {
    if (this == &o) return *this;
    free(a);
    asiz = o.asiz;
    rsiz = o.rsiz;
    a = o.a;
    o.a = nullptr;
    o.asiz = 0;
    o.rsiz = 0;
    return *this;
}

precn_t::~precn_t() { free(a); }

precn_t::operator std::string() const {
    if (rsiz == 0) return std::string("0");
    ensure_init();
    std::string s(rsiz * 20 + 2, '\0');
    size_t n = 0;
    check(pg_to_decimal(&s[0], s.size(), &n, a, rsiz));
    s.resize(n);
    return s;
}

// ---- products ----------------------------------------------------------------------

void mul_into(precn_t &r, const precn_t &a, const precn_t &b) {
    if (&a == &b) {   // mul_basic.cpp:244-246: an aliased square goes to precn_sqr
        precn_t t = precn_sqr(a);
        r = std::move(t);
        return;
    }
    gpu_mul(r, a, b, PG_PATH_AUTO);
}

precn_t operator*(const precn_t &a, const precn_t &b) {   // mul_basic.cpp:278-284
    precn_t r;
    mul_into(r, a, b);
    return r;
}

precn_t precn_sqr(const precn_t &a) {
    // mul_basic.cpp:193-205: the library sees one operand (same pointer);
    // above 192 limbs it squares on the NTT (preclib: mul_fft)
    precn_t r;
    gpu_mul(r, a, a, PG_PATH_AUTO);
    return r;
}

precn_t mul_ntt(const precn_t &a, const precn_t &b) {
    precn_t r;
    gpu_mul(r, a, b, PG_PATH_NTT);
    return r;
}

// COUNT_FFTS statistics (mul_fft.cpp:32-37, 359-382): one FFT product per
// mul_fft call above 192 limbs (below, mul_fft multiplies with mul_basic),
// classified by its largest rounding distance.
uint64_t total_fftmuls = 0;
uint64_t danger_fftmuls = 0;
uint64_t danger_fftmuls_1_4 = 0;
uint64_t danger_fftmuls_3_8 = 0;
double max_fft_rounding_error = 0.0;

static void count_fft(const precn_t &a, const precn_t &b, double err) {
    if (!a.rsiz || !b.rsiz || std::max(a.rsiz, b.rsiz) <= 192) return;
    ++total_fftmuls;
    if (err > 0.125) ++danger_fftmuls;
    if (err > 0.25) ++danger_fftmuls_1_4;
    if (err > 0.375) ++danger_fftmuls_3_8;
    if (err > max_fft_rounding_error) max_fft_rounding_error = err;
}

precn_t mul_fft(const precn_t &a, const precn_t &b) {
    precn_t r;
    double err = 0.0;
    gpu_mul(r, a, b, PG_PATH_FFT, &err);
    count_fft(a, b, err);
    return r;
}

precn_t mul_fft_round(const precn_t &a, const precn_t &b, double *max_round) {
    precn_t r;
    double err = 0.0;
    gpu_mul(r, a, b, PG_PATH_FFT, &err);
    count_fft(a, b, err);
    if (max_round) *max_round = err;
    return r;
}

precn_t mul_u64(const precn_t &a, uint64_t b) {
    precn_t r;
    if (!a.rsiz || !b) return r;
    reserve(r, a.rsiz + 1);
    unsigned __int128 c = 0;
    for (size_t i = 0; i < a.rsiz; ++i) {
        c += (unsigned __int128)a.a[i] * b;
        r.a[i] = (uint64_t)c;
        c >>= 64;
    }
    r.a[a.rsiz] = (uint64_t)c;
    r.rsiz = a.rsiz + 1;
    trim(r);
    return r;
}

precn_t mul_u32(const precn_t &a, uint32_t b) { return mul_u64(a, b); }

// The reference's alternative multipliers (schoolbook mul_basic.cpp:286,
// Karatsuba mul_kara.cpp:51, Toom-Cook mul_toom.cpp:457-471) return the
// same product as the dispatcher; here they all run on the GPU engine.
precn_t mul_basic(const precn_t &a, const precn_t &b) { return a * b; }
precn_t mul_karatsuba(const precn_t &a, const precn_t &b) { return a * b; }
precn_t mul_toom(const precn_t &a, const precn_t &b) { return a * b; }
precn_t mul_toom23(const precn_t &a, const precn_t &b) { return a * b; }
precn_t mul_toom24(const precn_t &a, const precn_t &b) { return a * b; }
precn_t mul_toom33(const precn_t &a, const precn_t &b) { return a * b; }
// mul_ssa.cpp:153-158: the public entry point multiplies with mul_ntt
precn_t mul_ssa(const precn_t &a, const precn_t &b) { return mul_ntt(a, b); }

precn_t precn_divexact_2(const precn_t &a) { return a >> 1; }

// The 2-adic quotient a * 3^-1 mod 2^(64 a.rsiz), which is a / 3 when 3
// divides a: limb by limb with the inverse of 3 modulo 2^64 and a borrow of
// 0..2 (Jebelean's exact division).
precn_t precn_divexact_3(const precn_t &a) {
    if (!a.rsiz) return precn_t();
    const uint64_t inv3 = 0xAAAAAAAAAAAAAAABull;   // 3 * inv3 = 1 mod 2^64
    precn_t r;
    reserve(r, a.rsiz);
    uint64_t borrow = 0;
    for (size_t i = 0; i < a.rsiz; ++i) {
        const uint64_t x = a.a[i] - borrow;
        const uint64_t q = x * inv3;
        r.a[i] = q;
        // 3 q = x + borrow' 2^64: borrow' = floor(3 q / 2^64) plus the wrap of x
        borrow = (uint64_t)(((unsigned __int128)q * 3) >> 64) + (x > a.a[i]);
    }
    r.rsiz = a.rsiz;
    trim(r);
    return r;
}

// ---- division, square root, conversion -------------------------------------------

namespace {

// q = floor(a / b), r = a mod b for b of two or more limbs (Newton's
// division on the GPU, div_newton.cpp:89-129); either may be null.
void gpu_div(const precn_t &a, const precn_t &b, precn_t *q, precn_t *r) {
    ensure_init();
    std::vector<uint64_t> qb(a.rsiz + 1), rb(b.rsiz + a.rsiz + 1);
    size_t ql = 0, rl = 0;
    check(pg_div(q ? qb.data() : nullptr, qb.size(), &ql, r ? rb.data() : nullptr, rb.size(), &rl, a.a, a.rsiz,
                 b.a, b.rsiz));
    if (q) *q = from_limbs(qb.data(), ql);
    if (r) *r = from_limbs(rb.data(), rl);
}

// Quotient by one machine word (div_basic.cpp:135-178): returns a mod d;
// q (if set) gets floor(a / d).  d == 0 gives 0 for both.
uint64_t div_word(const precn_t &a, uint64_t d, precn_t *q) {
    if (!a.rsiz || !d) {
        if (q) *q = precn_t();
        return 0;
    }
    precn_t t;
    if (q) reserve(t, a.rsiz);
    unsigned __int128 rem = 0;
    for (size_t i = a.rsiz; i-- > 0;) {
        const unsigned __int128 cur = (rem << 64) | a.a[i];
        if (q) t.a[i] = (uint64_t)(cur / d);
        rem = cur % d;
    }
    if (q) {
        t.rsiz = a.rsiz;
        trim(t);
        *q = std::move(t);
    }
    return (uint64_t)rem;
}

// div_into / mod_into / divmod_into (div_basic.cpp:334-407) without the
// aliasing cases: zero operand or divisor -> 0; one-limb divisor on the
// host; a < b -> (0, a); otherwise the GPU.
void divmod(const precn_t &a, const precn_t &b, precn_t *q, precn_t *r) {
    if (!a.rsiz || !b.rsiz) {
        if (q) *q = precn_t();
        if (r) *r = precn_t();
        return;
    }
    if (b.rsiz == 1) {
        const uint64_t rem = div_word(a, b.a[0], q);
        if (r) *r = precn_t(rem);
        return;
    }
    if (cmp(a, b) < 0) {
        if (q) *q = precn_t();
        if (r) *r = a;
        return;
    }
    gpu_div(a, b, q, r);
}

}  // namespace

precn_t div_mulinv(const precn_t &a, const precn_t &b) {
    precn_t q;
    gpu_div(a, b, &q, nullptr);
    return q;
}

precn_t mod_mulinv(const precn_t &a, const precn_t &b) {
    precn_t r;
    gpu_div(a, b, nullptr, &r);
    return r;
}

precn_t div_u64(const precn_t &a, uint64_t b) {
    precn_t q;
    div_word(a, b, &q);
    return q;
}
precn_t div_u32(const precn_t &a, uint32_t b) { return div_u64(a, b); }
precn_t mod_u64(const precn_t &a, uint64_t b) { return precn_t(div_word(a, b, nullptr)); }
precn_t mod_u32(const precn_t &a, uint32_t b) { return mod_u64(a, b); }

void div_into(precn_t &q, const precn_t &a, const precn_t &b) {
    precn_t t;
    divmod(a, b, &t, nullptr);
    q = std::move(t);   // q may alias a or b
}

void mod_into(precn_t &r, const precn_t &a, const precn_t &b) {
    precn_t t;
    divmod(a, b, nullptr, &t);
    r = std::move(t);
}

void divmod_into(precn_t &q, precn_t &r, const precn_t &a, const precn_t &b) {
    if (&q == &r) return;   // div_basic.cpp:335
    precn_t tq, tr;
    divmod(a, b, &tq, &tr);
    q = std::move(tq);
    r = std::move(tr);
}

void divmod_schoolbook_into(precn_t &q, precn_t &r, const precn_t &a, const precn_t &b) {
    divmod_into(q, r, a, b);
}

bool div_dc_into(precn_t &q, precn_t &r, const precn_t &a, const precn_t &b) {
    if (b.rsiz < 64 || a < b || a.rsiz > 2 * b.rsiz) return false;   // div_dc.cpp:93-94
    divmod_into(q, r, a, b);
    return true;
}

precn_t div_schoolbook(const precn_t &a, const precn_t &b) {
    precn_t q;
    divmod(a, b, &q, nullptr);
    return q;
}

precn_t mod_schoolbook(const precn_t &a, const precn_t &b) {
    precn_t r;
    divmod(a, b, nullptr, &r);
    return r;
}

precn_t operator/(const precn_t &a, const precn_t &b) {
    precn_t q;
    div_into(q, a, b);
    return q;
}

precn_t operator%(const precn_t &a, const precn_t &b) {
    precn_t r;
    mod_into(r, a, b);
    return r;
}

precn_t gcd(const precn_t &a, const precn_t &b) {   // gcd.cpp:5-16
    precn_t x(a), y(b);
    while (y.rsiz) {
        precn_t t;
        mod_into(t, x, y);
        x = std::move(y);
        y = std::move(t);
    }
    return x;
}

precn_t precn_reciprocal_newton(const precn_t &b, size_t bits) {
    ensure_init();
    std::vector<uint64_t> x(bits / 64 + 2);
    size_t xl = 0;
    check(pg_recip(x.data(), x.size(), &xl, b.a, b.rsiz, bits));
    return from_limbs(x.data(), xl);
}

precn_t precn_sqrt(const precn_t &a) {
    ensure_init();
    std::vector<uint64_t> x(a.rsiz / 2 + 2);
    size_t xl = 0;
    check(pg_sqrt(x.data(), x.size(), &xl, a.a, a.rsiz));
    return from_limbs(x.data(), xl);
}

precn_t::precn_t(std::string o) {
    a = (uint64_t *)malloc(sizeof(uint64_t));
    asiz = 1;
    rsiz = 0;
    *a = 0;
    bool any = false;
    for (char c : o) any |= c >= '1' && c <= '9';
    if (!any) return;   // no digits, or only zeros
    ensure_init();
    std::vector<uint64_t> v(o.size() * 3322 / 64000 + 2);
    size_t n = 0;
    check(pg_from_decimal(v.data(), v.size(), &n, o.data(), o.size()));
    *this = from_limbs(v.data(), n);
}

namespace {

// Digits of a in base `base` (not a power of two), least significant first:
// split by P^(2^i), P = base^k the largest power below 2^64, down to single
// chunks (base_convert.cpp:56-105's recursion; every division and product
// of two or more limbs runs on the GPU), then each chunk into k digits.
void chunk_digits(const precn_t &x, int level, const std::vector<precn_t> &pw, bool fixed, uint64_t P,
                  std::vector<uint64_t> &chunks) {
    if (level < 0) {
        chunks.push_back(x.rsiz ? x.a[0] : 0);
        return;
    }
    if (!fixed && x < pw[level]) {
        chunk_digits(x, level - 1, pw, false, P, chunks);
        return;
    }
    precn_t q, r;
    divmod_into(q, r, x, pw[level]);
    chunk_digits(r, level - 1, pw, true, P, chunks);
    chunk_digits(q, level - 1, pw, fixed, P, chunks);
}

}  // namespace

void precn_base_convert(const precn_t &a, uint32_t base, uint32_t *out, size_t &out_siz) {
    out_siz = 0;
    if (a.rsiz == 0 || base < 2) return;
    if (!(base & (base - 1))) {   // base_convert.cpp:111-128: bit slicing
        unsigned bits = 0;
        while ((1u << bits) < base) ++bits;
        const size_t nb = (a.rsiz - 1) * 64 + 64 - (size_t)__builtin_clzll(a.a[a.rsiz - 1]);
        out_siz = (nb + bits - 1) / bits;
        if (!out) return;
        const uint64_t mask = ((uint64_t)1 << bits) - 1;
        for (size_t i = 0; i < out_siz; ++i) {
            const size_t bit = i * bits, limb = bit / 64;
            const unsigned off = (unsigned)(bit % 64);
            uint64_t d = a.a[limb] >> off;
            if (off + bits > 64 && limb + 1 < a.rsiz) d |= a.a[limb + 1] << (64 - off);
            out[i] = (uint32_t)(d & mask);
        }
        return;
    }
    if (base == 10) {   // the GPU's conversion tree (pg_to_decimal)
        const std::string s = (std::string)a;
        out_siz = s.size();
        if (out)
            for (size_t i = 0; i < out_siz; ++i) out[i] = (uint32_t)(s[out_siz - 1 - i] - '0');
        return;
    }
    uint64_t P = base;
    size_t k = 1;
    while (P <= UINT64_MAX / base) {
        P *= base;
        ++k;
    }
    std::vector<precn_t> pw{precn_t(P)};
    while (pw.back().rsiz * 2 <= a.rsiz + 1) pw.push_back(precn_sqr(pw.back()));
    std::vector<uint64_t> chunks;
    chunk_digits(a, (int)pw.size() - 1, pw, false, P, chunks);
    std::vector<uint32_t> digits;
    digits.reserve(chunks.size() * k);
    for (uint64_t c : chunks)
        for (size_t i = 0; i < k; ++i) {
            digits.push_back((uint32_t)(c % base));
            c /= base;
        }
    while (!digits.empty() && digits.back() == 0) digits.pop_back();
    out_siz = digits.size();
    if (out) std::copy(digits.begin(), digits.end(), out);
}

// ---- linear steps ------------------------------------------------------------------

bool operator==(const precn_t &a, const precn_t &b) { return cmp(a, b) == 0; }
bool operator!=(const precn_t &a, const precn_t &b) { return cmp(a, b) != 0; }
bool operator<(const precn_t &a, const precn_t &b) { return cmp(a, b) < 0; }
bool operator>(const precn_t &a, const precn_t &b) { return cmp(a, b) > 0; }
bool operator<=(const precn_t &a, const precn_t &b) { return cmp(a, b) <= 0; }
bool operator>=(const precn_t &a, const precn_t &b) { return cmp(a, b) >= 0; }

precn_t operator+(const precn_t &a, const precn_t &b) {
    const precn_t &x = a.rsiz >= b.rsiz ? a : b, &y = a.rsiz >= b.rsiz ? b : a;
    precn_t r;
    reserve(r, x.rsiz + 1);
    uint64_t c = 0;
    for (size_t i = 0; i < x.rsiz; ++i) {
        const uint64_t yi = i < y.rsiz ? y.a[i] : 0;
        const uint64_t s = x.a[i] + yi;
        const uint64_t s2 = s + c;
        c = (s < yi) | (s2 < s);
        r.a[i] = s2;
    }
    r.a[x.rsiz] = c;
    r.rsiz = x.rsiz + 1;
    trim(r);
    return r;
}

precn_t add_u64(const precn_t &a, uint64_t b) { return a + precn_t(b); }

precn_t operator+(const precn_t &a) { return a; }

precn_t operator++(precn_t &a) {
    a = add_u64(a, 1);
    return a;
}

precn_t operator++(precn_t &a, int) {
    precn_t old(a);
    a = add_u64(a, 1);
    return old;
}

precn_t operator--(precn_t &a) {
    a = a - precn_t(1);
    return a;
}

precn_t operator--(precn_t &a, int) {
    precn_t old(a);
    a = a - precn_t(1);
    return old;
}

namespace {
// limb-wise a op b over max(rsiz) limbs (missing limbs read as 0), trimmed
template <class Op>
precn_t bitwise(const precn_t &a, const precn_t &b, Op op) {
    const size_t n = std::max(a.rsiz, b.rsiz);
    precn_t r;
    reserve(r, n);
    for (size_t i = 0; i < n; ++i) r.a[i] = op(i < a.rsiz ? a.a[i] : 0, i < b.rsiz ? b.a[i] : 0);
    r.rsiz = n;
    trim(r);
    return r;
}
}  // namespace

precn_t operator&(const precn_t &a, const precn_t &b) {
    return bitwise(a, b, [](uint64_t x, uint64_t y) { return x & y; });
}
precn_t operator|(const precn_t &a, const precn_t &b) {
    return bitwise(a, b, [](uint64_t x, uint64_t y) { return x | y; });
}
precn_t operator^(const precn_t &a, const precn_t &b) {
    return bitwise(a, b, [](uint64_t x, uint64_t y) { return x ^ y; });
}

precn_t operator-(const precn_t &a, const precn_t &b) {
    if (cmp(a, b) < 0) return precn_t();   // saturates at zero (sub.cpp:14)
    precn_t r;
    reserve(r, a.rsiz);
    uint64_t br = 0;
    for (size_t i = 0; i < a.rsiz; ++i) {
        const uint64_t yi = i < b.rsiz ? b.a[i] : 0;
        const uint64_t d = a.a[i] - yi - br;
        br = (a.a[i] < yi) || (a.a[i] - yi < br);
        r.a[i] = d;
    }
    r.rsiz = a.rsiz;
    trim(r);
    return r;
}

precn_t operator<<(const precn_t &a, size_t bits) {
    precn_t r;
    if (!a.rsiz) return r;
    const size_t q = bits / 64;
    const unsigned s = (unsigned)(bits % 64);
    reserve(r, a.rsiz + q + 1);
    std::memset(r.a, 0, (a.rsiz + q + 1) * sizeof(uint64_t));
    for (size_t i = 0; i < a.rsiz; ++i) {
        r.a[i + q] |= a.a[i] << s;
        if (s) r.a[i + q + 1] = a.a[i] >> (64 - s);
    }
    r.rsiz = a.rsiz + q + 1;
    trim(r);
    return r;
}

precn_t operator>>(const precn_t &a, size_t bits) {
    precn_t r;
    const size_t q = bits / 64;
    if (q >= a.rsiz) return r;
    const unsigned s = (unsigned)(bits % 64);
    const size_t n = a.rsiz - q;
    reserve(r, n);
    for (size_t i = 0; i < n; ++i)
        r.a[i] = (a.a[i + q] >> s) | (s && i + q + 1 < a.rsiz ? a.a[i + q + 1] << (64 - s) : 0);
    r.rsiz = n;
    trim(r);
    return r;
}
