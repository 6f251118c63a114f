// precn.cpp -- precn_t (include/precn.hpp) over the libpgpu.so C ABI.
//
// Lifecycle follows src/init_basic.cpp:47-135 (malloc / realloc / free,
// moved-from objects hold a == nullptr).  Products, divisions, square roots
// and base-10 conversion go through pg_mul / pg_div / pg_recip / pg_sqrt /
// pg_to_decimal; the dispatcher rules of mul_into (mul_basic.cpp:238-276)
// are applied by the library (PG_PATH_AUTO): the reference's structure, with
// the NTT where preclib's CPU thresholds pick mul_fft (pgpu.h).  Additions, subtractions, shifts and comparisons are
// the linear host steps of src/add.cpp, src/sub.cpp and
// src/compare_shift.cpp.
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

precn_t mul_fft(const precn_t &a, const precn_t &b) {
    precn_t r;
    gpu_mul(r, a, b, PG_PATH_FFT);
    return r;
}

precn_t mul_fft_round(const precn_t &a, const precn_t &b, double *max_round) {
    precn_t r;
    gpu_mul(r, a, b, PG_PATH_FFT, max_round);
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

// ---- division, square root, conversion -------------------------------------------

static void gpu_div(const precn_t &a, const precn_t &b, precn_t *q, precn_t *r) {
    ensure_init();
    std::vector<uint64_t> qb(a.rsiz + 1), rb(b.rsiz + a.rsiz + 1);
    size_t ql = 0, rl = 0;
    check(pg_div(q ? qb.data() : nullptr, qb.size(), &ql, r ? rb.data() : nullptr, rb.size(), &rl, a.a, a.rsiz,
                 b.a, b.rsiz));
    if (q) *q = from_limbs(qb.data(), ql);
    if (r) *r = from_limbs(rb.data(), rl);
}

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

precn_t operator/(const precn_t &a, const precn_t &b) { return div_mulinv(a, b); }
precn_t operator%(const precn_t &a, const precn_t &b) { return mod_mulinv(a, b); }

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
    if (base != 10) throw precn_gpu_error(PG_EINVAL, "precn_base_convert: bases other than 10 and 2^k");
    const std::string s = (std::string)a;
    out_siz = s.size();
    if (out)
        for (size_t i = 0; i < out_siz; ++i) out[i] = (uint32_t)(s[out_siz - 1 - i] - '0');
}

// ---- linear steps ------------------------------------------------------------------

int precn_cmp(const precn_t &a, const precn_t &b) {
    if (a.rsiz != b.rsiz) return a.rsiz < b.rsiz ? -1 : 1;
    for (size_t i = a.rsiz; i-- > 0;)
        if (a.a[i] != b.a[i]) return a.a[i] < b.a[i] ? -1 : 1;
    return 0;
}

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

precn_t operator-(const precn_t &a, const precn_t &b) {
    if (precn_cmp(a, b) < 0) std::abort();   // unsigned: a >= b required
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
