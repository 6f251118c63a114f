// test_precn.cpp -- precn_t over libpgpu.so (include/precn.hpp) against the
// oracle (oracle/oracle.h), in the manner of the reference's test/test_prec.cpp:
// the same entry points, the same shapes and known answers.  Needs a GPU.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "oracle.h"
#include "precn.hpp"

static int failures = 0, checks = 0;

static void expect(bool ok, const char *what, size_t a = 0, size_t b = 0) {
    ++checks;
    if (!ok) {
        ++failures;
        std::printf("FAIL %s (%zu, %zu)\n", what, a, b);
    }
}

static precn_t limbs(const std::vector<uint64_t> &v) {
    precn_t r;
    r.a = (uint64_t *)realloc(r.a, std::max<size_t>(v.size(), 1) * 8);
    r.asiz = std::max<size_t>(v.size(), 1);
    if (!v.empty()) std::memcpy(r.a, v.data(), v.size() * 8);
    r.rsiz = v.size();
    while (r.rsiz && !r.a[r.rsiz - 1]) --r.rsiz;
    return r;
}

static precn_t fill(size_t n, uint64_t seed, int dist = 0) {
    std::vector<uint64_t> v(n);
    if (n) or_fill(v.data(), n, seed, dist);
    return limbs(v);
}

static bool same(const precn_t &x, const std::vector<uint64_t> &v, size_t n) {
    if (x.rsiz != n) return false;
    return n == 0 || std::memcmp(x.a, v.data(), n * 8) == 0;
}

static std::vector<uint64_t> oracle_mul(const precn_t &a, const precn_t &b, size_t *n, int sqr = 0) {
    std::vector<uint64_t> r(a.rsiz + b.rsiz + 1);
    *n = or_mul_into(r.data(), r.size(), a.a, a.rsiz, b.a, b.rsiz, sqr);
    return r;
}

static void test_known_answers() {
    // test_prec.cpp:184-195: (2^32 - 1)^2 through every entry
    precn_t x((uint64_t)0xFFFFFFFFu);
    const uint64_t want = 0xFFFFFFFE00000001ull;
    for (const precn_t &r : {x * x, mul_ntt(x, x), mul_fft(x, x), precn_sqr(x)})
        expect(r.rsiz == 1 && r.a[0] == want, "(2^32-1)^2");
    precn_t z;
    expect((z * x).rsiz == 0 && (x * z).rsiz == 0, "zero operand");
    expect((std::string)precn_t(0) == "0" && (std::string)precn_t(1234567890123ull) == "1234567890123",
           "decimal text");
}

static void test_all_mul_eq() {
    // test_prec.cpp:429-478 shapes: base x {1, 1.25, 1.33, 1.5, 2, 3} base
    for (size_t base : {32, 128, 512, 2048}) {
        const size_t mults[6][2] = {{1, 1}, {5, 4}, {4, 3}, {3, 2}, {2, 1}, {3, 1}};
        for (auto &m : mults) {
            const size_t bn = base * m[0] / m[1];
            precn_t a = fill(base, 3000 + base), b = fill(bn, 4009 + bn);
            size_t n = 0;
            std::vector<uint64_t> e = oracle_mul(a, b, &n);
            expect(same(a * b, e, n), "operator*", base, bn);
            expect(same(mul_ntt(a, b), e, n), "mul_ntt", base, bn);
            expect(same(mul_fft(a, b), e, n), "mul_fft", base, bn);
            precn_t r;
            mul_into(r, a, b);
            expect(same(r, e, n), "mul_into", base, bn);
        }
    }
}

static void test_dispatcher_shapes() {
    // every branch of mul_into (mul_basic.cpp:238-276) and both sqr routes
    const size_t shapes[][2] = {{1, 1}, {1, 5000}, {24, 9000}, {25, 400}, {192, 192}, {193, 200},
                                {300, 1000}, {1000, 3100}, {6144, 6144}, {6145, 6145}, {7000, 20000},
                                {40000, 40000}};
    for (auto &s : shapes)
        for (int dist = 0; dist < 3; ++dist) {
            precn_t a = fill(s[0], 11 + s[0], dist), b = fill(s[1], 29 + s[1], dist);
            size_t n = 0;
            std::vector<uint64_t> e = oracle_mul(a, b, &n);
            expect(same(a * b, e, n), "dispatch", s[0], s[1]);
            e = oracle_mul(a, a, &n, 1);
            expect(same(precn_sqr(a), e, n), "precn_sqr", s[0], s[0]);
            precn_t c = a;
            mul_into(c, c, c);   // aliased square
            expect(same(c, e, n), "aliased square", s[0], s[0]);
        }
}

static void test_aliasing_and_lifecycle() {
    precn_t a = fill(3000, 7), b = fill(1200, 8);
    size_t n = 0;
    std::vector<uint64_t> e = oracle_mul(a, b, &n);
    precn_t r = a;
    mul_into(r, r, b);
    expect(same(r, e, n), "r aliases a");
    r = b;
    mul_into(r, a, r);
    expect(same(r, e, n), "r aliases b");
    precn_t m(std::move(r));
    expect(r.a == nullptr && same(m, e, n), "move leaves nullptr (init_basic.cpp:60-65)");
    precn_t c(m);
    expect(c == m && c.a != m.a, "copy");
    c = std::move(m);
    expect(m.a == nullptr && same(c, e, n), "move assignment");
}

static void test_division() {
    // test_prec.cpp:243-255
    expect(precn_reciprocal_newton(precn_t(3), 8) == precn_t(85), "recip(3, 8)");
    expect(precn_reciprocal_newton(precn_t(5), 3) == precn_t(1), "recip(5, 3)");
    expect(precn_reciprocal_newton(precn_t(9), 3).rsiz == 0, "recip(9, 3)");
    expect(precn_reciprocal_newton(precn_t(), 128).rsiz == 0, "recip(0, 128)");
    expect(div_mulinv(precn_t(7), precn_t(3)) == precn_t(2), "7 / 3");
    expect(div_mulinv(precn_t(3), precn_t(7)).rsiz == 0, "3 / 7");
    expect(div_mulinv(precn_t(9), precn_t()).rsiz == 0, "9 / 0");
    // test_prec.cpp:628-653: dividend = divisor * quotient + (divisor - 1)
    for (size_t k : {1, 16, 300, 4096, 40000}) {
        precn_t d = fill(k, 12000 + k), q = fill(k, 13000 + k);
        precn_t p = d * q + (d - precn_t(1));
        expect(div_mulinv(p, d) == q, "div_mulinv identity", k);
        expect(mod_mulinv(p, d) == d - precn_t(1), "mod_mulinv identity", k);
        expect(p / d == q && p % d == d - precn_t(1), "operator/ %", k);
    }
    precn_t a = fill(9000, 77), b = fill(3333, 78);
    std::vector<uint64_t> q(9000), r(9000);
    size_t ql = 0, rl = 0;
    or_divmod(q.data(), q.size(), &ql, r.data(), r.size(), &rl, a.a, a.rsiz, b.a, b.rsiz);
    expect(same(a / b, q, ql) && same(a % b, r, rl), "division vs oracle");
}

static void test_sqrt_and_text() {
    for (size_t n : {1, 4, 5, 33, 2000}) {
        precn_t a = fill(n, 500 + n);
        std::vector<uint64_t> x(n + 2);
        const size_t xl = or_sqrt(x.data(), x.size(), a.a, a.rsiz);
        expect(same(precn_sqrt(a), x, xl), "precn_sqrt", n);
    }
    precn_t t(1);
    for (int i = 0; i < 1000; ++i) t = t * precn_t(10);
    expect((std::string)t == "1" + std::string(1000, '0'), "10^1000 as text");
    std::vector<uint32_t> d(1100);
    size_t nd = 0;
    precn_base_convert(t, 10, d.data(), nd);
    expect(nd == 1001 && d[1000] == 1 && d[0] == 0, "precn_base_convert base 10");
    precn_base_convert(precn_t(0xF0F0), 16, d.data(), nd);
    expect(nd == 4 && d[0] == 0 && d[1] == 0xF && d[3] == 0xF, "precn_base_convert base 16");
}

static void test_linear_and_scalar() {
    // sub.cpp:14: a - b saturates to 0 when a < b
    precn_t a = fill(300, 901), b = fill(301, 902);
    expect((a - b).rsiz == 0 && (precn_t(3) - precn_t(7)).rsiz == 0, "a - b saturates");
    expect(b - a + a == b, "b - a + a");
    precn_t z;
    expect(--z == precn_t() && z.rsiz == 0, "--0 stays 0");
    precn_t c(41);
    expect(c++ == precn_t(41) && c == precn_t(42) && ++c == precn_t(43), "increments");
    expect(c-- == precn_t(43) && --c == precn_t(41), "decrements");
    precn_t ff = limbs({~0ull, ~0ull});
    expect(++ff == limbs({0, 0, 1}), "++ carries");
    expect((limbs({0xF0F0, 0xFF}) & limbs({0xFF00})) == precn_t(0xF000), "and");
    expect((limbs({0xF0F0}) | limbs({0x0F0F, 1})) == limbs({0xFFFF, 1}), "or");
    expect((limbs({5, 7}) ^ limbs({5, 7})).rsiz == 0, "xor trims");
    expect(+a == a, "unary plus");
    // scalar division by one word (div_basic.cpp:135-198); by 0 -> 0
    std::vector<uint64_t> q(300), r(2);
    size_t ql = 0, rl = 0;
    const uint64_t dw = 0xFEDCBA9876543211ull;
    or_divmod(q.data(), q.size(), &ql, r.data(), r.size(), &rl, a.a, a.rsiz, &dw, 1);
    expect(same(div_u64(a, dw), q, ql) && same(mod_u64(a, dw), r, rl), "div_u64 / mod_u64");
    expect(same(a / dw, q, ql) && same(a % dw, r, rl), "operator/ % by a word");
    expect(div_u64(a, 0).rsiz == 0 && mod_u64(a, 0).rsiz == 0, "by zero word");
    expect(div_u32(limbs({~0ull}), 2) == limbs({0x7FFFFFFFFFFFFFFFull}), "div_u32");
    expect(mul_u64(a, 3) == a * 3 && a * 3 == 3 * a && a + 5 == 5 + a, "scalar templates");
    expect(precn_t(100) / precn_t(7) == precn_t(14) && 100 / precn_t(7) == precn_t(14) &&
               100 % precn_t(7) == precn_t(2), "scalar left operands");
    // exact division helpers (mul_toom.cpp:38-70)
    precn_t t3 = mul_u64(a, 3);
    expect(precn_divexact_3(t3) == a && precn_divexact_2(a + a) == a, "divexact");
}

static void test_division_family() {
    // div_basic.cpp:318-407: every entry gives floor(a / b) and a mod b; a
    // zero operand gives 0; q and r may alias the operands
    for (size_t an : {1, 2, 40, 300, 5000})
        for (size_t bn : {1, 2, 3, 64, 200, 4000}) {
            precn_t a = fill(an, 7000 + an), b = fill(bn, 8000 + bn);
            std::vector<uint64_t> q(an + 2), r(an + bn + 2);
            size_t ql = 0, rl = 0;
            or_divmod(q.data(), q.size(), &ql, r.data(), r.size(), &rl, a.a, a.rsiz, b.a, b.rsiz);
            precn_t x, y;
            div_into(x, a, b);
            mod_into(y, a, b);
            expect(same(x, q, ql) && same(y, r, rl), "div_into / mod_into", an, bn);
            divmod_into(x, y, a, b);
            expect(same(x, q, ql) && same(y, r, rl), "divmod_into", an, bn);
            divmod_schoolbook_into(x, y, a, b);
            expect(same(x, q, ql) && same(y, r, rl), "divmod_schoolbook_into", an, bn);
            expect(same(div_schoolbook(a, b), q, ql) && same(mod_schoolbook(a, b), r, rl), "schoolbook", an, bn);
            precn_t aa = a;
            div_into(aa, aa, b);
            expect(same(aa, q, ql), "div_into aliased", an, bn);
            aa = a;
            mod_into(aa, aa, b);
            expect(same(aa, r, rl), "mod_into aliased", an, bn);
            precn_t dq, dr;
            const bool dc = div_dc_into(dq, dr, a, b);
            const bool want_dc = bn >= 64 && !(a < b) && an <= 2 * bn;   // div_dc.cpp:93-94
            expect(dc == want_dc && (!dc || (same(dq, q, ql) && same(dr, r, rl))), "div_dc_into", an, bn);
        }
    precn_t a = fill(50, 1);
    precn_t q1, r1;
    divmod_into(q1, r1, a, precn_t());
    expect(q1.rsiz == 0 && r1.rsiz == 0 && (a / precn_t()).rsiz == 0 && (a % precn_t()).rsiz == 0, "by zero");
    precn_t g = gcd(fill(40, 5) * fill(30, 6), fill(40, 5) * fill(20, 9));
    expect(g % fill(40, 5) == precn_t() && gcd(precn_t(12), precn_t(18)) == precn_t(6) &&
               gcd(precn_t(), precn_t(5)) == precn_t(5), "gcd");
}

static void test_parse_and_bases() {
    // init_basic.cpp:66-92: digits only, other characters skipped
    expect(precn_t(std::string("4294967296")) == limbs({1ull << 32}), "parse 2^32");
    expect(precn_t(std::string("12a34")) == precn_t(1234), "parse skips letters");
    expect(precn_t(std::string("")).rsiz == 0 && precn_t(std::string("000")).rsiz == 0 &&
               precn_t(std::string("-+x")).rsiz == 0, "parse empty / zeros");
    for (size_t n : {1, 5, 100, 3000, 40000}) {
        precn_t a = fill(n, 31 + n);
        const std::string s = (std::string)a;
        expect(precn_t(s) == a, "decimal round trip", n);
        expect(precn_t("00" + s) == a, "leading zeros", n);
    }
    // other bases (base_convert.cpp:107-172), least significant digit first
    for (uint32_t base : {3u, 7u, 10u, 16u, 1000u, 4294967291u}) {
        precn_t a = fill(60, 90 + base);
        size_t nd = 0;
        precn_base_convert(a, base, nullptr, nd);
        std::vector<uint32_t> d(nd + 1);
        size_t nd2 = 0;
        precn_base_convert(a, base, d.data(), nd2);
        precn_t back;
        for (size_t i = nd2; i-- > 0;) back = mul_u64(back, base) + precn_t(d[i]);
        expect(nd == nd2 && nd && d[nd - 1] != 0 && back == a, "base round trip", base);
    }
}

static void test_signed() {
    // precz.cpp semantics over every sign: truncating quotient, remainder
    // with the dividend's sign (as C++'s int64 / and %), zero never negative
    for (long x = -100; x <= 100; ++x)
        for (long y = -100; y <= 100; ++y) {
            const precz_t a(x), b(y);
            bool ok = (long)(a + b) == x + y && (long)(a - b) == x - y && (long)(a * b) == x * y;
            ok = ok && ((a < b) == (x < y)) && ((a == b) == (x == y)) && ((a >= b) == (x >= y));
            if (y) ok = ok && (long)(a / b) == x / y && (long)(a % b) == x % y;
            else ok = ok && (a / b).is_zero() && (a % b).is_zero();
            ok = ok && !(a * precz_t(0)).is_negative() && (a - a).is_zero() && !(a - a).is_negative();
            expect(ok, "precz arithmetic", (size_t)(x + 100), (size_t)(y + 100));
        }
    expect((std::string)precz_t(std::string("  -123")) == "-123" && (std::string)precz_t(std::string("-0")) == "0" &&
               (std::string)precz_t(-45) == "-45" && (long)precz_t(std::string("77")) == 77, "precz text");
    precz_t big = precz_t(std::string("-123456789012345678901234567890"));
    expect((std::string)(big * big) == "15241578753238836750495351562536198787501905199875019052100", "precz big");
    expect(abs(big) == -big && (precz_t(-8) << 2) == precz_t(-32) && (precz_t(-33) >> 2) == precz_t(-8) &&
               gcd(precz_t(-12), precz_t(18)) == precz_t(6) && precz_sqrt(precz_t(-4)).is_zero() &&
               precz_sqrt(precz_t(17)) == precz_t(4), "precz helpers");
    precz_t i(-1);
    expect((i++ == precz_t(-1)) && i.is_zero() && (--i == precz_t(-1)), "precz increments");
}

int main() {
    precn_gpu_init(0);
    test_linear_and_scalar();
    test_division_family();
    test_parse_and_bases();
    test_signed();
    test_known_answers();
    test_all_mul_eq();
    test_dispatcher_shapes();
    test_aliasing_and_lifecycle();
    test_division();
    test_sqrt_and_text();
    std::printf("%d checks, %d failures\n", checks, failures);
    if (!failures) std::printf("all passed\n");
    return failures ? 1 : 0;
}
