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

int main() {
    precn_gpu_init(0);
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
