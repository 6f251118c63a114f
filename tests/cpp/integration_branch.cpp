// The dispatcher branch INTEGRATION.md adds to preclib's mul_into (option 2),
// compiled against this repository's headers (precn_t has preclib's layout,
// prec.hpp:49-84) and run on the GPU: products, an aliased output, a square
// and the zero operand, each checked against a schoolbook product computed
// here on the CPU.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "pgpu.h"
#include "precn.hpp"

// Capacity of r for n limbs (the contract of mul_basic.cpp:14-18's helper).
static void reserve_limbs(precn_t &r, size_t n) {
    if (r.asiz >= n) return;
    r.a = (uint64_t *)realloc(r.a, n * sizeof(uint64_t));
    r.asiz = n;
}

// ---- the branch, as printed in INTEGRATION.md -----------------------------
// r = a * b on the GPU.  Returns false when the library declines; the status and
// pg_last_error() say why.
static bool mul_gpu(precn_t &r, const precn_t &a, const precn_t &b, int path) {
    size_t need = a.rsiz + b.rsiz;
    reserve_limbs(r, need);                   // r may alias a or b
    size_t len = 0;
    int st = pg_mul(r.a, r.asiz, &len, a.a, a.rsiz, b.a, b.rsiz, path, nullptr);
    if (st != PG_OK) return false;
    r.rsiz = len;
    return true;
}
// ---------------------------------------------------------------------------

static precn_t limbs(size_t n, uint64_t seed) {
    precn_t x;
    reserve_limbs(x, n ? n : 1);
    uint64_t s = seed;
    for (size_t i = 0; i < n; ++i) {   // splitmix64
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        x.a[i] = z ^ (z >> 31);
    }
    if (n) x.a[n - 1] |= 1;   // nonzero top limb
    x.rsiz = n;
    return x;
}

static std::vector<uint64_t> schoolbook(const precn_t &a, const precn_t &b) {
    std::vector<uint64_t> r(a.rsiz + b.rsiz, 0);
    for (size_t i = 0; i < a.rsiz; ++i) {
        unsigned __int128 carry = 0;
        for (size_t j = 0; j < b.rsiz; ++j) {
            const unsigned __int128 t = (unsigned __int128)a.a[i] * b.a[j] + r[i + j] + carry;
            r[i + j] = (uint64_t)t;
            carry = t >> 64;
        }
        r[i + b.rsiz] = (uint64_t)carry;
    }
    while (!r.empty() && r.back() == 0) r.pop_back();
    return r;
}

static int check(const char *what, const precn_t &r, const std::vector<uint64_t> &want) {
    if (r.rsiz != want.size() || (r.rsiz && std::memcmp(r.a, want.data(), r.rsiz * 8) != 0)) {
        std::printf("FAIL %s: %zu limbs, want %zu\n", what, r.rsiz, want.size());
        return 1;
    }
    return 0;
}

int main() {
    if (pg_init(0, 0) != PG_OK) {
        std::printf("FAIL pg_init: %s\n", pg_last_error());
        return 1;
    }
    int bad = 0;
    const size_t shapes[][2] = {{1, 1}, {20, 700}, {193, 193}, {300, 500}, {3000, 3000}, {1500, 7000}};
    for (const auto &sh : shapes) {
        precn_t a = limbs(sh[0], 11 + sh[0]), b = limbs(sh[1], 23 + sh[1]), r;
        const std::vector<uint64_t> want = schoolbook(a, b);
        if (!mul_gpu(r, a, b, PG_PATH_AUTO)) {
            std::printf("FAIL declined %zu x %zu: %s\n", sh[0], sh[1], pg_last_error());
            return 1;
        }
        char what[64];
        std::snprintf(what, sizeof what, "%zu x %zu", sh[0], sh[1]);
        bad |= check(what, r, want);
        // the output aliasing the first operand
        if (!mul_gpu(a, a, b, PG_PATH_AUTO)) return 1;
        bad |= check("aliased output", a, want);
    }
    precn_t s = limbs(4000, 5);
    const std::vector<uint64_t> sq = schoolbook(s, s);
    precn_t r2;
    if (!mul_gpu(r2, s, s, PG_PATH_AUTO)) return 1;
    bad |= check("square", r2, sq);
    precn_t z;   // zero: the empty product (mul_basic.cpp:239-242)
    if (!mul_gpu(r2, z, s, PG_PATH_AUTO)) return 1;
    bad |= check("zero operand", r2, {});
    std::printf(bad ? "integration branch FAILED\n" : "integration branch ok\n");
    return bad;
}
