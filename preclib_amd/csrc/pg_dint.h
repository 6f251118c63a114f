// pg_dint.h -- device-resident integers: the linear steps between products
// and Newton division (src/div_newton.cpp) on HBM operands.
#pragma once
#include <vector>
#include "pg_runtime.h"

namespace pg {

inline size_t limbs_for_bits(size_t bits) { return (bits + 63) / 64; }

// Pool-owned device limbs (move-only).
class DBuf {
public:
    DBuf() = default;
    explicit DBuf(size_t limbs) { reset(limbs); }
    ~DBuf() { release(); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : d(o.d), cap(o.cap) { o.d = nullptr; o.cap = 0; }
    DBuf &operator=(DBuf &&o) noexcept {
        if (this != &o) { release(); d = o.d; cap = o.cap; o.d = nullptr; o.cap = 0; }
        return *this;
    }
    bool reset(size_t limbs);
    void release();
    uint64_t *d = nullptr;
    size_t cap = 0;
};

// A linear node writing out[0, out_len) (out_len = m unless changed).
CarryNode lin_node(uint64_t *out, size_t m);
// One carry launch over n linear nodes.  Outputs must not overlap inputs.
int lin_run(const CarryNode *nodes, size_t n, const char *stage = "linear");
inline int lin_one(const CarryNode &n, const char *stage = "linear") { return lin_run(&n, 1, stage); }

// r[0, an + bn) = a * b through the size dispatcher (upper-bound lengths).
int dmul(uint64_t *r, const uint64_t *a, size_t an, const uint64_t *b, size_t bn);
// r[0, n) = value (value limbs, zero padded).
int dset(uint64_t *r, size_t n, const uint64_t *value, size_t vn);
int dcopy(uint64_t *r, const uint64_t *a, size_t n);

// floor(2^e / d) for a nonzero host integer d of a few limbs.
std::vector<uint64_t> pow2_div_small(size_t e, const std::vector<uint64_t> &d);

// Waiting reads.
int top_len(const uint64_t *a, size_t n, size_t *len);   // trimmed length
int read_words(const uint64_t *src, size_t n, uint64_t *host);
int bit_length(const uint64_t *a, size_t trimmed, size_t *bits);

// newton64_reciprocal_approx (div_newton.cpp:36-68): x <= floor(2^bits / b)
// within a few units; b trimmed (bn limbs, nb bits).  x gets xl limbs.
// The seed floor(2^(keep + known) / top_ceil(b)) is the reference's, with
// known = NEWTON_SEED_BITS instead of 64 (the reference is limited to one
// machine word there): a host long division of about 0.3 ms that saves five
// precision doublings.
constexpr size_t NEWTON_SEED_BITS = 2048;
int recip_approx(DBuf &x, size_t &xl, const uint64_t *b, size_t bn, size_t nb, size_t bits);
// precn_reciprocal_newton (div_newton.cpp:70-87): exact floor(2^bits / b).
int recip_exact(DBuf &x, size_t &xl, const uint64_t *b, size_t bn, size_t bits);
// From an estimate x (xl limbs) within a few units of floor(2^bits / b), the
// exact floor: r = 2^bits - b x, then the unit corrections of
// div_newton.cpp:78-85.
int recip_finish(DBuf &x, size_t &xl, const uint64_t *b, size_t bn, size_t bits);
// div_mulinv_impl (div_newton.cpp:89-129): q = floor(a / b), r = a mod b.
// a, b trimmed; q gets ql trimmed limbs, r (if non-null) rl trimmed limbs.
int div_mulinv(DBuf &q, size_t &ql, DBuf *r, size_t *rl, const uint64_t *a, size_t an, const uint64_t *b,
               size_t bn);
// Test hook: what 0 adds `amount` units to the next division's first
// quotient estimate; what 1 multiplies the next reciprocal's seed by
// `amount` (an estimate too large, which takes the overshoot branch).
int debug_perturb(int what, int64_t amount);

}  // namespace pg
