// pg_pi.hip -- Chudnovsky binary splitting, integer square root, decimal
// conversion and the pi program (pi/pi_chudnovsky.cpp) on device operands.
//
// Binary splitting (pi_chudnovsky.cpp:117-148).  P, Q and T of a term range
// do not depend on where the range is split, so the tree is built level by
// level over aligned power-of-two blocks instead of the reference's
// recursion.  Levels 1-5 run one thread per node (operands of at most a few
// dozen limbs, fixed slots); above that every level is one batched product
// launch plus one linear carry launch for T.  Sign of T(a, b) is (-1)^a:
// consecutive terms shrink by about 10^14, so the first term dominates.
// Aligned blocks therefore have T >= 0 from level 1 on, the only
// subtraction is in the level-1 merge, and every later T is a sum.
//
// Square root: Newton's iteration for 1/sqrt at doubling precision (no
// divisions), a final product and the unit corrections of src/sqrt.cpp:48-58.
// Decimal conversion: divide and conquer by W_i = (10^19)^(2^i) as
// src/base_convert.cpp:56-105, with one exact reciprocal per level shared by
// every node and a one-step quotient correction, then one thread per block
// of 32 chunks of 19 digits.
#include "pg_pi.h"
#include <algorithm>
#include <chrono>
#include <cstring>
#include "pg_batch.h"
#include "pgpu.h"

namespace pg {

constexpr uint64_t CHUD_C3_OVER_24 = 10939058860032000ull;   // pi_chudnovsky.cpp:11
constexpr uint64_t CHUD_A = 13591409u, CHUD_B = 545140134u;   // 12-13
constexpr int CHUD_THREAD_LEVELS = 2;                         // per-thread levels 1..2
constexpr int CHUD_DEVICE_LEVELS = 8;                         // one wave per merge: levels 3..8
constexpr uint64_t TEN19 = 10000000000000000000ull;
constexpr int DEC_LEAF = 32;                                  // chunks per conversion thread
constexpr size_t DEC_WAVE_MAX_LIMBS = 256;                    // split levels with W_i of at most this many limbs: one wave per division

// ---- per-thread integers -----------------------------------------------------------

__device__ __forceinline__ void mac(uint64_t a, uint64_t b, uint64_t add, uint64_t &c, uint64_t &out) {
    uint64_t lo = a * b, hi = __umul64hi(a, b);
    lo += c;
    hi += lo < c;
    lo += add;
    hi += lo < add;
    out = lo;
    c = hi;
}

// out[0, an + bn) = a * b
__device__ void t_mul(uint64_t *out, const uint64_t *a, uint32_t an, const uint64_t *b, uint32_t bn) {
    for (uint32_t k = 0; k < an + bn; ++k) out[k] = 0;
    for (uint32_t i = 0; i < an; ++i) {
        uint64_t c = 0;
        for (uint32_t j = 0; j < bn; ++j) mac(a[i], b[j], out[i + j], c, out[i + j]);
        out[i + bn] = c;
    }
}

// out[0, on) += a * b or -= a * b (the result must fit in on limbs and stay >= 0)
template <bool SUB>
__device__ void t_addmul(uint64_t *out, uint32_t on, const uint64_t *a, uint32_t an, const uint64_t *b,
                         uint32_t bn) {
    for (uint32_t i = 0; i < an; ++i) {
        uint64_t c = 0;
        uint32_t k = i;
        for (uint32_t j = 0; j < bn; ++j, ++k) {
            uint64_t p;
            mac(a[i], b[j], 0, c, p);
            if (SUB) {
                const uint64_t o = out[k];
                out[k] = o - p;
                c += o < p;
            } else {
                const uint64_t s = out[k] + p;
                c += s < p;
                out[k] = s;
            }
        }
        for (; c && k < on; ++k) {
            const uint64_t o = out[k];
            if (SUB) {
                out[k] = o - c;
                c = o < c;
            } else {
                out[k] = o + c;
                c = out[k] < c;
            }
        }
    }
}

__device__ __forceinline__ uint32_t t_len(const uint64_t *a, uint32_t n) {
    while (n && !a[n - 1]) --n;
    return n;
}

// Leaf k (pi_chudnovsky.cpp:118-135): p, q, |t| as two limbs each.
__device__ void chud_leaf(uint64_t k, uint64_t p[2], uint64_t q[2], uint64_t t[2]) {
    if (k == 0) {
        p[0] = 1; p[1] = 0;
        q[0] = 1; q[1] = 0;
        t[0] = CHUD_A; t[1] = 0;
        return;
    }
    const uint64_t x = (6 * k - 5) * (2 * k - 1);   // < 2^43 for k < 2^20
    p[0] = x * (6 * k - 1);
    p[1] = __umul64hi(x, 6 * k - 1);
    const uint64_t k3 = k * k * k;                  // < 2^60
    q[0] = k3 * CHUD_C3_OVER_24;
    q[1] = __umul64hi(k3, CHUD_C3_OVER_24);
    const uint64_t m = CHUD_B * k + CHUD_A;         // < 2^50
    uint64_t c = 0;
    mac(p[0], m, 0, c, t[0]);
    t[1] = p[1] * m + c;                            // |t| < 2^115
}

struct ChudLevel {
    uint64_t *P, *Q, *T;     // node i at + i * slot
    uint32_t *lp, *lq, *lt;
    uint32_t slot;
    uint32_t count;
};

// Level 1 from the leaves: node i merges terms 2i (even, t >= 0) and 2i + 1
// (odd, t <= 0): Q = q_a q_b, T = |t_a| q_b - |t_b| p_a, P = p_a p_b.
__global__ void __launch_bounds__(256) k_chud_base(uint64_t nterms, ChudLevel out) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= out.count) return;
    uint64_t pa[2], qa[2], ta[2];
    chud_leaf(2 * i, pa, qa, ta);
    uint64_t *P = out.P + i * out.slot, *Q = out.Q + i * out.slot, *T = out.T + i * out.slot;
    if (2 * i + 1 >= nterms) {
        for (int k = 0; k < 2; ++k) { P[k] = pa[k]; Q[k] = qa[k]; T[k] = ta[k]; }
        out.lp[i] = t_len(P, 2); out.lq[i] = t_len(Q, 2); out.lt[i] = t_len(T, 2);
        return;
    }
    uint64_t pb[2], qb[2], tb[2];
    chud_leaf(2 * i + 1, pb, qb, tb);
    t_mul(Q, qa, 2, qb, 2);
    t_mul(P, pa, 2, pb, 2);
    t_mul(T, ta, 2, qb, 2);
    t_addmul<true>(T, 4, tb, 2, pa, 2);
    out.lp[i] = t_len(P, 4); out.lq[i] = t_len(Q, 4); out.lt[i] = t_len(T, 4);
}

// Level j -> j + 1 (both children have T >= 0): Q = lq rq, T = lt rq + rt lp,
// P = lp rp; an unpaired last node moves up unchanged.
__global__ void __launch_bounds__(256) k_chud_merge(ChudLevel in, ChudLevel out) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= out.count) return;
    const uint64_t l = 2 * i, r = l + 1;
    uint64_t *P = out.P + i * out.slot, *Q = out.Q + i * out.slot, *T = out.T + i * out.slot;
    const uint64_t *lP = in.P + l * in.slot, *lQ = in.Q + l * in.slot, *lT = in.T + l * in.slot;
    const uint32_t llp = in.lp[l], llq = in.lq[l], llt = in.lt[l];
    if (r >= in.count) {
        for (uint32_t k = 0; k < llp; ++k) P[k] = lP[k];
        for (uint32_t k = 0; k < llq; ++k) Q[k] = lQ[k];
        for (uint32_t k = 0; k < llt; ++k) T[k] = lT[k];
        out.lp[i] = llp; out.lq[i] = llq; out.lt[i] = llt;
        return;
    }
    const uint64_t *rP = in.P + r * in.slot, *rQ = in.Q + r * in.slot, *rT = in.T + r * in.slot;
    const uint32_t rlp = in.lp[r], rlq = in.lq[r], rlt = in.lt[r];
    t_mul(Q, lQ, llq, rQ, rlq);
    t_mul(P, lP, llp, rP, rlp);
    const uint32_t tn = (llt + rlq > rlt + llp ? llt + rlq : rlt + llp) + 1;
    t_mul(T, lT, llt, rQ, rlq);
    for (uint32_t k = llt + rlq; k < tn; ++k) T[k] = 0;
    t_addmul<false>(T, tn, rT, rlt, lP, llp);
    out.lp[i] = t_len(P, llp + rlp); out.lq[i] = t_len(Q, llq + rlq); out.lt[i] = t_len(T, tn);
}

// ---- one wave per merge (levels CHUD_THREAD_LEVELS + 1 .. CHUD_DEVICE_LEVELS) ----------
//
// A level-j node has at most 3 * 2^j limbs (slot): one thread per merge took
// 350 / 579 / 1911 us for levels 3 / 4 / 5 at 10^7 digits (level 5 on 88
// workgroups), and host-planned batches of tens of thousands of small products
// cost more on the host than on the GPU.  Here a wave forms each product 64
// limbs at a time from column sums (wave_product): levels 3..8 take 159, 159,
// 131, 208, 345 and 645 us.

// Value of lane (this lane - d) (d = 1, 2); lanes below d read 0.
__device__ __forceinline__ uint64_t lane_down(uint64_t v, uint32_t lane, uint32_t d) {
    const int src = (int)(lane >= d ? lane - d : 0) << 2;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)(uint32_t)(v >> 32));
    return lane >= d ? ((uint64_t)hi << 32 | lo) : 0;
}

__device__ __forceinline__ uint64_t lane_at(uint64_t v, int l) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
    return (uint64_t)hi << 32 | lo;
}

// Adds the column sum of column c = c0 + lane of a * b (a global, read with
// uniform indices; b in LDS; na, nb >= 1) into (s0, s1, s2): the 128-bit
// products a_i b_(c-i).
__device__ __forceinline__ void col_sum(uint64_t &s0, uint64_t &s1, uint64_t &s2, const uint64_t *a, uint32_t na,
                                        const uint64_t *b, uint32_t nb, uint32_t c0, uint32_t lane) {
    const uint32_t c = c0 + lane;
    // the i some lane of the chunk uses: c - (nb - 1) <= i <= c
    const uint32_t ilo = c0 + 1 > nb ? c0 + 1 - nb : 0;
    const uint32_t ihi = min(na - 1, c0 + 63);
#pragma unroll 4
    for (uint32_t i = ilo; i <= ihi; ++i) {
        const uint32_t j = c - i;   // (wraps when i > c: then j >= nb)
        const uint64_t x = a[i];
        if (i <= c && j < nb) {
            const uint64_t y = b[j];
            const uint64_t lo = x * y, hi = __umul64hi(x, y);
            s0 += lo;
            const uint64_t k0 = s0 < lo;
            s1 += hi;
            const uint64_t k1 = s1 < hi;
            s1 += k0;
            s2 += k1 + (s1 < k0);
        }
    }
}

// out[0, n) = x * y (+ u * v when un, vn > 0), the value below 2^(64 n);
// returns the trimmed length.  y and v are staged in LDS (B, B2; slot words
// each).  64 limbs at a time: lane l forms the column sum (s0, s1, s2) of
// column c = c0 + l (complete within the chunk), limb c gathers s0 of column
// c, s1 of column c - 1 and s2 of column c - 2 (from the lanes below, or the
// previous chunk's top lanes) as lo + 2^64 e (e <= 2), adds e of the limb
// below: x = t + 2^64 g; the carries into the chunk's limbs follow from the
// generate (g) and propagate (t = 2^64 - 1) masks by one integer addition:
// with A = G | P and B = G, the carry into bit i of A + B + cin is bit i of
// (A + B + cin) ^ P.
__device__ uint32_t wave_product(uint64_t *out, uint32_t n, const uint64_t *x, uint32_t xn, const uint64_t *y,
                                 uint32_t yn, const uint64_t *u, uint32_t un, const uint64_t *v, uint32_t vn,
                                 uint64_t *B, uint64_t *B2, uint32_t lane) {
    const bool p1 = xn && yn, p2 = un && vn;
    for (uint32_t k = lane; k < (p1 ? yn : 0); k += 64) B[k] = y[k];
    for (uint32_t k = lane; k < (p2 ? vn : 0); k += 64) B2[k] = v[k];
    __syncthreads();
    uint64_t s1_63 = 0, s2_62 = 0, s2_63 = 0;   // the previous chunk's top columns
    uint64_t e_in = 0, cin = 0;                  // e of the limb below the chunk, carry into it
    uint32_t top = 0;
    for (uint32_t c0 = 0; c0 < n; c0 += 64) {
        const uint32_t c = c0 + lane;
        uint64_t s0 = 0, s1 = 0, s2 = 0;
        if (p1) col_sum(s0, s1, s2, x, xn, B, yn, c0, lane);
        if (p2) col_sum(s0, s1, s2, u, un, B2, vn, c0, lane);
        const uint64_t d1 = lane_down(s1, lane, 1), d2 = lane_down(s2, lane, 2);
        const uint64_t w1 = lane ? d1 : s1_63;
        const uint64_t w2 = lane >= 2 ? d2 : (lane ? s2_63 : s2_62);
        uint64_t lo = s0 + w1;
        uint64_t e = lo < w1;
        lo += w2;
        e += lo < w2;
        const uint64_t de = lane_down(e, lane, 1);   // (every lane takes part in the permute)
        const uint64_t eb = lane ? de : e_in;
        const uint64_t t = lo + eb;
        const uint64_t G = __ballot(t < eb), Pm = __ballot(t == ~0ull);
        const uint64_t Cv = (((G | Pm) + G) + cin) ^ Pm;
        const uint64_t r = t + ((Cv >> lane) & 1);
        if (c < n) out[c] = r;
        const uint64_t nz = __ballot(c < n && r != 0);
        if (nz) top = c0 + 64 - (uint32_t)__clzll((long long)nz);
        cin = ((G >> 63) & 1) | ((Pm >> 63) & (Cv >> 63) & 1);
        e_in = lane_at(e, 63);
        s1_63 = lane_at(s1, 63);
        s2_62 = lane_at(s2, 62);
        s2_63 = lane_at(s2, 63);
    }
    __syncthreads();   // B, B2 are restaged by the next product
    return top;
}

// dst[0, store) = the low limbs of a + (neg ? ~b : b) + cin over m limbs (a:
// an limbs, b: bn limbs, zero beyond; dst may be null with store 0); returns
// the carry out of limb m - 1.  64 limbs at a time, the carries from the
// generate / propagate ballots as in wave_product.
__device__ uint32_t wave_add(uint64_t *dst, uint32_t store, const uint64_t *a, uint32_t an, const uint64_t *b,
                             uint32_t bn, bool neg, uint32_t cin, uint32_t m, uint32_t lane) {
    uint64_t c = cin;
    for (uint32_t c0 = 0; c0 < m; c0 += 64) {
        const uint32_t k = c0 + lane;
        const bool in = k < m;
        const uint64_t x = (in && k < an) ? a[k] : 0;
        uint64_t y = (in && k < bn) ? b[k] : 0;
        if (neg && in) y = ~y;
        const uint64_t s = x + y;
        const uint64_t G = __ballot(in && s < x), Pm = __ballot(in && s == ~0ull);
        const uint64_t Cv = (((G | Pm) + G) + c) ^ Pm;
        if (in && k < store) dst[k] = s + ((Cv >> lane) & 1);
        const uint32_t top = m - c0 < 64 ? m - c0 : 64;   // limbs of this chunk
        c = top == 64 ? (((G >> 63) & 1) | ((Pm >> 63) & (Cv >> 63) & 1)) : ((Cv >> top) & 1);
    }
    return (uint32_t)c;
}

// One division of the decimal conversion's split (to_decimal): y = q W + r by
// the level's reciprocal V = floor(2^sb / W) (sb = 2 bits(W)), the steps of the
// batched levels in one wave, intermediates in LDS:
//   ax = y V,  q = ax >> sb (qlen limbs),  p = q W,
//   r = y - p over max(n, qlen + nw) + 1 limbs (flag r >= 0: the carry out),
//   t = r - W (f = r >= W),  r2 = f ? t : r,  q2 = q + f,  check r2 < W.
// flags[0] = (r >= 0), flags[1] = f, flags[2] = (r2 >= W); the host reads them
// after the last level.
struct DecSplitJob {
    const uint64_t *y;
    uint64_t *r2, *q2;   // children: nw and qlen limbs
    uint64_t *flags;
    uint32_t n, qlen;
};

// LDS words of k_dec_split_wave for a level (nmax, qmax: the largest n and
// qlen of its jobs).
inline size_t dec_split_lds_words(size_t nmax, size_t qmax, size_t nw, size_t vb) {
    return 2 * std::max(nw, vb) + (nmax + vb) + qmax + (qmax + nw) + 2 * (nw + 1);
}

__global__ void __launch_bounds__(64) k_dec_split_wave(const DecSplitJob *jobs, uint32_t nj, const uint64_t *W,
                                                       uint32_t nw, const uint64_t *V, uint32_t vb, uint32_t sb,
                                                       uint32_t nmax, uint32_t qmax) {
    extern __shared__ uint64_t sh[];
    const uint32_t jb = blockIdx.x, lane = threadIdx.x;
    if (jb >= nj) return;
    const DecSplitJob J = jobs[jb];
    const uint32_t bm = nw > vb ? nw : vb;
    uint64_t *B = sh, *B2 = B + bm, *AX = B2 + bm, *Q = AX + (nmax + vb), *Pp = Q + qmax, *R = Pp + (qmax + nw),
             *T = R + (nw + 1);
    const uint32_t axn = J.n + vb;
    wave_product(AX, axn, J.y, J.n, V, vb, nullptr, 0, nullptr, 0, B, B2, lane);
    {
        const uint32_t s0 = sb >> 6, r = sb & 63;
        for (uint32_t k = lane; k < J.qlen; k += 64) {
            const uint32_t i = k + s0;
            const uint64_t lo = i < axn ? AX[i] : 0, hi = i + 1 < axn ? AX[i + 1] : 0;
            Q[k] = r ? (lo >> r) | (hi << (64 - r)) : lo;
        }
    }
    __syncthreads();
    wave_product(Pp, J.qlen + nw, Q, J.qlen, W, nw, nullptr, 0, nullptr, 0, B, B2, lane);
    const uint32_t m = (J.n > J.qlen + nw ? J.n : J.qlen + nw) + 1;
    const uint32_t f0 = wave_add(R, nw + 1, J.y, J.n, Pp, J.qlen + nw, true, 1, m, lane);
    __syncthreads();
    const uint32_t f = wave_add(T, nw + 1, R, nw + 1, W, nw, true, 1, nw + 1, lane);
    __syncthreads();
    const uint64_t *R2 = f ? T : R;
    for (uint32_t k = lane; k < nw; k += 64) J.r2[k] = R2[k];
    wave_add(J.q2, J.qlen, Q, J.qlen, nullptr, 0, false, f, J.qlen, lane);
    const uint32_t ge = wave_add(nullptr, 0, R2, nw + 1, W, nw, true, 1, nw + 1, lane);
    if (lane == 0) {
        J.flags[0] = f0;
        J.flags[1] = f;
        J.flags[2] = ge;
        J.flags[3] = 0;
    }
}

// Level j -> j + 1 with one 64-thread workgroup per merge (the arithmetic of
// k_chud_merge): Q = lq rq, P = lp rp, T = lt rq + rt lp; an unpaired last
// node moves up unchanged.  Dynamic LDS: 2 slot words of the input level.
__global__ void __launch_bounds__(64) k_chud_merge_wave(ChudLevel in, ChudLevel out) {
    extern __shared__ uint64_t sh[];
    const uint32_t i = blockIdx.x, lane = threadIdx.x;
    if (i >= out.count) return;
    const uint32_t l = 2 * i, r = l + 1;
    uint64_t *P = out.P + (size_t)i * out.slot, *Q = out.Q + (size_t)i * out.slot, *T = out.T + (size_t)i * out.slot;
    const uint64_t *lP = in.P + (size_t)l * in.slot, *lQ = in.Q + (size_t)l * in.slot, *lT = in.T + (size_t)l * in.slot;
    const uint32_t llp = in.lp[l], llq = in.lq[l], llt = in.lt[l];
    if (r >= in.count) {
        for (uint32_t k = lane; k < llp; k += 64) P[k] = lP[k];
        for (uint32_t k = lane; k < llq; k += 64) Q[k] = lQ[k];
        for (uint32_t k = lane; k < llt; k += 64) T[k] = lT[k];
        if (lane == 0) {
            out.lp[i] = llp;
            out.lq[i] = llq;
            out.lt[i] = llt;
        }
        return;
    }
    const uint64_t *rP = in.P + (size_t)r * in.slot, *rQ = in.Q + (size_t)r * in.slot, *rT = in.T + (size_t)r * in.slot;
    const uint32_t rlp = in.lp[r], rlq = in.lq[r], rlt = in.lt[r];
    uint64_t *B = sh, *B2 = sh + in.slot;
    const uint32_t q = wave_product(Q, llq + rlq, lQ, llq, rQ, rlq, nullptr, 0, nullptr, 0, B, B2, lane);
    const uint32_t p = wave_product(P, llp + rlp, lP, llp, rP, rlp, nullptr, 0, nullptr, 0, B, B2, lane);
    const uint32_t tn = (llt + rlq > rlt + llp ? llt + rlq : rlt + llp) + 1;
    const uint32_t t = wave_product(T, tn, lT, llt, rQ, rlq, rT, rlt, lP, llp, B, B2, lane);
    if (lane == 0) {
        out.lp[i] = p;
        out.lq[i] = q;
        out.lt[i] = t;
    }
}

struct LenReq {
    const uint64_t *p;
    uint32_t n;
};
__global__ void k_lengths(const LenReq *req, uint32_t n, uint32_t *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t k = req[i].n;
    while (k && !req[i].p[k - 1]) --k;
    out[i] = k;
}

// ---- binary splitting ------------------------------------------------------------

namespace {

struct Node {
    const uint64_t *P = nullptr, *Q = nullptr, *T = nullptr;
    size_t lp = 0, lq = 0, lt = 0;
};

unsigned blocks_for(size_t n, unsigned t = 256) { return (unsigned)((n + t - 1) / t); }

// Uploads a host vector through the stream (image kept alive until sync).
template <class T>
int upload(DBuf &dst, const std::vector<T> &v) {
    const size_t bytes = std::max<size_t>(v.size() * sizeof(T), 8);
    if (!dst.reset((bytes + 7) / 8)) return set_error(PG_ENOMEM, "descriptor upload");
    HostImg img = ctx().take_img(bytes);
    std::memset(img.data(), 0, bytes);
    if (!v.empty()) std::memcpy(img.data(), v.data(), v.size() * sizeof(T));
    HIP_TRY(hipMemcpyAsync(dst.d, img.data(), bytes, hipMemcpyHostToDevice, ctx().stream));
    ctx().hold(std::move(img));
    return PG_OK;
}

// Trimmed lengths of many device arrays (one launch, one read).
int lengths(const std::vector<LenReq> &req, std::vector<uint32_t> &out) {
    Context &C = ctx();
    out.assign(req.size(), 0);
    if (req.empty()) return PG_OK;
    DBuf dreq, dout((req.size() * 4 + 7) / 8);
    int st = upload(dreq, req);
    if (st) return st;
    if (!dout.d) return set_error(PG_ENOMEM, "lengths");
    hipLaunchKernelGGL(k_lengths, dim3(blocks_for(req.size())), dim3(256), 0, C.stream, (const LenReq *)dreq.d,
                       (uint32_t)req.size(), (uint32_t *)dout.d);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out.data(), dout.d, req.size() * 4, hipMemcpyDeviceToHost, C.stream));
    return C.sync();
}

}  // namespace

int chud_split(size_t nterms, DBuf &Pout, size_t &lp, DBuf &Qout, size_t &lq, DBuf &Tout, size_t &lt,
               bool need_p, std::vector<DBuf> &keep) {
    Context &C = ctx();
    lp = lq = lt = 0;
    if (!nterms) return set_error(PG_EINVAL, "binary splitting needs at least one term");
    if (nterms > ((size_t)1 << 20)) return set_error(PG_ERANGE, "binary splitting limited to 2^20 terms");
    // per-thread levels, then one wave per merge
    ChudLevel lv[CHUD_DEVICE_LEVELS + 1]{};
    size_t count = (nterms + 1) / 2;
    int st;
    for (int j = 1; j <= CHUD_DEVICE_LEVELS; ++j) {
        ChudLevel &L = lv[j];
        L.count = (uint32_t)count;
        L.slot = 3u << j;
        keep.emplace_back(3 * count * L.slot);
        keep.emplace_back((3 * count * 4 + 7) / 8);
        if (!keep[keep.size() - 2].d || !keep.back().d) return set_error(PG_ENOMEM, "binary splitting");
        uint64_t *base = keep[keep.size() - 2].d;
        uint32_t *lens = (uint32_t *)keep.back().d;
        L.P = base; L.Q = base + count * L.slot; L.T = base + 2 * count * L.slot;
        L.lp = lens; L.lq = lens + count; L.lt = lens + 2 * count;
        hipEvent_t e0 = nullptr;
        C.stage_begin(e0);
        if (j == 1)
            hipLaunchKernelGGL(k_chud_base, dim3(blocks_for(count)), dim3(256), 0, C.stream, (uint64_t)nterms, L);
        else if (j <= CHUD_THREAD_LEVELS)
            hipLaunchKernelGGL(k_chud_merge, dim3(blocks_for(count)), dim3(256), 0, C.stream, lv[j - 1], L);
        else
            hipLaunchKernelGGL(k_chud_merge_wave, dim3((unsigned)count), dim3(64), 2 * (size_t)lv[j - 1].slot * 8,
                               C.stream, lv[j - 1], L);
        C.stage_end(j <= CHUD_THREAD_LEVELS ? "chud_thread" : "chud_wave", e0, 0.0);
        HIP_TRY(hipGetLastError());
        if (count == 1 && j < CHUD_DEVICE_LEVELS) {   // done early: carry the root up
            lv[CHUD_DEVICE_LEVELS] = L;
            break;
        }
        count = (count + 1) / 2;
    }
    // nodes of the last device level
    const ChudLevel &top = lv[CHUD_DEVICE_LEVELS];
    std::vector<uint32_t> lens(3 * (size_t)top.count);
    HIP_TRY(hipMemcpyAsync(lens.data(), top.lp, lens.size() * 4, hipMemcpyDeviceToHost, C.stream));
    if ((st = C.sync())) return st;
    std::vector<Node> nodes(top.count);
    for (size_t i = 0; i < top.count; ++i) {
        Node &n = nodes[i];
        n.P = top.P + i * top.slot; n.Q = top.Q + i * top.slot; n.T = top.T + i * top.slot;
        n.lp = lens[i]; n.lq = lens[top.count + i]; n.lt = lens[2 * top.count + i];
    }
    // batched levels
    while (nodes.size() > 1) {
        const size_t nm = nodes.size() / 2, np = (nodes.size() + 1) / 2;
        size_t total = 0;
        for (size_t k = 0; k < nm; ++k) {
            const Node &l = nodes[2 * k], &r = nodes[2 * k + 1];
            const bool wantp = need_p || k + 1 < np;
            total += (wantp ? l.lp + r.lp : 0) + l.lq + r.lq + 2 * (l.lt + r.lq) + 2 * (r.lt + l.lp) + 1;
        }
        keep.emplace_back(total + 1);
        DBuf &arena = keep.back();
        if (!arena.d) return set_error(PG_ENOMEM, "binary splitting");
        std::vector<pg_job> jobs;
        std::vector<CarryNode> lin;
        std::vector<Node> up(np);
        std::vector<size_t> bound(3 * np);
        uint64_t *cur = arena.d;
        auto job = [&](const uint64_t *a, size_t an, const uint64_t *b, size_t bn) {
            uint64_t *r = cur;
            cur += an + bn;
            if (an && bn) {
                pg_job j{};
                j.a = a; j.an = an; j.b = b; j.bn = bn; j.r = r; j.r_cap = an + bn; j.path = PG_PATH_AUTO;
                jobs.push_back(j);
            } else {
                lin.push_back(lin_node(r, std::max<size_t>(an + bn, 1)));   // zero fill
            }
            return r;
        };
        for (size_t k = 0; k < nm; ++k) {
            const Node &l = nodes[2 * k], &r = nodes[2 * k + 1];
            const bool wantp = need_p || k + 1 < np;
            Node &u = up[k];
            if (wantp) u.P = job(l.P, l.lp, r.P, r.lp);
            u.Q = job(l.Q, l.lq, r.Q, r.lq);
            const uint64_t *t1 = job(l.T, l.lt, r.Q, r.lq);
            const uint64_t *t2 = job(r.T, r.lt, l.P, l.lp);
            const size_t tn = std::max(l.lt + r.lq, r.lt + l.lp) + 1;
            uint64_t *t = cur;
            cur += tn;
            CarryNode c = lin_node(t, tn);
            c.x = t1; c.xn = l.lt + r.lq; c.y = t2; c.yn = r.lt + l.lp;
            lin.push_back(c);
            u.T = t;
            bound[3 * k] = wantp ? l.lp + r.lp : 0; bound[3 * k + 1] = l.lq + r.lq; bound[3 * k + 2] = tn;
        }
        if (np > nm) {
            up[nm] = nodes.back();
            bound[3 * nm] = up[nm].lp; bound[3 * nm + 1] = up[nm].lq; bound[3 * nm + 2] = up[nm].lt;
        }
        // the zero-fill nodes must run before the sums read them: products
        // of an empty operand are written by the fills, the sums come after
        std::vector<CarryNode> fills, sums;
        for (auto &c : lin) (c.x ? sums : fills).push_back(c);
        if (!fills.empty() && (st = lin_run(fills.data(), fills.size(), "chud_sum"))) return st;
        if (!jobs.empty()) {
            Batch b;
            b.use_graph = false;
            if ((st = b.build(jobs.data(), jobs.size()))) return st;
            if ((st = b.run())) return st;
        }
        if ((st = lin_run(sums.data(), sums.size(), "chud_sum"))) return st;
        // the next level is planned on the length bounds (a product of la and lb
        // limbs has la + lb; T one more), not on trimmed lengths read back, so
        // the host plans it while this one runs: the operands only carry a few
        // zero top limbs more
        for (size_t k = 0; k < np; ++k) {
            up[k].lp = up[k].P ? bound[3 * k] : 0;
            up[k].lq = bound[3 * k + 1];
            up[k].lt = bound[3 * k + 2];
        }
        nodes.swap(up);
    }
    Node root = nodes[0];
    {   // the root's trimmed lengths
        std::vector<LenReq> req{{root.P, (uint32_t)root.lp}, {root.Q, (uint32_t)root.lq}, {root.T, (uint32_t)root.lt}};
        std::vector<uint32_t> got;
        if ((st = lengths(req, got))) return st;
        root.lp = root.P ? got[0] : 0;
        root.lq = got[1];
        root.lt = got[2];
    }
    lp = root.lp; lq = root.lq; lt = root.lt;
    if (!Pout.reset(std::max<size_t>(lp, 1)) || !Qout.reset(std::max<size_t>(lq, 1)) ||
        !Tout.reset(std::max<size_t>(lt, 1)))
        return set_error(PG_ENOMEM, "binary splitting");
    if ((st = dcopy(Pout.d, root.P, lp)) || (st = dcopy(Qout.d, root.Q, lq)) || (st = dcopy(Tout.d, root.T, lt)))
        return st;
    return C.sync();
}

// ---- integer square root ----------------------------------------------------------

namespace {

std::vector<uint64_t> small_isqrt(std::vector<uint64_t> a) {
    // digit-by-digit binary square root on at most a few limbs
    auto cmp = [](const std::vector<uint64_t> &x, const std::vector<uint64_t> &y) {
        for (size_t i = std::max(x.size(), y.size()); i-- > 0;) {
            const uint64_t u = i < x.size() ? x[i] : 0, v = i < y.size() ? y[i] : 0;
            if (u != v) return u < v ? -1 : 1;
        }
        return 0;
    };
    const size_t n = a.size() + 1;
    a.resize(n, 0);
    std::vector<uint64_t> res(n, 0), bit(n, 0), t(n, 0);
    size_t hb = 0;
    for (size_t i = 0; i < 64 * n; ++i)
        if ((a[i / 64] >> (i % 64)) & 1) hb = i;
    hb &= ~(size_t)1;
    bit[hb / 64] = 1ull << (hb % 64);
    for (;;) {
        uint64_t c = 0;
        for (size_t i = 0; i < n; ++i) {   // t = res + bit
            const uint64_t s = res[i] + bit[i];
            const uint64_t s2 = s + c;
            c = (s < res[i]) | (s2 < s);
            t[i] = s2;
        }
        const bool ge = cmp(a, t) >= 0;
        if (ge) {
            uint64_t br = 0;
            for (size_t i = 0; i < n; ++i) {
                const uint64_t d = a[i] - t[i] - br;
                br = (a[i] < t[i]) || (a[i] - t[i] < br);
                a[i] = d;
            }
        }
        for (size_t i = 0; i < n; ++i) res[i] = (res[i] >> 1) | (i + 1 < n ? res[i + 1] << 63 : 0);
        if (ge)
            for (size_t i = 0; i < n; ++i) {   // res += bit
                const uint64_t s = res[i] + bit[i];
                const uint64_t cc = s < res[i];
                res[i] = s;
                for (size_t k = i + 1; cc && k < n; ++k)
                    if (++res[k]) break;
            }
        bool zero = true;
        for (size_t i = 0; i < n; ++i) {
            bit[i] = (bit[i] >> 2) | (i + 1 < n ? bit[i + 1] << 62 : 0);
            zero &= bit[i] == 0;
        }
        if (zero) break;
    }
    while (!res.empty() && !res.back()) res.pop_back();
    return res;
}

// Makes an estimate x (xl limbs, within a few units of the root) the floor
// root of a: r = a - x^2, then the two unit loops of sqrt.cpp:48-58.
int sqrt_fix(DBuf &x, size_t &xl, const uint64_t *a, size_t an) {
    int st;
    if ((st = top_len(x.d, xl, &xl))) return st;
    if (!xl) xl = 1;
    // r = a - x^2, then the two loops of sqrt.cpp:48-58
    const size_t rm = std::max(an, 2 * xl) + 2;
    DBuf sq(2 * xl), r(rm), r2(rm), x2(xl + 1);
    if (!sq.d || !r.d || !r2.d || !x2.d) return set_error(PG_ENOMEM, "sqrt");
    if ((st = dmul(sq.d, x.d, xl, x.d, xl))) return st;
    uint64_t *flag = ctx().words + 17;
    CarryNode c = lin_node(r.d, rm);
    c.x = a; c.xn = an; c.y = sq.d; c.yn = 2 * xl; c.yneg = 1; c.cin = 1; c.flag = flag;
    if ((st = lin_one(c, "sqrt"))) return st;
    DBuf xx(xl + 1);
    if (!xx.d || (st = dset(xx.d, xl + 1, nullptr, 0)) || (st = dcopy(xx.d, x.d, xl))) return st ? st : PG_ENOMEM;
    size_t xn = xl + 1;
    for (int it = 0; it < 64; ++it) {
        uint64_t f = 0;
        if ((st = read_words(flag, 1, &f))) return st;
        if (f == 0) {   // x^2 > a: x -= 1, r += 2 x_new + 1
            CarryNode m = lin_node(x2.d, xn);
            m.x = xx.d; m.xn = xn; m.yneg = 1;
            if ((st = lin_one(m, "sqrt"))) return st;
            CarryNode u = lin_node(r2.d, rm);
            u.x = r.d; u.xn = rm; u.y = x2.d; u.yn = xn; u.ysh = -1; u.cin = 1; u.flag = flag;
            if ((st = lin_one(u, "sqrt"))) return st;
        } else {        // (x + 1)^2 <= a  <=>  r >= 2 x + 1
            CarryNode u = lin_node(r2.d, rm);
            u.x = r.d; u.xn = rm; u.y = xx.d; u.yn = xn; u.ysh = -1; u.yneg = 1; u.flag = flag;
            if ((st = lin_one(u, "sqrt"))) return st;
            uint64_t g = 0;
            if ((st = read_words(flag, 1, &g))) return st;
            if (g == 0) {
                x = std::move(xx);
                return top_len(x.d, xn, &xl);
            }
            CarryNode m = lin_node(x2.d, xn);
            m.x = xx.d; m.xn = xn; m.cin = 1;
            if ((st = lin_one(m, "sqrt"))) return st;
        }
        std::swap(xx, x2);
        std::swap(r, r2);
    }
    return set_error(PG_EINTERNAL, "square root correction did not converge");
}

// floor(sqrt(a)) for a of 5 or more limbs.  With B = bit_length(a), E = B
// rounded up to even and m = a / 2^E in [1/4, 1), y = Y / 2^P approximates
// 1/sqrt(m) from below by Newton's iteration y <- y + y (1 - m y^2) / 2 at
// doubling precision P: each step squares Y, multiplies by the top Q bits of
// a rounded up (so the correction term never overshoots; its sign word zeroes
// it when y is already exact to that rounding), multiplies by Y and shifts.
// Then s = floor(a Y / 2^(P + E/2)) is the root or one below it, and
// sqrt_fix makes it exact.  Every multiply runs on the GPU product engine.
int isqrt_newton(DBuf &x, size_t &xl, const uint64_t *a, size_t an) {
    int st;
    size_t B = 0;
    if ((st = bit_length(a, an, &B))) return st;
    const size_t E = B + (B & 1);
    // seed: s_t = isqrt of the top 128 bits (Mt = floor(a / 2^(E - 128))),
    // Y0 = floor(2^(P0 + 64) / (s_t + 1)) <= 2^(P0 + E/2) / sqrt(a)
    const size_t QT = 128, drop = E - QT, l0 = drop / 64;
    std::vector<uint64_t> w(an - l0), mt(3, 0);
    if ((st = read_words(a + l0, w.size(), w.data()))) return st;
    {
        const int r = (int)(drop % 64);
        for (size_t i = 0; i < 3 && i < w.size(); ++i)
            mt[i] = r ? (w[i] >> r) | (i + 1 < w.size() ? w[i + 1] << (64 - r) : 0) : w[i];
        mt.resize(2);
    }
    std::vector<uint64_t> stt = small_isqrt(mt);
    stt.push_back(0);
    for (auto &v : stt)
        if (++v) break;
    while (!stt.empty() && !stt.back()) stt.pop_back();
    size_t P = 60;
    const std::vector<uint64_t> y0 = pow2_div_small(P + QT / 2, stt);
    size_t yl = limbs_for_bits(P + 2) + 1;
    DBuf Y(yl), F(1);
    if (!Y.d || !F.d || (st = dset(Y.d, yl, y0.data(), y0.size()))) return st ? st : set_error(PG_ENOMEM, "sqrt");
    const size_t target = B / 2 + 64;
    while (P < target) {
        const size_t P2 = std::min(target, 2 * P - 8), Q = P2 + 8;
        // mt = top Q bits of a, rounded up
        const size_t mtl = limbs_for_bits(Q + 1);
        DBuf M(mtl), Y2(2 * yl), T(mtl + 2 * yl);
        if (!M.d || !Y2.d || !T.d) return set_error(PG_ENOMEM, "sqrt");
        CarryNode c = lin_node(M.d, mtl);
        c.x = a; c.xn = an; c.xsh = (int64_t)(E - Q); c.cin = 1;
        if ((st = lin_one(c, "sqrt"))) return st;
        if ((st = dmul(Y2.d, Y.d, yl, Y.d, yl)) || (st = dmul(T.d, M.d, mtl, Y2.d, 2 * yl))) return st;
        // D = 2^(Q + 2P) - mt Y^2 (sign word F), Du = D if D >= 0 else 0
        const size_t dm = std::max(limbs_for_bits(Q + 2 * P + 1), mtl + 2 * yl) + 1;
        DBuf D(dm), Du(dm);
        if (!D.d || !Du.d) return set_error(PG_ENOMEM, "sqrt");
        CarryNode d = lin_node(D.d, dm);
        d.pow2 = (int64_t)(Q + 2 * P); d.y = T.d; d.yn = mtl + 2 * yl; d.yneg = 1; d.cin = 1; d.flag = F.d;
        CarryNode u = lin_node(Du.d, dm);
        u.y = D.d; u.yn = dm; u.cond = F.d; u.cond_eq = 1;
        if ((st = lin_one(d, "sqrt")) || (st = lin_one(u, "sqrt"))) return st;
        size_t dl = 0;
        if ((st = top_len(Du.d, dm, &dl))) return st;
        // Y' = Y 2^(P2 - P) + floor(Y Du / 2^(Q + 3P - P2 + 1))
        const size_t yl2 = limbs_for_bits(P2 + 2) + 1;
        DBuf C(yl + std::max<size_t>(dl, 1)), Yn(yl2);
        if (!C.d || !Yn.d) return set_error(PG_ENOMEM, "sqrt");
        if ((st = dmul(C.d, Y.d, yl, Du.d, std::max<size_t>(dl, 1)))) return st;
        CarryNode n = lin_node(Yn.d, yl2);
        n.x = Y.d; n.xn = yl; n.xsh = -(int64_t)(P2 - P);
        n.y = C.d; n.yn = yl + std::max<size_t>(dl, 1); n.ysh = (int64_t)(Q + 3 * P - P2 + 1);
        if ((st = lin_one(n, "sqrt"))) return st;
        Y = std::move(Yn);
        yl = yl2;
        P = P2;
    }
    // s = floor(a Y / 2^(P + E/2))
    xl = limbs_for_bits(B / 2 + 2) + 1;
    DBuf AY(an + yl);
    if (!AY.d || !x.reset(xl)) return set_error(PG_ENOMEM, "sqrt");
    if ((st = dmul(AY.d, a, an, Y.d, yl))) return st;
    CarryNode c = lin_node(x.d, xl);
    c.x = AY.d; c.xn = an + yl; c.xsh = (int64_t)(P + E / 2);
    if ((st = lin_one(c, "sqrt"))) return st;
    return sqrt_fix(x, xl, a, an);
}

}  // namespace

int isqrt(DBuf &x, size_t &xl, const uint64_t *a, size_t an) {
    xl = 0;
    int st;
    if (!an) return PG_OK;
    if (an <= 4) {   // sqrt.cpp:78: small operands on the host
        std::vector<uint64_t> h(an);
        if ((st = read_words(a, an, h.data()))) return st;
        const std::vector<uint64_t> r = small_isqrt(h);
        xl = r.size();
        if (!x.reset(std::max<size_t>(xl, 1))) return set_error(PG_ENOMEM, "sqrt");
        return dset(x.d, std::max<size_t>(xl, 1), r.data(), r.size());
    }
    return isqrt_newton(x, xl, a, an);
}

// ---- decimal conversion ------------------------------------------------------------

struct DecLeaf {
    const uint64_t *p;
    uint32_t n;          // <= DEC_LEAF limbs, value < 10^(19 DEC_LEAF)
    uint64_t chunk0;     // index of its lowest 19-digit chunk
};

// Division of (u1, u0) by the normalized invariant d with v = floor((2^128 - 1) / d) - 2^64:
// Moller and Granlund, "Improved division by invariant integers" (IEEE TC 60(2), 2011), Alg. 4.
__device__ __forceinline__ uint64_t div2by1(uint64_t u1, uint64_t u0, uint64_t d, uint64_t v, uint64_t &r) {
    uint64_t q0 = v * u1, q1 = __umul64hi(v, u1);
    const uint64_t s = q0 + u0;
    q1 += u1 + (s < q0);
    q0 = s;
    q1 += 1;
    r = u0 - q1 * d;
    if (r > q0) {
        q1 -= 1;
        r += d;
    }
    if (r >= d) {
        q1 += 1;
        r -= d;
    }
    return q1;
}

// One thread per leaf: DEC_LEAF chunks of 19 digits into out (most
// significant digit first; total length dcap).
__global__ void __launch_bounds__(64) k_dec_leaf(const DecLeaf *leaves, uint32_t n, uint64_t v, char *out,
                                                 uint64_t dcap) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const DecLeaf L = leaves[i];
    uint64_t y[DEC_LEAF];
#pragma unroll
    for (int k = 0; k < DEC_LEAF; ++k) y[k] = k < (int)L.n ? L.p[k] : 0;
    for (int c = 0; c < DEC_LEAF; ++c) {
        uint64_t r = 0;
#pragma unroll
        for (int k = DEC_LEAF - 1; k >= 0; --k) y[k] = div2by1(r, y[k], TEN19, v, r);
        const uint64_t pos = (L.chunk0 + c) * 19;
        uint64_t x = r;
        for (int t = 0; t < 19; ++t) {
            const uint64_t q = __umul64hi(x, 0xCCCCCCCCCCCCCCCDull) >> 3;   // x / 10
            out[dcap - 1 - (pos + t)] = (char)('0' + (x - q * 10));
            x = q;
        }
    }
}

__global__ void k_fill_chars(char *p, uint64_t n, char c) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = c;
}

namespace {
size_t bits_of_top(uint64_t top, size_t limbs) { return limbs ? (limbs - 1) * 64 + 64 - (size_t)__builtin_clzll(top) : 0; }
}  // namespace

int to_decimal(const uint64_t *x, size_t xn, std::string &out, size_t point, char *dst, size_t cap, size_t *dlen) {
    Context &C = ctx();
    out.clear();
    // the result, appended piece by piece to dst (up to cap) or to out
    size_t pos = 0;
    bool over = false;
    auto put = [&](const char *p, size_t n) {
        if (!dst) out.append(p, n);
        else if (pos + n > cap) over = true;
        else std::memcpy(dst + pos, p, n);
        pos += n;
    };
    auto fill = [&](char ch, size_t n) {
        if (!dst) out.append(n, ch);
        else if (pos + n > cap) over = true;
        else std::memset(dst + pos, ch, n);
        pos += n;
    };
    auto done = [&]() {
        if (dlen) *dlen = pos;
        return over ? set_error(PG_EINVAL, "output capacity too small") : PG_OK;
    };
    int st;
    size_t xb = 0;
    if ((st = top_len(x, xn, &xn))) return st;
    if (!xn) {
        if (point) {
            put("0.", 2);
            fill('0', point);
        }
        return done();
    }
    if ((st = bit_length(x, xn, &xb))) return st;
    // chunks of 19 digits: 2^L >= ceil(digits / 19) with digits <= bits * log10(2) + 1
    const size_t maxdig = (size_t)((double)xb * 0.30103) + 2;
    const size_t nchunk = (maxdig + 18) / 19;
    int L = 0;
    while (((size_t)1 << L) < nchunk) ++L;
    const int leaf_log = 5;   // DEC_LEAF = 2^5
    const int top = std::max(L, leaf_log);
    // W_i = (10^19)^(2^i), i < top; V_i = floor(2^(2 bits(W_i)) / W_i) for i >= leaf_log
    std::vector<DBuf> W(top), V(top);
    std::vector<size_t> wl(top), wb(top), vl(top);
    {
        const uint64_t ten19 = TEN19;
        if (!W[0].reset(1) || (st = dset(W[0].d, 1, &ten19, 1))) return st ? st : set_error(PG_ENOMEM, "powers");
        wl[0] = 1;
        wb[0] = 64;
        for (int i = 1; i < top; ++i) {
            if (!W[i].reset(2 * wl[i - 1])) return set_error(PG_ENOMEM, "powers");
            if ((st = dmul(W[i].d, W[i - 1].d, wl[i - 1], W[i - 1].d, wl[i - 1]))) return st;
            if ((st = top_len(W[i].d, 2 * wl[i - 1], &wl[i])) || (st = bit_length(W[i].d, wl[i], &wb[i]))) return st;
        }
        // V_top-1 by Newton's iteration; below it, since W_(i+1) = W_i^2,
        //   2^(2 b_i) / W_i = W_i 2^(2 b_i) / W_(i+1),
        // so floor(W_i V_(i+1) / 2^(2 b_(i+1) - 2 b_i)) is V_i or V_i - 1
        // (V_(i+1) is within 1 of 2^(2 b_(i+1)) / W_(i+1), and W_i is below
        // 2^(2 b_(i+1) - 2 b_i)); recip_finish makes it exact.
        for (int i = top - 1; i >= leaf_log; --i) {
            if (i == top - 1) {
                if ((st = recip_exact(V[i], vl[i], W[i].d, wl[i], 2 * wb[i]))) return st;
                continue;
            }
            const size_t pl = wl[i] + vl[i + 1], sh = 2 * wb[i + 1] - 2 * wb[i];
            DBuf P(pl);
            if (!P.d) return set_error(PG_ENOMEM, "powers");
            if ((st = dmul(P.d, W[i].d, wl[i], V[i + 1].d, vl[i + 1]))) return st;
            vl[i] = limbs_for_bits(wb[i] + 2) + 1;
            if (!V[i].reset(vl[i])) return set_error(PG_ENOMEM, "powers");
            CarryNode c = lin_node(V[i].d, vl[i]);
            c.x = P.d;
            c.xn = pl;
            c.xsh = (int64_t)sh;
            if ((st = lin_one(c, "decimal")) || (st = recip_finish(V[i], vl[i], W[i].d, wl[i], 2 * wb[i]))) return st;
        }
    }
    struct DNode {
        const uint64_t *p;
        size_t n, bits, chunk0;
    };
    std::vector<DNode> nodes{{x, xn, xb, 0}};
    std::vector<DBuf> keep;
    struct Check {
        const uint64_t *flags;   // 4 words per division: r >= 0 (1), -, r2 >= W (0), -
        size_t n;
    };
    std::vector<Check> checks;
    for (int i = top - 1; i >= leaf_log; --i) {
        const size_t nw = wl[i], vb = vl[i], sb = 2 * wb[i];
        std::vector<DNode> next;
        std::vector<size_t> jobs_of;   // node index of each division
        for (size_t k = 0; k < nodes.size(); ++k) {
            if (nodes[k].bits + 1 <= wb[i]) next.push_back(nodes[k]);   // value < 2^(wb-1) <= W_i
            else jobs_of.push_back(k);
        }
        const size_t nj = jobs_of.size();
        if (!nj) {
            nodes.swap(next);
            continue;
        }
        std::vector<size_t> qlen(nj), off(nj);
        size_t nmax = 0, qmax = 0;
        for (size_t j = 0; j < nj; ++j) {
            const DNode &y = nodes[jobs_of[j]];
            qlen[j] = limbs_for_bits(y.bits + 2 - wb[i]) + 1;
            nmax = std::max(nmax, y.n);
            qmax = std::max(qmax, qlen[j]);
        }
        if (nw <= DEC_WAVE_MAX_LIMBS && dec_split_lds_words(nmax, qmax, nw, vb) * 8 <= 64 * 1024) {
            // one wave per division (k_dec_split_wave): only the children r2 (nw
            // limbs) and q2 (ql) are stored
            size_t total = 0;
            for (size_t j = 0; j < nj; ++j) {
                off[j] = total;
                total += nw + qlen[j];
            }
            keep.emplace_back(total);
            DBuf fl(4 * nj), dj;
            if (!keep.back().d || !fl.d) return set_error(PG_ENOMEM, "decimal conversion");
            std::vector<DecSplitJob> jobs(nj);
            for (size_t j = 0; j < nj; ++j) {
                const DNode &y = nodes[jobs_of[j]];
                uint64_t *r2 = keep.back().d + off[j], *q2 = r2 + nw;
                jobs[j] = DecSplitJob{y.p, r2, q2, fl.d + 4 * j, (uint32_t)y.n, (uint32_t)qlen[j]};
                const size_t qb = std::min(y.bits + 2 - wb[i], wb[i]);
                next.push_back({r2, nw, wb[i], y.chunk0});
                next.push_back({q2, std::min(qlen[j], limbs_for_bits(qb)), qb, y.chunk0 + ((size_t)1 << i)});
            }
            if ((st = upload(dj, jobs))) return st;
            const size_t lds = dec_split_lds_words(nmax, qmax, nw, vb) * 8;
            hipEvent_t e0 = nullptr;
            C.stage_begin(e0);
            hipLaunchKernelGGL(k_dec_split_wave, dim3((unsigned)nj), dim3(64), lds, C.stream, (const DecSplitJob *)dj.d,
                               (uint32_t)nj, W[i].d, (uint32_t)nw, V[i].d, (uint32_t)vb, (uint32_t)sb, (uint32_t)nmax,
                               (uint32_t)qmax);
            C.stage_end("decimal_wave", e0, 0.0);
            HIP_TRY(hipGetLastError());
            checks.push_back({fl.d, nj});
            keep.push_back(std::move(fl));
            keep.push_back(std::move(dj));
            nodes.swap(next);
            continue;
        }
        // per job: ax (n + vb), q (ql), p (ql + nw), r (nw + 1), t (nw + 1), r2 (nw + 1), q2 (ql)
        size_t total = 0;
        for (size_t j = 0; j < nj; ++j) {
            const DNode &y = nodes[jobs_of[j]];
            qlen[j] = limbs_for_bits(y.bits + 2 - wb[i]) + 1;
            off[j] = total;
            total += (y.n + vb) + qlen[j] + (qlen[j] + nw) + 3 * (nw + 1) + qlen[j] + 1;
        }
        keep.emplace_back(total);
        DBuf fl(4 * nj);
        if (!keep.back().d || !fl.d) return set_error(PG_ENOMEM, "decimal conversion");
        uint64_t *A = keep.back().d, *F = fl.d;
        std::vector<pg_job> j1(nj), j2(nj);
        std::vector<CarryNode> lq(nj), lr(nj), lt(nj), lc(2 * nj), lv(nj);
        for (size_t j = 0; j < nj; ++j) {
            const DNode &y = nodes[jobs_of[j]];
            uint64_t *ax = A + off[j], *q = ax + y.n + vb, *p = q + qlen[j], *r = p + qlen[j] + nw;
            uint64_t *t = r + nw + 1, *r2 = t + nw + 1, *q2 = r2 + nw + 1;
            j1[j] = pg_job{};
            j1[j].a = y.p; j1[j].an = y.n; j1[j].b = V[i].d; j1[j].bn = vb; j1[j].r = ax; j1[j].r_cap = y.n + vb;
            lq[j] = lin_node(q, qlen[j]);
            lq[j].x = ax; lq[j].xn = y.n + vb; lq[j].xsh = (int64_t)sb;
            j2[j] = pg_job{};
            j2[j].a = q; j2[j].an = qlen[j]; j2[j].b = W[i].d; j2[j].bn = nw; j2[j].r = p; j2[j].r_cap = qlen[j] + nw;
            // r = y - q0 W >= 0 (flag 1), fits nw + 1 limbs
            lr[j] = lin_node(r, std::max(y.n, qlen[j] + nw) + 1);
            lr[j].out_len = nw + 1;
            lr[j].x = y.p; lr[j].xn = y.n; lr[j].y = p; lr[j].yn = qlen[j] + nw; lr[j].yneg = 1; lr[j].cin = 1;
            lr[j].flag = F + 4 * j;
            // t = r - W, flag f = (r >= W)
            lt[j] = lin_node(t, nw + 1);
            lt[j].x = r; lt[j].xn = nw + 1; lt[j].y = W[i].d; lt[j].yn = nw; lt[j].yneg = 1; lt[j].cin = 1;
            lt[j].flag = F + 4 * j + 1;
            // r2 = f ? r - W : r,  q2 = q + f
            lc[2 * j] = lin_node(r2, nw + 1);
            lc[2 * j].x = r; lc[2 * j].xn = nw + 1; lc[2 * j].y = W[i].d; lc[2 * j].yn = nw; lc[2 * j].yneg = 1;
            lc[2 * j].cin = 1; lc[2 * j].cond = F + 4 * j + 1; lc[2 * j].cond_eq = 1;
            lc[2 * j + 1] = lin_node(q2, qlen[j]);
            lc[2 * j + 1].x = q; lc[2 * j + 1].xn = qlen[j]; lc[2 * j + 1].cin = 1;
            lc[2 * j + 1].cond = F + 4 * j + 1; lc[2 * j + 1].cond_eq = 1;
            // check: r2 < W (flag 0)
            lv[j] = lin_node(t, nw + 1);
            lv[j].x = r2; lv[j].xn = nw + 1; lv[j].y = W[i].d; lv[j].yn = nw; lv[j].yneg = 1; lv[j].cin = 1;
            lv[j].flag = F + 4 * j + 2;
            // children: r < W_i and q = floor(y / W_i) < W_i, q < 2^(bits - wb + 2)
            const size_t qb = std::min(y.bits + 2 - wb[i], wb[i]);
            next.push_back({r2, nw, wb[i], y.chunk0});
            next.push_back({q2, std::min(qlen[j], limbs_for_bits(qb)), qb, y.chunk0 + ((size_t)1 << i)});
        }
        {
            Batch b;
            b.use_graph = false;
            if ((st = b.build(j1.data(), nj)) || (st = b.run())) return st;
        }
        if ((st = lin_run(lq.data(), nj, "decimal"))) return st;
        {
            Batch b;
            b.use_graph = false;
            if ((st = b.build(j2.data(), nj)) || (st = b.run())) return st;
        }
        if ((st = lin_run(lr.data(), nj, "decimal")) || (st = lin_run(lt.data(), nj, "decimal")) ||
            (st = lin_run(lc.data(), 2 * nj, "decimal")) || (st = lin_run(lv.data(), nj, "decimal")))
            return st;
        // the split checks are read once, after the last level: the host plans
        // the next level from bit bounds alone, so levels queue back to back
        checks.push_back({F, nj});
        keep.push_back(std::move(fl));
        nodes.swap(next);
    }
    // leaves
    const size_t kcap = (size_t)1 << top;
    const uint64_t dcap = 19 * (uint64_t)kcap;
    std::vector<DecLeaf> leaves;
    for (const DNode &n : nodes) {
        if (n.n > (size_t)DEC_LEAF) return set_error(PG_EINTERNAL, "decimal leaf too long");
        leaves.push_back({n.p, (uint32_t)n.n, n.chunk0});
    }
    DBuf dl, chars((dcap + 7) / 8);
    if (!chars.d) return set_error(PG_ENOMEM, "decimal string");
    if ((st = upload(dl, leaves))) return st;
    const uint64_t v = (uint64_t)((~(unsigned __int128)0) / TEN19 - ((unsigned __int128)1 << 64));
    hipEvent_t e0 = nullptr;
    C.stage_begin(e0);
    hipLaunchKernelGGL(k_fill_chars, dim3(2048), dim3(256), 0, C.stream, (char *)chars.d, dcap, '0');
    hipLaunchKernelGGL(k_dec_leaf, dim3(blocks_for(leaves.size(), 64)), dim3(64), 0, C.stream,
                       (const DecLeaf *)dl.d, (uint32_t)leaves.size(), v, (char *)chars.d, dcap);
    C.stage_end("decimal_leaf", e0, (double)dcap);
    HIP_TRY(hipGetLastError());
    // the string into page-locked memory (one DMA; the image is recycled for
    // the next conversion), then into `out` in one copy
    HostImg img = C.take_img(dcap);
    HIP_TRY(hipMemcpyAsync(img.data(), chars.d, dcap, hipMemcpyDeviceToHost, C.stream));
    std::vector<std::vector<uint64_t>> fh(checks.size());
    for (size_t c = 0; c < checks.size(); ++c) {
        fh[c].resize(4 * checks[c].n);
        HIP_TRY(hipMemcpyAsync(fh[c].data(), checks[c].flags, fh[c].size() * 8, hipMemcpyDeviceToHost, C.stream));
    }
    if ((st = C.sync())) return st;
    for (const auto &f : fh)
        for (size_t j = 0; j < f.size(); j += 4)
            if (f[j] != 1 || f[j + 2] != 0) return set_error(PG_EINTERNAL, "decimal split check");
    const char *p = img.data();
    size_t first = 0;
    while (first < dcap && p[first] == '0') ++first;
    const char *sig = p + first;
    size_t n = dcap - first;
    if (!point) {
        if (!dst) out.reserve(n);
        put(sig, n);
    } else {
        if (!n) {   // zero: "0" before the padding
            sig = "0";
            n = 1;
        }
        const size_t total = std::max(n, point + 1), zeros = total - n;
        if (!dst) out.reserve(total + 1);
        if (zeros) {   // integer part "0", then zeros - 1 zeros and the digits
            put("0.", 2);
            fill('0', zeros - 1);
            put(sig, n);
        } else {
            put(sig, n - point);
            put(".", 1);
            put(sig + (n - point), point);
        }
    }
    C.hold(std::move(img));
    return done();
}

// ---- decimal parse (precn_t(std::string), src/init_basic.cpp:7-45, 66-92) -------------
// The inverse of to_decimal's tree: the digit string is cut into chunks of 19
// digits from its least significant end, DEC_LEAF chunks to a leaf, and
// leaf / node values combine pairwise, level by level:
//   node_(i+1)[k] = node_i[2k + 1] * U_i + node_i[2k],  U_i = 10^(19 DEC_LEAF 2^i),
// with U_(i+1) = U_i^2.  A node of level i is below U_i, so it fits the
// limbs of U_i; every level is one batched product launch (all nodes times
// the same U_i) and one linear launch for the sums.  The reference combines
// chunks of 9 digits with the same halving recursion (init_basic.cpp:30-45);
// the value is the same.

// One thread per leaf: chunks [DEC_LEAF i, DEC_LEAF (i + 1)) of the string
// (chunk c is digits [n - 19 (c + 1), n - 19 c), the first one possibly
// shorter) by Horner's rule in base 10^19, into out + DEC_LEAF i.
__global__ void __launch_bounds__(64) k_dec_parse_leaf(const char *digits, uint64_t n, uint64_t nchunk,
                                                       uint32_t nleaf, uint64_t *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nleaf) return;
    uint64_t acc[DEC_LEAF];
#pragma unroll
    for (int k = 0; k < DEC_LEAF; ++k) acc[k] = 0;
    const uint64_t c0 = (uint64_t)i * DEC_LEAF;
    for (int c = DEC_LEAF - 1; c >= 0; --c) {
        const uint64_t ch = c0 + c;
        if (ch >= nchunk) continue;
        const uint64_t end = n - 19 * ch, beg = end > 19 ? end - 19 : 0;
        uint64_t v = 0;
        for (uint64_t p = beg; p < end; ++p) v = v * 10 + (uint64_t)(digits[p] - '0');
        // acc = acc * 10^19 + v
        uint64_t carry = v;
#pragma unroll
        for (int k = 0; k < DEC_LEAF; ++k) {
            const unsigned __int128 t = (unsigned __int128)acc[k] * TEN19 + carry;
            acc[k] = (uint64_t)t;
            carry = (uint64_t)(t >> 64);
        }
    }
#pragma unroll
    for (int k = 0; k < DEC_LEAF; ++k) out[c0 + k] = acc[k];
}

int from_decimal(const char *digits, size_t n, std::vector<uint64_t> &out) {
    Context &C = ctx();
    out.clear();
    size_t lead = 0;
    while (lead < n && digits[lead] == '0') ++lead;
    digits += lead;
    n -= lead;
    if (!n) return PG_OK;
    int st;
    const size_t nchunk = (n + 18) / 19, nleaf = (nchunk + DEC_LEAF - 1) / DEC_LEAF;
    int top = 0;
    while (((size_t)1 << top) < nleaf) ++top;
    // U_0 = 10^(19 DEC_LEAF) on the host; its limbs bound a leaf
    std::vector<uint64_t> u0(1, 1);
    for (int k = 0; k < DEC_LEAF; ++k) {
        uint64_t carry = 0;
        for (auto &w : u0) {
            const unsigned __int128 t = (unsigned __int128)w * TEN19 + carry;
            w = (uint64_t)t;
            carry = (uint64_t)(t >> 64);
        }
        if (carry) u0.push_back(carry);
    }
    if (u0.size() != (size_t)DEC_LEAF) return set_error(PG_EINTERNAL, "decimal parse: leaf bound");
    DBuf dstr((n + 7) / 8), leaves(nleaf * DEC_LEAF);
    if (!dstr.d || !leaves.d) return set_error(PG_ENOMEM, "decimal parse");
    {
        HostImg img = C.take_img(n);
        std::memcpy(img.data(), digits, n);
        HIP_TRY(hipMemcpyAsync(dstr.d, img.data(), n, hipMemcpyHostToDevice, C.stream));
        C.hold(std::move(img));
    }
    hipEvent_t e0 = nullptr;
    C.stage_begin(e0);
    hipLaunchKernelGGL(k_dec_parse_leaf, dim3(blocks_for(nleaf, 64)), dim3(64), 0, C.stream, (const char *)dstr.d,
                       (uint64_t)n, (uint64_t)nchunk, (uint32_t)nleaf, leaves.d);
    C.stage_end("decimal_parse_leaf", e0, (double)n + 8.0 * DEC_LEAF * nleaf);
    HIP_TRY(hipGetLastError());
    // node k of the current level: limbs [p, p + len) (len = the level's bound)
    std::vector<const uint64_t *> nodes(nleaf);
    for (size_t k = 0; k < nleaf; ++k) nodes[k] = leaves.d + k * DEC_LEAF;
    size_t len = DEC_LEAF;
    DBuf U(DEC_LEAF);
    if (!U.d || (st = dset(U.d, DEC_LEAF, u0.data(), DEC_LEAF))) return st ? st : set_error(PG_ENOMEM, "decimal parse");
    size_t ul = DEC_LEAF;   // limbs of U_i (its top limb is nonzero)
    std::vector<DBuf> keep;
    for (int i = 0; i < top; ++i) {
        const size_t nm = nodes.size() / 2, np = (nodes.size() + 1) / 2;
        // products hi * U_i (len + ul limbs) and the level's nodes (2 len limbs each)
        const size_t nlen = len + ul;
        keep.emplace_back(nm * nlen + np * nlen);
        if (!keep.back().d) return set_error(PG_ENOMEM, "decimal parse");
        uint64_t *prods = keep.back().d, *lvl = prods + nm * nlen;
        std::vector<pg_job> jobs(nm);
        std::vector<CarryNode> sums(nm);
        std::vector<const uint64_t *> up(np);
        for (size_t k = 0; k < nm; ++k) {
            pg_job &j = jobs[k];
            j = pg_job{};
            j.a = nodes[2 * k + 1]; j.an = len; j.b = U.d; j.bn = ul;
            j.r = prods + k * nlen; j.r_cap = nlen; j.path = PG_PATH_AUTO;
            CarryNode c = lin_node(lvl + k * nlen, nlen);
            c.x = prods + k * nlen; c.xn = nlen; c.y = nodes[2 * k]; c.yn = len;
            sums[k] = c;
            up[k] = lvl + k * nlen;
        }
        if (np > nm) {   // the most significant node has no partner: zero-extend it
            CarryNode c = lin_node(lvl + nm * nlen, nlen);
            c.x = nodes.back(); c.xn = len;
            sums.push_back(c);
            up[nm] = lvl + nm * nlen;
        }
        if (nm) {
            Batch b;
            b.use_graph = false;
            if ((st = b.build(jobs.data(), nm)) || (st = b.run())) return st;
        }
        if ((st = lin_run(sums.data(), sums.size(), "decimal_parse"))) return st;
        nodes.swap(up);
        len = nlen;
        if (i + 1 < top) {   // U_(i+1) = U_i^2
            DBuf U2(2 * ul);
            if (!U2.d) return set_error(PG_ENOMEM, "decimal parse");
            if ((st = dmul(U2.d, U.d, ul, U.d, ul))) return st;
            size_t ul2 = 0;
            if ((st = top_len(U2.d, 2 * ul, &ul2))) return st;
            keep.push_back(std::move(U));
            U = std::move(U2);
            ul = ul2;
        }
    }
    size_t rl = 0;
    if ((st = top_len(nodes[0], len, &rl))) return st;
    out.resize(rl);
    if (rl) HIP_TRY(hipMemcpyAsync(out.data(), nodes[0], rl * 8, hipMemcpyDeviceToHost, C.stream));
    return C.sync();
}

// ---- the pi program ------------------------------------------------------------------

int pi_digits(size_t digits, std::string &out, double phase_ms[4], char *dst, size_t cap, size_t *dlen) {
    out.clear();
    int st;
    auto now = [] { return std::chrono::steady_clock::now(); };
    auto ms = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
        return std::chrono::duration<double, std::milli>(b - a).count();
    };
    const size_t guard = 10, work = digits + guard, terms = work / 14 + 2;   // pi_chudnovsky.cpp:163-165
    auto t0 = now();
    std::vector<DBuf> keep;
    DBuf P, Q, T;
    size_t lp = 0, lq = 0, lt = 0;
    if ((st = chud_split(terms, P, lp, Q, lq, T, lt, false, keep))) return st;
    keep.clear();
    if ((st = ctx().sync())) return st;
    auto t1 = now();
    // s = floor(sqrt(10005 * 10^(2 work)))   (pi_chudnovsky.cpp:109-113, 172-174)
    DBuf pw, s;
    size_t pwl = 0, sl = 0;
    {
        // pow_u32(10, 2 work) (pi_chudnovsky.cpp:62-71): r = r * b on odd bits, b = b * b
        size_t e = 2 * work;
        DBuf r(1), b(1);
        size_t rl = 1, bl = 1;
        const uint64_t one = 1, ten = 10;
        if (!r.d || !b.d || (st = dset(r.d, 1, &one, 1)) || (st = dset(b.d, 1, &ten, 1)))
            return st ? st : set_error(PG_ENOMEM, "power");
        while (e) {
            if (e & 1) {
                DBuf t(rl + bl);
                if (!t.d || (st = dmul(t.d, r.d, rl, b.d, bl)) || (st = top_len(t.d, rl + bl, &rl)))
                    return st ? st : set_error(PG_ENOMEM, "power");
                r = std::move(t);
            }
            e >>= 1;
            if (e) {
                DBuf t(2 * bl);
                if (!t.d || (st = dmul(t.d, b.d, bl, b.d, bl)) || (st = top_len(t.d, 2 * bl, &bl)))
                    return st ? st : set_error(PG_ENOMEM, "power");
                b = std::move(t);
            }
        }
        pwl = rl + 1;
        if (!pw.reset(pwl)) return set_error(PG_ENOMEM, "power");
        CarryNode c = lin_node(pw.d, pwl);   // mul_u32(10^(2 work), 10005)
        c.x = r.d; c.xn = rl; c.xmul = 10005;
        if ((st = lin_one(c, "pi")) || (st = top_len(pw.d, pwl, &pwl))) return st;
        if ((st = isqrt(s, sl, pw.d, pwl))) return st;
    }
    if ((st = ctx().sync())) return st;
    auto t2 = now();
    // pi_scaled = floor(floor(Q 426880 s / T) / 10^10) = floor(Q 426880 s / (T 10^10))
    DBuf num, tq, x;
    size_t xl = 0;
    {
        DBuf q2(lq + 1), t2b(lt + 1);
        if (!q2.d || !t2b.d) return set_error(PG_ENOMEM, "final division");
        CarryNode c[2] = {lin_node(q2.d, lq + 1), lin_node(t2b.d, lt + 1)};
        c[0].x = Q.d; c[0].xn = lq; c[0].xmul = 426880;
        c[1].x = T.d; c[1].xn = lt; c[1].xmul = 10000000000ull;
        if ((st = lin_run(c, 2, "pi"))) return st;
        if (!num.reset(lq + 1 + sl)) return set_error(PG_ENOMEM, "final division");
        if ((st = dmul(num.d, q2.d, lq + 1, s.d, sl))) return st;
        size_t nl = 0, tl = 0;
        if ((st = top_len(num.d, lq + 1 + sl, &nl)) || (st = top_len(t2b.d, lt + 1, &tl))) return st;
        if ((st = div_mulinv(x, xl, nullptr, nullptr, num.d, nl, t2b.d, tl))) return st;
    }
    if ((st = ctx().sync())) return st;
    auto t3 = now();
    size_t n = 0;
    st = to_decimal(x.d, xl, out, digits, dst, cap, &n);
    if (dlen) *dlen = n;
    if (st) return st;
    auto t4 = now();
    if (phase_ms) {
        phase_ms[0] = ms(t0, t1);
        phase_ms[1] = ms(t1, t2);
        phase_ms[2] = ms(t2, t3);
        phase_ms[3] = ms(t3, t4);
    }
    if (!n) {   // to_dec, pi_chudnovsky.cpp:154 (digits == 0 and a zero value)
        if (!dst) {
            out = "0";
        } else if (cap) {
            dst[0] = '0';
        } else {
            if (dlen) *dlen = 1;
            return set_error(PG_EINVAL, "output capacity too small");
        }
        if (dlen) *dlen = 1;
    }
    return PG_OK;
}

}  // namespace pg
