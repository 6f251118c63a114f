// pg_ntt_impl.h -- four-step NTT product kernels for gfx950 (templates; the
// default plans are instantiated in pg_ntt.hip, autotuner variants in
// pg_ntt_v1.hip .. pg_ntt_v15.hip).
//
// Product path of preclib's mul_ntt (src/mul_ntt.cpp:494-518): base-2^16
// digits (331-342), cyclic convolution modulo each prime (349-387), Garner
// CRT (405-412) and base-2^16 carry (438-461).  The transform is organised
// for HBM instead of the reference's radix-2 sweeps:
//
//   n = N1 * N2, digit index j = j1 * N2 + c  (row j1 < N1, column c < N2)
//   k_col_fwd : per tile of T columns, an N1-point DFT down every column,
//               then the four-step twiddle omega_n^(c * k1);      HBM sweep 1
//   k_row     : per row, N2-point DFTs of both operands, pointwise Montgomery
//               product, inverse N2-point DFT;                      HBM sweep 2
//   k_col_inv : per tile, inverse twiddle and 1/n, inverse N1-point DFT for
//               every prime, CRT, and assembly of 4 coefficients into one
//               128-bit limb sum S_j = sum_t c_{4j+t} 2^(16t);      HBM sweep 3
//   carry     : S -> normalized limbs (pg_carry.hip).
//
// Inside a workgroup each N-point DFT is a sequence of radix-R register
// passes over an LDS tile (element (row, col) at lds[row * T + col]); pass i
// with block size B does, for every column j2 < B/R of every block, an R-point
// DIF in registers and multiplies output k1 by omega_B^(j2 * k1) (Shoup
// twiddles from a per-(prime, B) table laid out [k1][j2]).  Outputs end in
// mixed-radix digit-reversed order; the inverse passes run the same network
// backwards, so no permutation pass is ever needed.
//
// Every size parameter is a template constant: one instantiation per
// transform length 2^8 .. 2^26 and plan variant V (pg_kernels.h), so strides
// and offsets fold into immediates.
#pragma once
#include "pg_modarith.h"
#include "pg_kernels.h"

namespace pg {

// LDS element i lives at word i ^ (9 * ((i >> 6) & 7)): within each 64-word
// row (one word per bank of gfx950's 64 banks) the bank is XORed with the row
// number in both 3-bit halves.  Across the radix passes a wave touches 64
// words of the same q-th input at strides 1, 8, 64 and above; the XOR sends
// each of those patterns to 64 distinct banks (32-word padding leaves 2- to
// 4-way conflicts in the passes of stride 8 and in the last pass).
__device__ __forceinline__ int lp(int i) { return i ^ (9 * ((i >> 6) & 7)); }

// Workgroup barrier for LDS traffic: waits for this wave's LDS operations and
// the other waves, not for its outstanding global loads.  Every exchange
// between threads of these kernels goes through LDS, so a prefetched global
// load can stay in flight across the radix passes.
__device__ __forceinline__ void lds_sync() {
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
#else
    lds_sync();
#endif
}

// threadIdx.x through an empty asm: a value the compiler cannot reuse across
// calls, so the index arithmetic of each transform is computed where it is
// used instead of being held in registers from one transform to the next.
__device__ __forceinline__ int fresh_tid() {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

// The workgroup's leaf descriptor, copied into registers once: read through
// a reference, its fields are reloaded by flat loads at every use, and each
// flat load waits for every global store issued before it.
// The fields are uniform: readfirstlane keeps them in scalar registers.
__device__ __forceinline__ NttLeaf leaf_of(const NttLaunch &L) {
    union {
        NttLeaf d;
        int w[sizeof(NttLeaf) / 4];
    } u;
    if (L.leaves) {
        const auto src = gptr((const int *)(L.leaves + blockIdx.z));
#pragma unroll
        for (int i = 0; i < (int)(sizeof(NttLeaf) / 4); ++i) u.w[i] = src[i];
    } else {
        u.d = L.one;
    }
#pragma unroll
    for (int i = 0; i < (int)(sizeof(NttLeaf) / 4); ++i) u.w[i] = __builtin_amdgcn_readfirstlane(u.w[i]);
    return u.d;
}

// A 64-bit value or pointer that is the same in every lane, in scalar
// registers (loads through such a pointer take a scalar base and a 32-bit
// vector offset).
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v), hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (uint64_t)hi << 32 | lo;
}
template <class T>
__device__ __forceinline__ T *uniform_ptr(T *p) {
    return (T *)uniform_u64((uint64_t)p);
}

// 16-byte vector of the compiler's own (loads through an address-space
// pointer need a builtin vector type; HIP's uint4 is a class)
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// ---- in-LDS radix passes --------------------------------------------------

// HALFZ (first pass only): inputs q >= R/2, the upper half of the
// transform, are zero and are neither read nor added.  LZ: lazy residues in
// [0, 2p) (pg_modarith.h), for a prime below 2^30.
template <int LR, int LOGN, int LOGT, int LOGB, int NT, bool HALFZ = false, bool LZ = false>
__device__ __forceinline__ void pass_fwd(uint32_t *lds, const uint2 *__restrict__ tw, const PrimeK &c, int tid) {
    constexpr int R = 1 << LR, LOGM = LOGB - LR, M = 1 << LOGM, T = 1 << LOGT;
    constexpr int NCOLS = 1 << (LOGN - LR + LOGT);
    constexpr int S = M * T;
#pragma unroll 1
    for (int col = tid; col < NCOLS; col += NT) {
        const int cc = col & (T - 1), r = col >> LOGT;
        const int blk = r >> LOGM, j2 = r & (M - 1);
        const int base = (((blk << LOGB) + j2) << LOGT) + cc;
        uint32_t x[R];
#pragma unroll
        for (int q = 0; q < (HALFZ ? R / 2 : R); ++q) x[q] = lds[lp(base + q * S)];
        dft_dif<R, HALFZ, LZ, (M > 1)>(x, c);
        if constexpr (M > 1) {
            const uint2 *t = tw + j2;
#pragma unroll
            for (int q = 1; q < R; ++q) {
                const uint2 w = t[bitrev<R>(q) * M];
                x[q] = mul_shoup<LZ>(x[q], w.x, w.y, c);
            }
        }
#pragma unroll
        for (int q = 0; q < R; ++q) lds[lp(base + bitrev<R>(q) * S)] = x[q];
    }
}

// pass_fwd over two tiles of the same shape (the row kernel's two operands):
// each column's twiddles are loaded once and used for both.
template <int LR, int LOGN, int LOGT, int LOGB, int NT, bool LZ = false>
__device__ __forceinline__ void pass_fwd2(uint32_t *lda, uint32_t *ldb, const uint2 *__restrict__ tw, const PrimeK &c,
                                          int tid) {
    constexpr int R = 1 << LR, LOGM = LOGB - LR, M = 1 << LOGM, T = 1 << LOGT;
    constexpr int NCOLS = 1 << (LOGN - LR + LOGT);
    constexpr int S = M * T;
#pragma unroll 1
    for (int col = tid; col < NCOLS; col += NT) {
        const int cc = col & (T - 1), r = col >> LOGT;
        const int blk = r >> LOGM, j2 = r & (M - 1);
        const int base = (((blk << LOGB) + j2) << LOGT) + cc;
        uint2 w[R];
        if constexpr (M > 1) {
            const uint2 *t = tw + j2;
#pragma unroll
            for (int q = 1; q < R; ++q) w[q] = t[bitrev<R>(q) * M];
        }
#pragma unroll
        for (int op = 0; op < 2; ++op) {
            uint32_t *lds = op ? ldb : lda;
            uint32_t x[R];
#pragma unroll
            for (int q = 0; q < R; ++q) x[q] = lds[lp(base + q * S)];
            dft_dif<R, false, LZ, (M > 1)>(x, c);
            if constexpr (M > 1) {
#pragma unroll
                for (int q = 1; q < R; ++q) x[q] = mul_shoup<LZ>(x[q], w[q].x, w[q].y, c);
            }
#pragma unroll
            for (int q = 0; q < R; ++q) lds[lp(base + bitrev<R>(q) * S)] = x[q];
        }
    }
}

template <int LR, int LOGN, int LOGT, int LOGB, int NT, bool LZ = false>
__device__ __forceinline__ void pass_inv(uint32_t *lds, const uint2 *__restrict__ itw, const PrimeK &c, int tid) {
    constexpr int R = 1 << LR, LOGM = LOGB - LR, M = 1 << LOGM, T = 1 << LOGT;
    constexpr int NCOLS = 1 << (LOGN - LR + LOGT);
    constexpr int S = M * T;
#pragma unroll 1
    for (int col = tid; col < NCOLS; col += NT) {
        const int cc = col & (T - 1), r = col >> LOGT;
        const int blk = r >> LOGM, j2 = r & (M - 1);
        const int base = (((blk << LOGB) + j2) << LOGT) + cc;
        uint32_t x[R];
#pragma unroll
        for (int q = 0; q < R; ++q) x[q] = lds[lp(base + bitrev<R>(q) * S)];
        if constexpr (M > 1) {
            const uint2 *t = itw + j2;
#pragma unroll
            for (int q = 1; q < R; ++q) {
                const uint2 w = t[bitrev<R>(q) * M];
                x[q] = mul_shoup<LZ>(x[q], w.x, w.y, c);
            }
        }
        dft_dit_inv<R, LZ>(x, c);
#pragma unroll
        for (int q = 0; q < R; ++q) lds[lp(base + q * S)] = x[q];
    }
}

// a * w mod p for a twiddle w held in Montgomery form (w 2^32 mod p), a < 2^32:
// one Montgomery product; LZ leaves the result in [0, 2p) (then a < 2p is
// required, so that a w < 2 p^2 < p 2^32).
template <bool LZ>
__device__ __forceinline__ uint32_t mul_mont_tw(uint32_t a, uint32_t wm, const PrimeK &c) {
    const uint64_t t = (uint64_t)a * wm;
    const uint32_t m = (uint32_t)t * c.ninv;
    const uint32_t r = (uint32_t)((t + (uint64_t)m * c.p) >> 32);   // [0, 2p)
    if constexpr (LZ) return r;
    else return min(r, r - c.p);
}

// pass_inv with the pass's twiddles in LDS, in Montgomery form, laid out
// [j2][k1] (R words per j2, k1 = 0 unused; fill_inv_tw), read 16 bytes at a
// time.
template <int LR, int LOGN, int LOGT, int LOGB, int NT, bool LZ>
__device__ __forceinline__ void pass_inv_lt(uint32_t *lds, const uint32_t *tw, const PrimeK &c, int tid) {
    constexpr int R = 1 << LR, LOGM = LOGB - LR, M = 1 << LOGM, T = 1 << LOGT;
    constexpr int NCOLS = 1 << (LOGN - LR + LOGT);
    constexpr int S = M * T;
    static_assert(M == 1 || R >= 4, "pass_inv_lt: 16-byte twiddle rows");
#pragma unroll 1
    for (int col = tid; col < NCOLS; col += NT) {
        const int cc = col & (T - 1), r = col >> LOGT;
        const int blk = r >> LOGM, j2 = r & (M - 1);
        const int base = (((blk << LOGB) + j2) << LOGT) + cc;
        uint32_t x[R];
#pragma unroll
        for (int q = 0; q < R; ++q) x[q] = lds[lp(base + bitrev<R>(q) * S)];
        if constexpr (M > 1) {
            uint32_t w[R];
#pragma unroll
            for (int k = 0; k < R; k += 4) {
                const u32x4 v = *(const u32x4 *)(tw + j2 * R + k);
                w[k] = v.x; w[k + 1] = v.y; w[k + 2] = v.z; w[k + 3] = v.w;
            }
#pragma unroll
            for (int q = 1; q < R; ++q) x[q] = mul_mont_tw<LZ>(x[q], w[bitrev<R>(q)], c);
        }
        dft_dit_inv<R, LZ>(x, c);
#pragma unroll
        for (int q = 0; q < R; ++q) lds[lp(base + q * S)] = x[q];
    }
}

// LDS words of the twiddles of inverse passes I, I + 1, .. of a 2^LOGN
// transform whose pass I has block size 2^LOGB (passes with M = 1 have none).
constexpr int inv_tw_words(int logn, int rb, int i, int logb) {
    return split_lr(logn, i, rb) <= 0 ? 0
           : ((logb - split_lr(logn, i, rb)) > 0 ? 1 << logb : 0) +
                 inv_tw_words(logn, rb, i + 1, logb - split_lr(logn, i, rb));
}

// Inverse passes I .. in LDS, their twiddles at tw + OFF, pass by pass.
template <int LOGN, int LOGT, int NT, int LOGB, int I, int RB, bool LZ, int OFF>
__device__ __forceinline__ void inv_from_lt(uint32_t *lds, const uint32_t *tw, const PrimeK &c, int tid) {
    constexpr int LR = split_lr(LOGN, I, RB);
    if constexpr (LR > 0) {
        constexpr int M = 1 << (LOGB - LR);
        inv_from_lt<LOGN, LOGT, NT, LOGB - LR, I + 1, RB, LZ, OFF + (M > 1 ? (1 << LOGB) : 0)>(lds, tw, c, tid);
        pass_inv_lt<LR, LOGN, LOGT, LOGB, NT, LZ>(lds, tw + OFF, c, tid);
        lds_sync();
    }
}

// Copies the twiddles of inverse passes I .. of one prime from their Shoup
// tables (itw[I] laid out [k1][j2], (w, floor(w 2^32 / p))) into LDS in
// Montgomery form, w 2^32 mod p = -floor(w 2^32 / p) p mod 2^32 = wq (2^32 - p),
// laid out [j2][k1] (entry k1 + R j2 from table entry k1 M + j2).
template <int LOGN, int NT, int LOGB, int I, int RB, int OFF>
__device__ __forceinline__ void fill_inv_tw(uint32_t *tw, const uint2 *const *itw, const PrimeK &c, int tid) {
    constexpr int LR = split_lr(LOGN, I, RB);
    if constexpr (LR > 0) {
        constexpr int R = 1 << LR, M = 1 << (LOGB - LR), W = M > 1 ? R * M : 0;
        if constexpr (W > 0) {
            // every load of the thread issued before the first store
            constexpr int K = (W + NT - 1) / NT;
            const auto src = gptr(itw[I]);
            uint32_t q[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int i = tid + k * NT;   // LDS entry k1 + R j2
                q[k] = (W % NT == 0 || i < W) ? src[(i & (R - 1)) * M + i / R].y : 0;
            }
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (W % NT == 0 || tid + k * NT < W) tw[OFF + tid + k * NT] = q[k] * c.np;
        }
        fill_inv_tw<LOGN, NT, LOGB - LR, I + 1, RB, OFF + W>(tw, itw, c, tid);
    }
}

// Pass lists: forward runs passes 0..np-1 with shrinking block size; the
// inverse runs them backwards (post-order recursion).  Passes I >= STOP are
// left out (the row kernel runs its last pass in registers).
template <int LOGN, int LOGT, int NT, int LOGB, int I, int RB, int STOP, bool LZ = false>
__device__ __forceinline__ void fwd_from(uint32_t *lds, const uint2 *const *tw, const PrimeK &c, int tid) {
    constexpr int LR = split_lr(LOGN, I, RB);
    if constexpr (LR > 0 && I < STOP) {
        pass_fwd<LR, LOGN, LOGT, LOGB, NT, false, LZ>(lds, tw[I], c, tid);
        lds_sync();
        fwd_from<LOGN, LOGT, NT, LOGB - LR, I + 1, RB, STOP, LZ>(lds, tw, c, tid);
    }
}

template <int LOGN, int LOGT, int NT, int LOGB, int I, int RB, int STOP, bool LZ = false>
__device__ __forceinline__ void fwd2_from(uint32_t *lda, uint32_t *ldb, const uint2 *const *tw, const PrimeK &c,
                                          int tid) {
    constexpr int LR = split_lr(LOGN, I, RB);
    if constexpr (LR > 0 && I < STOP) {
        pass_fwd2<LR, LOGN, LOGT, LOGB, NT, LZ>(lda, ldb, tw[I], c, tid);
        lds_sync();
        fwd2_from<LOGN, LOGT, NT, LOGB - LR, I + 1, RB, STOP, LZ>(lda, ldb, tw, c, tid);
    }
}

template <int LOGN, int LOGT, int NT, int LOGB, int I, int RB, int STOP, bool LZ = false>
__device__ __forceinline__ void inv_from(uint32_t *lds, const uint2 *const *itw, const PrimeK &c, int tid) {
    constexpr int LR = split_lr(LOGN, I, RB);
    if constexpr (LR > 0 && I < STOP) {
        inv_from<LOGN, LOGT, NT, LOGB - LR, I + 1, RB, STOP, LZ>(lds, itw, c, tid);
        pass_inv<LR, LOGN, LOGT, LOGB, NT, LZ>(lds, itw[I], c, tid);
        lds_sync();
    }
}

// tid: the thread index (callers may pass an opaque copy of threadIdx.x so
// that index arithmetic is recomputed per call rather than kept live).
template <int LOGN, int LOGT, int NT, int RB, int STOP = MAX_PASS, bool LZ = false>
__device__ __forceinline__ void lds_forward(uint32_t *lds, const uint2 *const *tw, const PrimeK &c,
                                            int tid = threadIdx.x) {
    fwd_from<LOGN, LOGT, NT, LOGN, 0, RB, STOP, LZ>(lds, tw, c, tid);
}

template <int LOGN, int LOGT, int NT, int RB, int STOP = MAX_PASS, bool LZ = false>
__device__ __forceinline__ void lds_inverse(uint32_t *lds, const uint2 *const *itw, const PrimeK &c,
                                            int tid = threadIdx.x) {
    inv_from<LOGN, LOGT, NT, LOGN, 0, RB, STOP, LZ>(lds, itw, c, tid);
}

// Twiddles of one inverse pass for the columns a thread handles, loaded
// into registers ahead of the pass (before the barrier that precedes it), so
// the pass starts on its products instead of waiting for the loads.
template <int LR, int LOGN, int LOGT, int LOGB, int NT>
struct PassTw {
    static constexpr int R = 1 << LR, LOGM = LOGB - LR, M = 1 << LOGM, T = 1 << LOGT;
    static constexpr int NCOLS = 1 << (LOGN - LR + LOGT), IT = (NCOLS + NT - 1) / NT;
    uint2 w[M > 1 ? IT : 1][M > 1 ? R : 1];
    __device__ __forceinline__ void load(const uint2 *__restrict__ itw, int tid) {
        if constexpr (M > 1) {
#pragma unroll
            for (int it = 0; it < IT; ++it) {
                const int col = tid + it * NT;
                if (col < NCOLS) {
                    const uint2 *t = itw + ((col >> LOGT) & (M - 1));
#pragma unroll
                    for (int q = 1; q < R; ++q) w[it][q] = t[bitrev<R>(q) * M];
                }
            }
        }
    }
};

// pass_inv with the pass's twiddles already in registers.
template <int LR, int LOGN, int LOGT, int LOGB, int NT, bool LZ>
__device__ __forceinline__ void pass_inv_pre(uint32_t *lds, const PassTw<LR, LOGN, LOGT, LOGB, NT> &tw,
                                             const PrimeK &c, int tid) {
    using TW = PassTw<LR, LOGN, LOGT, LOGB, NT>;
    constexpr int R = TW::R, LOGM = TW::LOGM, M = TW::M, T = TW::T, S = M * T;
#pragma unroll
    for (int it = 0; it < TW::IT; ++it) {
        const int col = tid + it * NT;
        if (col < TW::NCOLS) {
            const int cc = col & (T - 1), r = col >> LOGT;
            const int blk = r >> LOGM, j2 = r & (M - 1);
            const int base = (((blk << LOGB) + j2) << LOGT) + cc;
            uint32_t x[R];
#pragma unroll
            for (int q = 0; q < R; ++q) x[q] = lds[lp(base + bitrev<R>(q) * S)];
            if constexpr (M > 1) {
#pragma unroll
                for (int q = 1; q < R; ++q) x[q] = mul_shoup<LZ>(x[q], tw.w[it][q].x, tw.w[it][q].y, c);
            }
            dft_dit_inv<R, LZ>(x, c);
#pragma unroll
            for (int q = 0; q < R; ++q) lds[lp(base + q * S)] = x[q];
        }
    }
}

// pass_fwd2 with the pass's twiddles already in registers.
template <int LR, int LOGN, int LOGT, int LOGB, int NT, bool LZ>
__device__ __forceinline__ void pass_fwd2_pre(uint32_t *lda, uint32_t *ldb, const PassTw<LR, LOGN, LOGT, LOGB, NT> &tw,
                                              const PrimeK &c, int tid) {
    using TW = PassTw<LR, LOGN, LOGT, LOGB, NT>;
    constexpr int R = TW::R, LOGM = TW::LOGM, M = TW::M, T = TW::T, S = M * T;
#pragma unroll
    for (int it = 0; it < TW::IT; ++it) {
        const int col = tid + it * NT;
        if (col < TW::NCOLS) {
            const int cc = col & (T - 1), r = col >> LOGT;
            const int blk = r >> LOGM, j2 = r & (M - 1);
            const int base = (((blk << LOGB) + j2) << LOGT) + cc;
#pragma unroll
            for (int op = 0; op < 2; ++op) {
                uint32_t *lds = op ? ldb : lda;
                uint32_t x[R];
#pragma unroll
                for (int q = 0; q < R; ++q) x[q] = lds[lp(base + q * S)];
                dft_dif<R, false, LZ, (M > 1)>(x, c);
                if constexpr (M > 1) {
#pragma unroll
                    for (int q = 1; q < R; ++q) x[q] = mul_shoup<LZ>(x[q], tw.w[it][q].x, tw.w[it][q].y, c);
                }
#pragma unroll
                for (int q = 0; q < R; ++q) lds[lp(base + bitrev<R>(q) * S)] = x[q];
            }
        }
    }
}

// Block size (log2) of forward pass I of a 2^LOGN transform.
template <int LOGN, int I, int RB>
constexpr int pass_logb() {
    int b = LOGN;
    for (int k = 0; k < I; ++k) b -= split_lr(LOGN, k, RB);
    return b;
}
template <int LOGN, int LOGT, int NT, int I, int RB>
using PassTwOf = PassTw<split_lr(LOGN, I, RB), LOGN, LOGT, pass_logb<LOGN, I, RB>(), NT>;

// Inverse passes I, I - 1, ..., 0 (each followed by a barrier), pass I with
// its twiddles in `cur`; each next pass's twiddles are loaded before the
// barrier that precedes it.
template <int LOGN, int LOGT, int NT, int I, int RB, bool LZ>
__device__ __forceinline__ void inv_pre_from(uint32_t *lds, const uint2 *const *itw, const PrimeK &c, int tid,
                                             const PassTwOf<LOGN, LOGT, NT, I, RB> &cur) {
    pass_inv_pre<split_lr(LOGN, I, RB), LOGN, LOGT, pass_logb<LOGN, I, RB>(), NT, LZ>(lds, cur, c, tid);
    if constexpr (I > 0) {
        PassTwOf<LOGN, LOGT, NT, I - 1, RB> nxt;
        nxt.load(itw[I - 1], tid);
        lds_sync();
        inv_pre_from<LOGN, LOGT, NT, I - 1, RB, LZ>(lds, itw, c, tid, nxt);
    } else {
        lds_sync();
    }
}

// Forward passes I .. STOP - 1 of both tiles (each followed by a barrier),
// pass I with its twiddles in `cur`; each next pass's twiddles are loaded
// before the barrier that precedes it.
template <int LOGN, int LOGT, int NT, int I, int RB, int STOP, bool LZ>
__device__ __forceinline__ void fwd2_pre_from(uint32_t *lda, uint32_t *ldb, const uint2 *const *tw, const PrimeK &c,
                                              int tid, const PassTwOf<LOGN, LOGT, NT, I, RB> &cur) {
    pass_fwd2_pre<split_lr(LOGN, I, RB), LOGN, LOGT, pass_logb<LOGN, I, RB>(), NT, LZ>(lda, ldb, cur, c, tid);
    if constexpr (I + 1 < STOP) {
        PassTwOf<LOGN, LOGT, NT, I + 1, RB> nxt;
        nxt.load(tw[I + 1], tid);
        lds_sync();
        fwd2_pre_from<LOGN, LOGT, NT, I + 1, RB, STOP, LZ>(lda, ldb, tw, c, tid, nxt);
    } else {
        lds_sync();
    }
}

// Whether every inverse pass below STOP keeps at most 15 twiddle pairs per
// thread (the register budget of the prefetch).
template <int LOGN, int LOGT, int NT, int RB, int STOP>
constexpr bool inv_pre_fits() {
    for (int i = 0; i < STOP; ++i) {
        const int lr = split_lr(LOGN, i, RB);
        const int ncols = 1 << (LOGN - lr + LOGT);
        if (lr <= 0 || ((ncols + NT - 1) / NT) * ((1 << lr) - 1) > 15) return false;
    }
    return true;
}

// lds_sync(); fwd2_from<LOGN, LOGT, NT, LOGN, 0, RB, STOP, LZ>(...), with the
// first pass's twiddles loaded before that barrier and each later pass's
// before its own (where they fit; otherwise exactly those two calls).
// PRE = false (a kernel held to 64 VGPRs) also takes the plain form.
template <int LOGN, int LOGT, int NT, int RB, int STOP, bool LZ, bool PRE>
__device__ __forceinline__ void sync_fwd2_pre(uint32_t *lda, uint32_t *ldb, const uint2 *const *tw, const PrimeK &c,
                                              int tid) {
    if constexpr (!PRE || !inv_pre_fits<LOGN, LOGT, NT, RB, STOP>()) {
        lds_sync();
        fwd2_from<LOGN, LOGT, NT, LOGN, 0, RB, STOP, LZ>(lda, ldb, tw, c, tid);
    } else if constexpr (STOP >= 1) {
        PassTwOf<LOGN, LOGT, NT, 0, RB> first;
        first.load(tw[0], tid);
        lds_sync();
        fwd2_pre_from<LOGN, LOGT, NT, 0, RB, STOP, LZ>(lda, ldb, tw, c, tid, first);
    } else {
        lds_sync();
    }
}

// lds_sync(); lds_inverse<LOGN, LOGT, NT, RB, STOP, LZ>(...), with the first
// inverse pass's twiddles loaded before that barrier and each later pass's
// before its own (where they fit; otherwise exactly those two calls).
template <int LOGN, int LOGT, int NT, int RB, int STOP, bool LZ, bool PRE>
__device__ __forceinline__ void sync_inverse_pre(uint32_t *lds, const uint2 *const *itw, const PrimeK &c, int tid) {
    if constexpr (!PRE || !inv_pre_fits<LOGN, LOGT, NT, RB, STOP>()) {
        lds_sync();
        lds_inverse<LOGN, LOGT, NT, RB, STOP, LZ>(lds, itw, c, tid);
    } else if constexpr (STOP >= 1) {
        PassTwOf<LOGN, LOGT, NT, STOP - 1, RB> first;
        first.load(itw[STOP - 1], tid);
        lds_sync();
        inv_pre_from<LOGN, LOGT, NT, STOP - 1, RB, LZ>(lds, itw, c, tid, first);
    } else {
        lds_sync();
    }
}

// Frequency index held at position rho after the forward passes of a 2^LOGN
// transform (mixed-radix digit reversal).
template <int LOGN, int RB>
__device__ __forceinline__ uint32_t freq_of_pos(uint32_t rho) {
    uint32_t k = 0;
    int w = 0, logB = LOGN;
#pragma unroll
    for (int i = 0; i < split_np(LOGN, RB); ++i) {
        const int lr = split_lr(LOGN, i, RB);
        logB -= lr;
        k |= ((rho >> logB) & ((1u << lr) - 1)) << w;
        w += lr;
    }
    return k;
}

__device__ __forceinline__ uint32_t twiddle_mont(const uint32_t *__restrict__ lo,
                                                 const uint32_t *__restrict__ hi, uint32_t e,
                                                 int lobits, const PrimeK &c) {
    return mont_mul(lo[e & ((1u << lobits) - 1)], hi[e >> lobits], c.p, c.ninv);
}

// Four-step twiddles of a column tile.  Thread t holds elements
// e = t + k * NT (k < E) of the tile: row rho = e >> LT, column
// C = c0 + (t mod T), twiddle omega_n^(C * freq(rho)).  When the rows of one
// thread differ only in the top radix digit (the first pass's), freq(rho)
// steps by 2^D from one element to the next, so each twiddle is the previous
// one times omega_n^(C * 2^D): one Montgomery product per element instead of
// forming each from the tables.  Otherwise every twiddle is formed directly.
template <int L1, int LT, int RBV, int NT>
struct ColTw {
    static constexpr int LNT = clog2(NT);
    static constexpr int E = ((1 << (L1 + LT)) + NT - 1) / NT;
    static constexpr int LR0 = split_lr(L1, 0, RBV);
    static constexpr bool STEP = LNT > 0 && E > 1 && (LNT - LT) >= (L1 - LR0);
    static constexpr int D = STEP ? (LNT - LT) - (L1 - LR0) : 0;
};

// Workgroups are dealt to the 8 XCDs round-robin by dispatch order; mapping
// workgroup i to tile (i % 8) * (ntiles / 8) + i / 8 keeps neighbouring tiles
// (which share 128-byte lines of every row) on one XCD's L2.
__device__ __forceinline__ int xcd_tile(int i, int ntiles) {
    return (ntiles & 7) ? i : (i & 7) * (ntiles >> 3) + (i >> 3);
}

// Column-kernel workgroup i of ntiles * np: tile and prime.  Consecutive
// workgroups of one XCD (i, i + 8, i + 16, ...) take the primes of one tile,
// then the next tile; each XCD covers a contiguous run of tiles.
__device__ __forceinline__ void xcd_tile_prime(int i, int ntiles, int np, int &tile, int &pr) {
    if (ntiles & 7) {
        tile = i / np;
        pr = i % np;
        return;
    }
    const int slot = i >> 3;
    pr = slot % np;
    tile = (i & 7) * (ntiles >> 3) + slot / np;
}

// x mod p for any x < 2^32 (Shoup's product by 1 with quotient floor(2^32/p)).
__device__ __forceinline__ uint32_t reduce32(uint32_t x, uint32_t red, uint32_t p) {
    const uint32_t r = x - __umulhi(x, red) * p;   // [0, 2p)
    return min(r, r - p);
}

// The digits of one limb, as residues mod the transform's prime, into tile
// elements i, i + 1, ...
template <int MODE>
__device__ __forceinline__ void put_digits(uint32_t *lds, int i, uint64_t v, uint32_t red, uint32_t p) {
    if constexpr (MODE == MODE_REF) {
        lds[lp(i)] = (uint32_t)(v & 0xFFFF);
        lds[lp(i + 1)] = (uint32_t)((v >> 16) & 0xFFFF);
        lds[lp(i + 2)] = (uint32_t)((v >> 32) & 0xFFFF);
        lds[lp(i + 3)] = (uint32_t)(v >> 48);
    } else {
        lds[lp(i)] = reduce32((uint32_t)v, red, p);
        lds[lp(i + 1)] = reduce32((uint32_t)(v >> 32), red, p);
    }
}

// Word offset of row j1, column c in a forward spectrum: rows in groups of
// G = 2^plan_gl, and within a group each tile of T columns stored as a G x T
// block (plan_gl = 0: row-major).
template <int LOGN, int V>
__device__ __forceinline__ size_t spec_off(uint32_t j1, uint32_t c) {
    constexpr int L2 = plan_l2(LOGN, V), LT = plan_lt(LOGN, V), GL = plan_gl(LOGN, V);
    return ((size_t)(j1 >> GL) << (L2 + GL)) + ((size_t)(c >> LT) << (LT + GL)) +
           ((j1 & ((1u << GL) - 1)) << LT) + (c & ((1u << LT) - 1));
}

// Forward column spectra of operand op, prime pr, of a leaf (see NttLeaf::ext).
template <int LOGN, int NP>
__device__ __forceinline__ const uint32_t *op_spec(const NttLeaf &D, int op, int pr) {
    const uint32_t *x = op ? D.ext[1] : D.ext[0];   // (a select: an indexed read would put D in scratch)
    return x ? x + ((size_t)pr << LOGN) : D.spec + ((size_t)(op * NP + pr) << LOGN);
}

// ---- kernel 1: forward column pass ------------------------------------------

// Column transforms of a staged tile (half: the upper half of every column
// is zero).
template <int L1, int LT, int RBV, int NT, bool LZ>
__device__ __forceinline__ void col_forward(uint32_t *lds, const uint2 *const *tw, const PrimeK &c, bool half) {
    constexpr int LR0 = split_lr(L1, 0, RBV);
    if (half) {
        pass_fwd<LR0, L1, LT, L1, NT, true, LZ>(lds, tw[0], c, threadIdx.x);
        lds_sync();
        fwd_from<L1, LT, NT, L1 - LR0, 1, RBV, MAX_PASS, LZ>(lds, tw, c, threadIdx.x);
    } else {
        lds_forward<L1, LT, NT, RBV, MAX_PASS, LZ>(lds, tw, c);
    }
}

// One workgroup per (tile, prime, operand), in the order L.order selects
// (COLF_*): with COLF_TILE_PRIME the workgroups of one tile's primes are dealt
// to the same XCD one after another (xcd_tile_prime), so the operand limbs
// that the first reads from HBM the others find in L2.
template <int LOGN, int MODE, int V>
__global__ void __launch_bounds__(col_threads(LOGN, V)) __attribute__((amdgpu_waves_per_eu(8))) k_col_fwd(const NttLaunch L) {
    constexpr int L1 = plan_l1(LOGN, V), L2 = plan_l2(LOGN, V), LT = plan_lt(LOGN, V), RBV = plan_rb(LOGN, V);
    constexpr int DPL = 64 / mode_digit_bits(MODE);   // digits per limb
    constexpr int N1 = 1 << L1, T = 1 << LT, LPR = T / DPL;
    constexpr uint32_t NMASK = (1u << LOGN) - 1;
    constexpr int NP = mode_primes(MODE);
    constexpr int NT = col_threads(LOGN, V);
    static_assert(NT >= T, "k_col_fwd: one column per thread");
    using TW = ColTw<L1, LT, RBV, NT>;
    extern __shared__ uint32_t lds[];
    const NttLeaf D = leaf_of(L);
    int tile, pr, op;
    switch (L.order & 3) {
    case COLF_OP_PRIME:
        tile = xcd_tile(blockIdx.x, gridDim.x);
        op = blockIdx.y / NP;
        pr = blockIdx.y % NP;
        break;
    case COLF_TILE_PRIME:
        xcd_tile_prime(blockIdx.x, gridDim.x / NP, NP, tile, pr);
        op = blockIdx.y;
        break;
    default: {
        const int t2 = xcd_tile(blockIdx.x, gridDim.x);
        tile = t2 >> 1;
        op = t2 & 1;
        pr = (L.order & 3) == COLF_TILE_OP ? blockIdx.y : NP - 1 - blockIdx.y;
    }
    }
    if (op ? D.ext[1] : D.ext[0]) return;   // transformed by another leaf (or operand) of this launch
    const uint32_t c0 = (uint32_t)tile << LT;
    const uint64_t *src = op ? D.b : D.a;
    const size_t nl = op ? D.bn : D.an;
    const uint32_t C = c0 + (threadIdx.x & (T - 1));
    // An operand of at most n/2 digits leaves rows N1/2 .. N1 of every column
    // zero: they are not staged and the first pass skips them.
    constexpr int LR0 = split_lr(L1, 0, RBV);
    const bool half = L1 > 0 && LR0 > 0 && (size_t)nl * DPL <= ((size_t)1 << (LOGN - 1));
    {
        const PrimeK &c = L.pk[pr];
        // digits of the tile: row j1 holds digits j1*N2 + c0 + [0, T).  (Loads
        // grouped ahead of their LDS writes measured slower here: 32 waves per
        // CU already hide the latency, and groups of 4 spill at 64 VGPRs.)
        const int ne = (half ? N1 / 2 : N1) * LPR;
        for (int e = threadIdx.x; e < ne; e += NT) {
            const int j1 = e / LPR, m = e % LPR;
            const size_t li = (((size_t)j1 << L2) + c0 + DPL * m) / DPL;
            put_digits<MODE>(lds, j1 * T + DPL * m, li < nl ? src[li] : 0, L.crt.red[pr], c.p);
        }
        lds_sync();
        if constexpr (LR0 > 0) {
            if (pr == LAZY_PRIME) col_forward<L1, LT, RBV, NT, true>(lds, L.col_tw[pr], c, half);
            else col_forward<L1, LT, RBV, NT, false>(lds, L.col_tw[pr], c, half);
        }
        // (global stores here instead of flat ones measured slower: 282 -> 291
        // us at 2^24 points)
        uint32_t *out = D.spec + ((size_t)(op * NP + pr) << LOGN);
        if constexpr (TW::STEP) {
            const uint32_t f0 = freq_of_pos<L1, RBV>(threadIdx.x >> LT);
            uint32_t w = twiddle_mont(L.glo[pr], L.ghi[pr], (C * f0) & NMASK, L.lobits, c);
            const uint32_t st = twiddle_mont(L.glo[pr], L.ghi[pr], (C << TW::D) & NMASK, L.lobits, c);
#pragma unroll 4
            for (int k = 0; k < TW::E; ++k) {
                const int e = threadIdx.x + k * NT;
                out[spec_off<LOGN, V>(e >> LT, C)] = mont_mul(lds[lp(e)], w, c.p, c.ninv);
                w = mont_mul(w, st, c.p, c.ninv);
            }
        } else {
            for (int e = threadIdx.x; e < N1 * T; e += NT) {
                const int rho = e >> LT, cc = e & (T - 1);
                const uint32_t ex = ((c0 + cc) * freq_of_pos<L1, RBV>(rho)) & NMASK;
                uint32_t x = lds[lp(e)];
                if (ex) x = mont_mul(x, twiddle_mont(L.glo[pr], L.ghi[pr], ex, L.lobits, c), c.p, c.ninv);
                out[spec_off<LOGN, V>(rho, c0 + cc)] = x;
            }
        }
    }
}

// ---- kernel 2: row pass (forward A, forward B, pointwise, inverse) -------------

// Operand rows op0 .. op0 + NOPS - 1 of the row kernel into lds + k * N2.
// Every thread issues all its loads (NOPS * N2 / 4 / ROW_THREADS of 16
// bytes) before the first LDS write, so a workgroup waits for memory once.
template <int LOGN, int MODE, int V, int NOPS>
__device__ __forceinline__ void load_rows(uint32_t *lds, const NttLaunch &L, const NttLeaf &D, int op0, int pr,
                                          int row) {
    constexpr int L1 = plan_l1(LOGN, V), L2 = plan_l2(LOGN, V), N2 = 1 << L2, ROW_THREADS = row_threads(LOGN, V);
    static_assert(L1 == 0 || plan_lt(LOGN, V) >= 2, "16-byte row loads need tiles of 4 or more columns");
    if constexpr (L1 == 0) {
        // single-row plan: the row is the digit vector itself
        constexpr int DPL = 64 / mode_digit_bits(MODE);
#pragma unroll
        for (int k = 0; k < NOPS; ++k) {
            const int op = op0 + k;
            const uint64_t *src = op ? D.b : D.a;
            const size_t nl = op ? D.bn : D.an;
            for (int m = threadIdx.x; m < N2 / DPL; m += ROW_THREADS)
                put_digits<MODE>(lds + k * N2, DPL * m, (size_t)m < nl ? src[m] : 0, L.crt.red[pr], L.pk[pr].p);
        }
    } else if constexpr (NOPS == 1 && (N2 / 4 + ROW_THREADS - 1) / ROW_THREADS > 4) {
        // the longest single rows: one load at a time (that kernel has no
        // registers to spare)
        const uint4 *in = (const uint4 *)op_spec<LOGN, mode_primes(MODE)>(D, op0, pr);
        for (int m = threadIdx.x; m < N2 / 4; m += ROW_THREADS) {
            const uint4 v = in[spec_off<LOGN, V>(row, 4 * m) >> 2];
            lds[lp(4 * m)] = v.x; lds[lp(4 * m + 1)] = v.y; lds[lp(4 * m + 2)] = v.z; lds[lp(4 * m + 3)] = v.w;
        }
    } else {
        // loads of 16 bytes, up to 4 per operand row in flight per thread
        // (E4 is a power of two)
        constexpr int E4 = (N2 / 4 + ROW_THREADS - 1) / ROW_THREADS;
        constexpr int GR = E4 < 4 ? E4 : 4;
#pragma unroll
        for (int i0 = 0; i0 < E4; i0 += GR) {
            u32x4 v[NOPS][GR];
#pragma unroll
            for (int k = 0; k < NOPS; ++k) {
                const auto in = gptr((const u32x4 *)op_spec<LOGN, mode_primes(MODE)>(D, op0 + k, pr));
#pragma unroll
                for (int i = 0; i < GR; ++i) {
                    const int m = threadIdx.x + (i0 + i) * ROW_THREADS;
                    if (m < N2 / 4) v[k][i] = in[spec_off<LOGN, V>(row, 4 * m) >> 2];
                }
            }
#pragma unroll
            for (int k = 0; k < NOPS; ++k) {
                uint32_t *l = lds + k * N2;
#pragma unroll
                for (int i = 0; i < GR; ++i) {
                    const int m = threadIdx.x + (i0 + i) * ROW_THREADS;
                    if (m < N2 / 4) {
                        l[lp(4 * m)] = v[k][i].x; l[lp(4 * m + 1)] = v[k][i].y;
                        l[lp(4 * m + 2)] = v[k][i].z; l[lp(4 * m + 3)] = v[k][i].w;
                    }
                }
            }
        }
    }
}

template <int LOGN, int MODE, int V, bool LZ>
__device__ __forceinline__ void row_body(const NttLaunch &L, const NttLeaf &D, int pr, int row) {
    constexpr int L2 = plan_l2(LOGN, V), N2 = 1 << L2, ROW_THREADS = row_threads(LOGN, V);
    // twiddle prefetch across barriers for rows of 2^13 points and up, whose
    // LDS (not registers) bounds the occupancy; not in kernels held to 64
    // VGPRs (measured slower on the shorter rows of batches)
    constexpr bool PRE = row_waves(LOGN, V) < 8 && L2 >= 13;
    constexpr int RBV = plan_rb(LOGN, V);
    // The last forward pass (block size R_L, no twiddles) runs in registers:
    // thread t holds the outputs of last-pass columns t + it * ROW_THREADS,
    // positions col * R_L + bitrev(q), through the pointwise product and the
    // inverse's first pass, which reads exactly those positions.
    constexpr int NPS = split_np(L2, RBV);
    constexpr int LRL = split_lr(L2, NPS - 1, RBV), RL = 1 << LRL;
    constexpr int NCL = N2 >> LRL, IL = (NCL + ROW_THREADS - 1) / ROW_THREADS;
    extern __shared__ uint32_t lds[];
    const PrimeK &c = L.pk[pr];
    if constexpr (row_dual(LOGN, V)) {
        // Both operands' rows in LDS (A at 0, B at N2): every forward pass
        // transforms a column of each with one load of its twiddles, then B's
        // last pass, the pointwise product and the inverse's first pass run in
        // registers, and the inverse finishes in B's tile.
        uint32_t *lda = lds, *ldb = lds + N2;
        load_rows<LOGN, MODE, V, 2>(lda, L, D, 0, pr, row);   // A at lda, B at lda + N2 = ldb
        sync_fwd2_pre<L2, 0, ROW_THREADS, RBV, NPS - 1, LZ, PRE>(lda, ldb, L.row_tw[pr], c, threadIdx.x);
#pragma unroll
        for (int it = 0; it < IL; ++it) {
            const int col = threadIdx.x + it * ROW_THREADS;
            if (col < NCL) {
                uint32_t xa[RL], x[RL];
#pragma unroll
                for (int q = 0; q < RL; ++q) xa[q] = lda[lp(col * RL + q)];
                dft_dif<RL, false, LZ>(xa, c);
#pragma unroll
                for (int q = 0; q < RL; ++q) x[q] = ldb[lp(col * RL + q)];
                dft_dif<RL, false, LZ>(x, c);
                // lazy operands: x xa < 4p^2 < p 2^32
#pragma unroll
                for (int q = 0; q < RL; ++q) x[q] = mont_mul(x[q], xa[q], c.p, c.ninv);
                dft_dit_inv<RL, LZ>(x, c);
#pragma unroll
                for (int q = 0; q < RL; ++q) ldb[lp(col * RL + q)] = x[q];
            }
        }
        sync_inverse_pre<L2, 0, ROW_THREADS, RBV, NPS - 1, LZ, PRE>(ldb, L.row_itw[pr], c, threadIdx.x);
        const auto out = gptr_mut((u32x4 *)(D.prod + ((size_t)pr << LOGN) + ((size_t)row << L2)));
        for (int m = threadIdx.x; m < N2 / 4; m += ROW_THREADS)
            out[m] = u32x4{ldb[lp(4 * m)], ldb[lp(4 * m + 1)], ldb[lp(4 * m + 2)], ldb[lp(4 * m + 3)]};
        return;
    }
    uint32_t keep[IL][RL];

    constexpr int E4 = (N2 / 4 + ROW_THREADS - 1) / ROW_THREADS;
    // larger rows would spill; radix-8 rows at 64 VGPRs rely on occupancy instead
    constexpr bool PF = plan_l1(LOGN, V) > 0 && E4 <= 4 && row_waves(LOGN, V) < 8;
    load_rows<LOGN, MODE, V, 1>(lds, L, D, 0, pr, row);
    // operand B's spectrum row is fetched now and lands during A's transform
    uint4 pb[PF ? E4 : 1];
    const uint4 *inb = (const uint4 *)op_spec<LOGN, mode_primes(MODE)>(D, 1, pr);
    if constexpr (PF) {
#pragma unroll
        for (int k = 0; k < E4; ++k) {
            const int m = threadIdx.x + k * ROW_THREADS;
            if (m < N2 / 4) pb[k] = inb[spec_off<LOGN, V>(row, 4 * m) >> 2];
        }
    }
    lds_sync();
    lds_forward<L2, 0, ROW_THREADS, RBV, NPS - 1, LZ>(lds, L.row_tw[pr], c);
#pragma unroll
    for (int it = 0; it < IL; ++it) {
        const int col = threadIdx.x + it * ROW_THREADS;
        if (col < NCL) {
#pragma unroll
            for (int q = 0; q < RL; ++q) keep[it][q] = lds[lp(col * RL + q)];
            dft_dif<RL, false, LZ>(keep[it], c);
        }
    }
    lds_sync();
    if constexpr (PF) {
#pragma unroll
        for (int k = 0; k < E4; ++k) {
            const int m = threadIdx.x + k * ROW_THREADS;
            if (m < N2 / 4) {
                lds[lp(4 * m)] = pb[k].x; lds[lp(4 * m + 1)] = pb[k].y;
                lds[lp(4 * m + 2)] = pb[k].z; lds[lp(4 * m + 3)] = pb[k].w;
            }
        }
    } else {
        load_rows<LOGN, MODE, V, 1>(lds, L, D, 1, pr, row);
    }
    lds_sync();
    lds_forward<L2, 0, ROW_THREADS, RBV, NPS - 1, LZ>(lds, L.row_tw[pr], c);
    // B's last pass, pointwise product and the inverse's first pass share the
    // registers; every thread reads its inputs before any writes its outputs
    // to the same positions, within one column
#pragma unroll
    for (int it = 0; it < IL; ++it) {
        const int col = threadIdx.x + it * ROW_THREADS;
        if (col < NCL) {
            uint32_t x[RL];
#pragma unroll
            for (int q = 0; q < RL; ++q) x[q] = lds[lp(col * RL + q)];
            dft_dif<RL, false, LZ>(x, c);
#pragma unroll
            for (int q = 0; q < RL; ++q) x[q] = mont_mul(x[q], keep[it][q], c.p, c.ninv);
            dft_dit_inv<RL, LZ>(x, c);
#pragma unroll
            for (int q = 0; q < RL; ++q) lds[lp(col * RL + q)] = x[q];
        }
    }
    sync_inverse_pre<L2, 0, ROW_THREADS, RBV, NPS - 1, LZ, PRE>(lds, L.row_itw[pr], c, threadIdx.x);
    const auto out = gptr_mut((u32x4 *)(D.prod + ((size_t)pr << LOGN) + ((size_t)row << L2)));
    for (int m = threadIdx.x; m < N2 / 4; m += ROW_THREADS)
        out[m] = u32x4{lds[lp(4 * m)], lds[lp(4 * m + 1)], lds[lp(4 * m + 2)], lds[lp(4 * m + 3)]};
}

// Row DFTs of both operands, pointwise product, inverse row DFT.  The lazy
// prime's residues leave in [0, 2p); the inverse column kernel's first
// Montgomery product takes them as they are.
template <int LOGN, int MODE, int V>
__global__ void __launch_bounds__(row_threads(LOGN, V)) __attribute__((amdgpu_waves_per_eu(row_waves(LOGN, V))))
k_row(const NttLaunch L) {
    const NttLeaf D = leaf_of(L);
    // consecutive rows on one XCD: the rows of a spectrum group share its lines
    const int row = xcd_tile(blockIdx.x, gridDim.x);
    const int pr = (L.order & ROW_PRIME_REV) ? (int)gridDim.y - 1 - (int)blockIdx.y : blockIdx.y;
    if (pr == LAZY_PRIME) row_body<LOGN, MODE, V, true>(L, D, pr, row);
    else row_body<LOGN, MODE, V, false>(L, D, pr, row);
}

// ---- kernel 3: inverse column pass, CRT, limb sums ------------------------------

// 128-bit accumulate: (lo, hi) += v * 2^sh for a 32-bit v and 0 <= sh < 128.
__device__ __forceinline__ void acc128(uint64_t &lo, uint64_t &hi, uint64_t v, int sh) {
    uint64_t plo, phi;
    if (sh == 0) { plo = v; phi = 0; }
    else if (sh < 64) { plo = v << sh; phi = v >> (64 - sh); }
    else { plo = 0; phi = v << (sh - 64); }
    const uint64_t s = lo + plo;
    hi += phi + (s < lo);
    lo = s;
}

// Prime pr's product residues of the tile, times the inverse four-step
// twiddle omega_n^(-C * freq(rho)) and 1/n (both folded into the ilo / ihi
// tables), into LDS.  Element e = threadIdx.x + k * COLINV_THREADS: row
// e >> LT, column c0 + (e mod T).
template <int LOGN, int V>
__device__ __forceinline__ void stage_prime(uint32_t *lds, const NttLaunch &L, const NttLeaf &D, int pr,
                                            uint32_t c0, int tid) {
    constexpr int L1 = plan_l1(LOGN, V), L2 = plan_l2(LOGN, V), LT = plan_lt(LOGN, V), RBV = plan_rb(LOGN, V);
    constexpr int T = 1 << LT, NE = (1 << L1) * T;
    constexpr uint32_t NMASK = (1u << LOGN) - 1;
    using TW = ColTw<L1, LT, RBV, COLINV_THREADS>;
    const PrimeK &c = L.pk[pr];
    static_assert(COLINV_THREADS >= T, "stage_prime: one column per thread");
    const auto in = gptr(D.prod + ((size_t)pr << LOGN));
    const uint32_t C = c0 + (tid & (T - 1));
    if constexpr (TW::STEP) {
        const uint32_t f0 = freq_of_pos<L1, RBV>(tid >> LT);
        uint32_t w = twiddle_mont(L.ilo[pr], L.ihi[pr], (C * f0) & NMASK, L.lobits, c);
        // omega_n^(-C * 2^D) in Montgomery form, from the forward tables
        const uint32_t st = twiddle_mont(L.glo[pr], L.ghi[pr], (0u - (C << TW::D)) & NMASK, L.lobits, c);
        // residues in groups of 8 loads in flight (E is a power of two)
        constexpr int G = TW::E < 8 ? TW::E : 8;
#pragma unroll 1
        for (int k0 = 0; k0 < TW::E; k0 += G) {
            uint32_t x[G];
#pragma unroll
            for (int k = 0; k < G; ++k) {
                const int e = tid + (k0 + k) * COLINV_THREADS;
                x[k] = in[((size_t)(e >> LT) << L2) + C];
            }
#pragma unroll
            for (int k = 0; k < G; ++k) {
                lds[lp(tid + (k0 + k) * COLINV_THREADS)] = mont_mul(x[k], w, c.p, c.ninv);
                w = mont_mul(w, st, c.p, c.ninv);
            }
        }
    } else {
        for (int e = tid; e < NE; e += COLINV_THREADS) {
            const int rho = e >> LT, cc = e & (T - 1);
            const uint32_t ex = ((c0 + cc) * freq_of_pos<L1, RBV>(rho)) & NMASK;
            const uint32_t x = in[((size_t)rho << L2) + c0 + cc];
            lds[lp(e)] = mont_mul(x, twiddle_mont(L.ilo[pr], L.ihi[pr], ex, L.lobits, c), c.p, c.ninv);
        }
    }
}

// Pipelined staging (inverse column kernel).  The tile's residues of the
// next prime are copied by LDS DMA (global_load_lds) into a staging tile
// B = lds + NE in plain order (element e = row * T + column at B[e]) while
// the current prime's passes run; stage_store then applies the four-step
// twiddles on the way from B to the transform tile.  DMA k of wave w carries
// 64 lanes x 16 bytes, elements [256 (16 k + w), + 256): lane l holds
// elements 4l .. 4l + 3 of that chunk, four columns of one row.
template <int LOGN, int V>
__device__ __forceinline__ void stage_dma(uint32_t *B, const uint32_t *prod, int pr, uint32_t c0, int tid) {
    constexpr int L1 = plan_l1(LOGN, V), L2 = plan_l2(LOGN, V), LT = plan_lt(LOGN, V);
    constexpr int NE = (1 << L1) << LT, NW = COLINV_THREADS / 64;
    static_assert(LT >= 2 && NE % 256 == 0, "stage_dma: rows of whole 16-byte pieces");
    const int w = tid >> 6, l = tid & 63;
    const uint32_t *in = prod + ((size_t)pr << LOGN);
#pragma unroll
    for (int k = 0; k < (NE / 256 + NW - 1) / NW; ++k) {
        const int chunk = k * NW + w;
        if (NE % (256 * NW) == 0 || chunk < NE / 256) {
            const uint32_t e = (uint32_t)chunk * 256 + 4 * l;
            const uint32_t off = ((e >> LT) << L2) + c0 + (e & ((1u << LT) - 1));
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(in + off),
                                             (__attribute__((address_space(3))) void *)(B + chunk * 256), 16, 0, 0);
        }
    }
}

// Table words of the four-step twiddle and its step for prime pr (loaded
// with the DMA, used by stage_store).
struct StageTw {
    uint32_t wlo, whi, slo, shi;
};
template <int LOGN, int V>
__device__ __forceinline__ void stage_tw(StageTw &s, const NttLaunch &L, int pr, uint32_t c0, int tid) {
    constexpr int L1 = plan_l1(LOGN, V), LT = plan_lt(LOGN, V), RBV = plan_rb(LOGN, V);
    constexpr uint32_t NMASK = (1u << LOGN) - 1;
    using TW = ColTw<L1, LT, RBV, COLINV_THREADS>;
    static_assert(TW::STEP, "stage_tw: stepped twiddles only");
    const uint32_t C = c0 + (tid & ((1 << LT) - 1));
    const uint32_t ew = (C * freq_of_pos<L1, RBV>(tid >> LT)) & NMASK, es = (0u - (C << TW::D)) & NMASK;
    const uint32_t lmask = (1u << L.lobits) - 1;
    s.wlo = gptr(L.ilo[pr])[ew & lmask];
    s.whi = gptr(L.ihi[pr])[ew >> L.lobits];
    s.slo = gptr(L.glo[pr])[es & lmask];
    s.shi = gptr(L.ghi[pr])[es >> L.lobits];
}

// The staged residues (B, after the DMA) times their four-step twiddles and
// 1/n into the transform tile, as stage_prime writes them.
template <int LOGN, int V>
__device__ __forceinline__ void stage_store(uint32_t *lds, const uint32_t *B, const StageTw &s, const PrimeK &c,
                                            int tid) {
    constexpr int E = ((1 << plan_l1(LOGN, V)) << plan_lt(LOGN, V)) / COLINV_THREADS;
    uint32_t w = mont_mul(s.wlo, s.whi, c.p, c.ninv);
    const uint32_t st = mont_mul(s.slo, s.shi, c.p, c.ninv);
#pragma unroll
    for (int k = 0; k < E; ++k) {
        const int e = tid + k * COLINV_THREADS;
        lds[lp(e)] = mont_mul(B[e], w, c.p, c.ninv);
        w = mont_mul(w, st, c.p, c.ninv);
    }
}

// Prime pr's residues through the inverse column transform; the result is in
// LDS (after a workgroup barrier).
// The lazy prime's residues come out in [0, 2p).
template <int LOGN, int V>
__device__ __forceinline__ void inv_column(uint32_t *lds, const NttLaunch &L, const NttLeaf &D, int pr, uint32_t c0,
                                           int tid = threadIdx.x) {
    constexpr int L1 = plan_l1(LOGN, V), LT = plan_lt(LOGN, V), RBV = plan_rb(LOGN, V);
    stage_prime<LOGN, V>(lds, L, D, pr, c0, tid);
    lds_sync();
    if constexpr (L1 > 0) {
        if (pr == LAZY_PRIME) lds_inverse<L1, LT, COLINV_THREADS, RBV, MAX_PASS, true>(lds, L.col_itw[pr], L.pk[pr], tid);
        else lds_inverse<L1, LT, COLINV_THREADS, RBV, MAX_PASS, false>(lds, L.col_itw[pr], L.pk[pr], tid);
    }
}

// Value of lane (this lane ^ 1), by a DPP quad permutation [1, 0, 3, 2].
__device__ __forceinline__ uint32_t pair_swap(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);
}

// Value of lane (this lane - 2) within each row of 16 lanes (DPP row_shr:2;
// the first two lanes of a row read 0).
__device__ __forceinline__ uint64_t lane_down2(uint64_t v) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, 0x112, 0xF, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), 0x112, 0xF, 0xF, false);
    return (uint64_t)hi << 32 | lo;
}

// Wide mode: three primes, Garner
//   t2 = (r2 - r1) p1^-1 mod p2,  t3 = (r3 - (r1 + p1 t2)) (p1 p2)^-1 mod p3,
//   x  = r1 + p1 t2 + p1 p2 t3 < 2^89 (the exact coefficient);
// a limb is columns (2m, 2m+1) of a row, x_even + 2^32 x_odd, held by lanes
// (2i, 2i+1), so the odd lane's coefficient reaches its partner by DPP and the
// even lane writes the 128-bit limb sum.  With at most 8 elements per
// thread this fits 64 VGPRs and two workgroups share a CU (col_inv_waves).
// (Halved tiles measured slower at 2^24 points: with 1024 threads they spill
// at 64 VGPRs; with 512 the occupancy stays and the rows are read in 16-byte
// pieces.)
// Two workgroups per CU (at most 64 VGPRs) where the Garner state (2 words per
// element) and a transform pass fit: 8 elements per thread or radix-8 passes.
template <int LOGN, int V>
constexpr int col_inv_waves() {
    return ((1 << (plan_l1(LOGN, V) + plan_lt(LOGN, V))) <= 8 * COLINV_THREADS || plan_rb(LOGN, V) == 3) ? 8 : 4;
}

// Whether the wide-mode inverse column kernel stages its primes by LDS DMA
// (plans with stepped four-step twiddles whose LDS fits, in kernels that
// hold a CU with one workgroup anyway; where two workgroups share a CU, they
// hide each other's loads, and twice the LDS would leave room for one): its
// LDS then holds
// the transform tile, the staging tile and the three primes' inverse pass
// twiddles, so its passes issue no global loads and the DMA of the next prime
// runs under them (global loads complete in order: a twiddle load issued
// after the DMA would wait for it).
template <int LOGN, int V>
constexpr int col_inv_tw_words() {
    return inv_tw_words(plan_l1(LOGN, V), plan_rb(LOGN, V), 0, plan_l1(LOGN, V));
}
template <int LOGN, int V>
constexpr size_t col_inv_pipe_lds() {
    return (2 * ((size_t)1 << (plan_l1(LOGN, V) + plan_lt(LOGN, V))) + 3 * (size_t)col_inv_tw_words<LOGN, V>()) * 4;
}
template <int LOGN, int V>
constexpr bool col_inv_pipe() {
    return col_inv_waves<LOGN, V>() == 4 && plan_l1(LOGN, V) > 0 &&
           ColTw<plan_l1(LOGN, V), plan_lt(LOGN, V), plan_rb(LOGN, V), COLINV_THREADS>::STEP &&
           col_inv_pipe_lds<LOGN, V>() <= 156 * 1024;
}

// One prime's step of Garner's CRT for the tile's elements (the prime's
// residues in LDS): prime 0 sets r0, prime 1 sets t2, prime 2 forms each
// coefficient and writes the limb sums of the tile starting at column c0.
template <int LOGN, int V, int E>
__device__ __forceinline__ void crt_step(const uint32_t *lds, const NttLaunch &L, int pr, uint32_t c0, int tid,
                                         uint32_t (&r0)[E], uint32_t (&t2)[E], uint64_t *slo, uint64_t *shi,
                                         size_t sum_len, int pack_log) {
    constexpr int L1 = plan_l1(LOGN, V), L2 = plan_l2(LOGN, V), LT = plan_lt(LOGN, V);
    constexpr int T = 1 << LT, NE = (1 << L1) * T;
    // packed sums: pieces of P2 = min(T / 2, 8) limbs, within one row of 16
    // lanes (a piece of 8 limbs is one such row: tiles start at multiples of T)
    constexpr int PL = sum_pack_log(MODE_WIDE, LT), P2 = 1 << PL;
    constexpr bool PACKABLE = PL > 0;
    static_assert(COLINV_THREADS >= T, "crt_step: one column per thread");
    const CrtK &K = L.crt;
    const uint32_t C = c0 + (tid & (T - 1));
    const bool even = (tid & 1) == 0;
    if (pr == 0) {
        const uint32_t p1 = L.pk[0].p;
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const int e = tid + k * COLINV_THREADS;
            const uint32_t r = e < NE ? lds[lp(e)] : 0;
            r0[k] = min(r, r - p1);   // lazy prime: [0, 2p) -> [0, p)
        }
    } else if (pr == 1) {
        const uint32_t p2 = L.pk[1].p;
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const int e = tid + k * COLINV_THREADS;
            const uint32_t r = e < NE ? lds[lp(e)] : 0;
            t2[k] = mul_shoup(mod_sub(r, r0[k], p2), K.inv12, K.inv12q, L.pk[1]);
        }
    } else {
        const uint32_t p3 = L.pk[2].p;
        const uint64_t p1 = L.pk[0].p, p12 = p1 * L.pk[1].p;
        const uint32_t p12lo = (uint32_t)p12, p12hi = (uint32_t)(p12 >> 32);
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const int e = tid + k * COLINV_THREADS;
            const uint32_t r = e < NE ? lds[lp(e)] : 0;
            // (r1 + p1 t2) mod p3, with r1 < p1 < p3 and t2 < p2 < p3
            const uint32_t x12m = mod_add(r0[k], mul_shoup(t2[k], K.p1_mod3, K.p1_mod3q, L.pk[2]), p3);
            const uint32_t t3 = mul_shoup(mod_sub(r, x12m, p3), K.inv123, K.inv123q, L.pk[2]);
            // x = a + 2^32 b
            const uint64_t x12 = (uint64_t)r0[k] + p1 * t2[k];
            const uint64_t s0 = x12 + (uint64_t)p12lo * t3;
            const uint32_t a = (uint32_t)s0;
            const uint64_t b = (s0 >> 32) + ((uint64_t)(s0 < x12) << 32) + (uint64_t)p12hi * t3;
            const uint32_t ao = pair_swap(a);
            const uint32_t bo_lo = pair_swap((uint32_t)b), bo_hi = pair_swap((uint32_t)(b >> 32));
            const uint64_t m = b + ao;
            const uint64_t lo = (uint64_t)a | (m << 32), hi = (m >> 32) + ((uint64_t)bo_hi << 32 | bo_lo);
            const size_t li = ((size_t)(e >> LT) << (L2 - 1)) + (C >> 1);
            if (PACKABLE && pack_log) {
                // limb q of a piece (P2 limbs, on the even lanes 0, 2, ..)
                // adds the high word of limb q - 1, two lanes down
                const uint32_t q = (C >> 1) & (P2 - 1);
                const uint64_t hp = lane_down2(hi);
                const uint64_t t = q ? lo + hp : lo;
                // bit q: limb q's addition carried; OR-ed up the even lanes
                // of the piece (shifts of 2, 4, 8 lanes within a row of 16),
                // so the last limb's lane holds the bits of limbs 1 .. q
                uint32_t gb = (q && t < lo) ? 1u << q : 0u;
                gb |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)gb, 0x112, 0xF, 0xF, false);
                if constexpr (P2 >= 4) gb |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)gb, 0x114, 0xF, 0xF, false);
                if constexpr (P2 >= 8) gb |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)gb, 0x118, 0xF, 0xF, false);
                if (even && e < NE && li < sum_len) {
                    gptr_mut(slo)[li] = t;
                    if (q == P2 - 1)   // the piece's last limb writes its high word and carries
                        gptr_mut(shi)[li >> PL] = hi | (uint64_t)gb << 56;
                }
            } else if (even && e < NE && li < sum_len) {
                gptr_mut(slo)[li] = lo;
                gptr_mut(shi)[li] = hi;
            }
        }
    }
}

template <int LOGN, int V>
__device__ __forceinline__ void col_inv_wide(uint32_t *lds, const NttLaunch &L, const NttLeaf &D, uint32_t c0) {
    constexpr int L1 = plan_l1(LOGN, V), LT = plan_lt(LOGN, V);
    constexpr int NE = (1 << L1) << LT, E = (NE + COLINV_THREADS - 1) / COLINV_THREADS;
    uint64_t *const slo = uniform_ptr(D.slo), *const shi = uniform_ptr(D.shi);
    const size_t sum_len = uniform_u64(D.sum_len);
    const int pack_log = __builtin_amdgcn_readfirstlane(D.pack_log);
    uint32_t r0[E], t2[E];
#pragma unroll 1
    for (int pr = 0; pr < 3; ++pr) {
        if (pr) lds_sync();   // the previous prime's residues have been read
        // The thread index is re-read each prime, so the index arithmetic of
        // the passes is recomputed instead of being held, beside the Garner
        // state, across the whole loop.
        const int tid = fresh_tid();
        inv_column<LOGN, V>(lds, L, D, pr, c0, tid);
        crt_step<LOGN, V>(lds, L, pr, c0, tid, r0, t2, slo, shi, sum_len, pack_log);
    }
}

// The pipelined form (col_inv_pipe): one workgroup per CU works through
// tiles i = blockIdx.x, blockIdx.x + gridDim.x, .. (tile xcd_tile(i)), three
// primes each, as one sequence of steps; the DMA of each step's residues is
// issued while the step before it runs, the first step's alone waits.  The
// three primes' pass twiddles are put in LDS once.
template <int LOGN, int V>
__device__ __forceinline__ void col_inv_wide_pipe(uint32_t *lds, const NttLaunch &L, const NttLeaf &D) {
    constexpr int L1 = plan_l1(LOGN, V), L2 = plan_l2(LOGN, V), LT = plan_lt(LOGN, V), RBV = plan_rb(LOGN, V);
    constexpr int NE = (1 << L1) << LT, E = NE / COLINV_THREADS, NTILES = (1 << L2) >> LT;
    constexpr int TWW = col_inv_tw_words<LOGN, V>();
    uint32_t *const B = lds + NE, *const TWL = lds + 2 * NE;
    // descriptor fields, read once into scalar registers (read through D at
    // each use, each is a load that later stores wait for)
    uint64_t *const slo = uniform_ptr(D.slo), *const shi = uniform_ptr(D.shi);
    const size_t sum_len = uniform_u64(D.sum_len);
    const int pack_log = __builtin_amdgcn_readfirstlane(D.pack_log);
    const uint32_t *const prod = uniform_ptr(D.prod);
    const int nsteps = 3 * (int)((NTILES - blockIdx.x + gridDim.x - 1) / gridDim.x);
    const auto col0 = [&](int step) {
        return (uint32_t)xcd_tile((int)blockIdx.x + step / 3 * (int)gridDim.x, NTILES) << LT;
    };
    StageTw stw;
    stage_dma<LOGN, V>(B, prod, 0, col0(0), threadIdx.x);
    stage_tw<LOGN, V>(stw, L, 0, col0(0), threadIdx.x);
#pragma unroll
    for (int p = 0; p < 3; ++p)
        fill_inv_tw<L1, COLINV_THREADS, L1, 0, RBV, 0>(TWL + p * TWW, L.col_itw[p], L.pk[p], threadIdx.x);
    uint32_t r0[E], t2[E];
#pragma unroll 1
    for (int step = 0; step < nsteps; ++step) {
        const int pr = step % 3;
        // this wave's DMA of the step has landed; after the barrier every
        // wave's has, and the previous step's CRT reads are done
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        lds_sync();
        // (the thread index is re-read each step, so the index arithmetic
        // is recomputed instead of being held beside the Garner state)
        const int tid = fresh_tid();
        stage_store<LOGN, V>(lds, B, stw, L.pk[pr], tid);
        lds_sync();
        // B is free: the next step's residues arrive while this one runs
        if (step + 1 < nsteps) {
            const int pn = (step + 1) % 3;
            stage_dma<LOGN, V>(B, prod, pn, col0(step + 1), tid);
            stage_tw<LOGN, V>(stw, L, pn, col0(step + 1), tid);
        }
        if (pr == LAZY_PRIME)
            inv_from_lt<L1, LT, COLINV_THREADS, L1, 0, RBV, true, 0>(lds, TWL + pr * TWW, L.pk[pr], tid);
        else
            inv_from_lt<L1, LT, COLINV_THREADS, L1, 0, RBV, false, 0>(lds, TWL + pr * TWW, L.pk[pr], tid);
        crt_step<LOGN, V>(lds, L, pr, col0(step), tid, r0, t2, slo, shi, sum_len, pack_log);
    }
}

// Reference mode: two primes, coefficient x = r1 + p1 t (< 2^60), four
// coefficients per limb, summed through LDS.
template <int LOGN, int V>
__device__ __forceinline__ void col_inv_ref(uint32_t *lds, const NttLaunch &L, const NttLeaf &D, uint32_t c0) {
    constexpr int L1 = plan_l1(LOGN, V), L2 = plan_l2(LOGN, V), LT = plan_lt(LOGN, V);
    constexpr int DPL = 4;
    constexpr int N1 = 1 << L1, T = 1 << LT, NE = N1 * T, LPR = T / DPL, NL = N1 * LPR;
    constexpr int E = (NE + COLINV_THREADS - 1) / COLINV_THREADS;
    constexpr int EL = (NL + COLINV_THREADS - 1) / COLINV_THREADS;
    const CrtK &K = L.crt;
    uint64_t slo[EL], shi[EL];
#pragma unroll
    for (int k = 0; k < EL; ++k) { slo[k] = 0; shi[k] = 0; }
    uint32_t res0[E];
    uint64_t coef[E];
    for (int pr = 0; pr < 2; ++pr) {
        inv_column<LOGN, V>(lds, L, D, pr, c0);
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const int i = threadIdx.x + k * COLINV_THREADS;
            if (i < NE) {
                const uint32_t r = lds[lp(i)];
                if (pr == 0) {
                    res0[k] = min(r, r - L.pk[0].p);   // lazy prime: [0, 2p) -> [0, p)
                } else {
                    // Garner (mul_ntt.cpp:405-412): x = r1 + p1 * ((r2 - r1) * p1^-1 mod p2)
                    const uint32_t d = mod_sub(r, res0[k], L.pk[1].p);   // r1 < p1 < p2
                    const uint32_t t = mul_shoup(d, K.inv12, K.inv12q, L.pk[1]);
                    coef[k] = (uint64_t)res0[k] + (uint64_t)L.pk[0].p * t;
                }
            }
        }
        lds_sync();
    }
#pragma unroll
    for (int half = 0; half < 2; ++half) {
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const int i = threadIdx.x + k * COLINV_THREADS;
            if (i < NE) lds[lp(i)] = half ? (uint32_t)(coef[k] >> 32) : (uint32_t)coef[k];
        }
        lds_sync();
#pragma unroll
        for (int k = 0; k < EL; ++k) {
            const int e = threadIdx.x + k * COLINV_THREADS;
            if (e < NL) {
                const int rho = e / LPR, m = e % LPR;
#pragma unroll
                for (int t = 0; t < 4; ++t) acc128(slo[k], shi[k], lds[lp(rho * T + 4 * m + t)], 16 * t + 32 * half);
            }
        }
        lds_sync();
    }
#pragma unroll
    for (int k = 0; k < EL; ++k) {
        const int e = threadIdx.x + k * COLINV_THREADS;
        if (e < NL) {
            const int rho = e / LPR, m = e % LPR;
            const size_t li = (((size_t)rho << L2) + c0 + DPL * m) / DPL;
            if (li < D.sum_len) {
                gptr_mut(D.slo)[li] = slo[k];
                gptr_mut(D.shi)[li] = shi[k];
            }
        }
    }
}

template <int LOGN, int MODE, int V>
__global__ void __launch_bounds__(COLINV_THREADS)
__attribute__((amdgpu_waves_per_eu(MODE == MODE_WIDE ? col_inv_waves<LOGN, V>() : 1)))
k_col_inv(const NttLaunch L) {
    extern __shared__ uint32_t lds[];
    // (read through the reference here: leaf_of's scalar registers push this
    // kernel to 106 SGPRs and its pass twiddle loads are then issued one at
    // a time, 208 -> 227 us at 2^24 points)
    const NttLeaf &D = L.leaves ? L.leaves[blockIdx.z] : L.one;
    const uint32_t c0 = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) << plan_lt(LOGN, V);
    if constexpr (MODE == MODE_WIDE && col_inv_pipe<LOGN, V>()) col_inv_wide_pipe<LOGN, V>(lds, L, D);
    else if constexpr (MODE == MODE_WIDE) col_inv_wide<LOGN, V>(lds, L, D, c0);
    else col_inv_ref<LOGN, V>(lds, L, D, c0);
}

// ---- host-side launch table ----------------------------------------------------

// Compute units of the current device (looked up once per device).
static inline int device_cus() {
    static int cus[64] = {};
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess || d < 0 || d >= 64) return 256;
    if (!cus[d]) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, d) != hipSuccess || n <= 0) n = 256;
        cus[d] = n;
    }
    return cus[d];
}

template <int LOGN, int MODE, int V>
static hipError_t launch_size(const NttLaunch &L, hipStream_t stream, const StageHook &h) {
    constexpr int L1 = plan_l1(LOGN, V), L2 = plan_l2(LOGN, V), LT = plan_lt(LOGN, V), NP = mode_primes(MODE);
    const int ntiles = (1 << L2) >> LT;
    const size_t col_lds = lds_words((size_t)1 << (L1 + LT)) * 4,
                 row_lds = lds_words((size_t)1 << L2) * 4 * (row_dual(LOGN, V) ? 2 : 1);
    const int nl = L.leaves ? L.nleaves : 1;
    // byte-model shares use the reference's P = 2 (see bench.py)
    const double n = (double)(1u << LOGN) * nl, io = L.io_bytes;
    if constexpr (L1 > 0) {
        h.begin(h.ctx, "ntt_col_fwd");
        const int of = L.order & 3;
        const dim3 g = of == COLF_OP_PRIME ? dim3(ntiles, 2 * NP, nl)
                       : of == COLF_TILE_PRIME ? dim3(ntiles * NP, 2, nl) : dim3(ntiles * 2, NP, nl);
        hipLaunchKernelGGL((k_col_fwd<LOGN, MODE, V>), g, dim3(col_threads(LOGN, V)), col_lds, stream, L);
        h.end(h.ctx, "ntt_col_fwd", io + 16.0 * NP * n);
    }
    h.begin(h.ctx, "ntt_row");
    hipLaunchKernelGGL((k_row<LOGN, MODE, V>), dim3(1 << L1, NP, nl), dim3(row_threads(LOGN, V)), row_lds, stream, L);
    h.end(h.ctx, "ntt_row", 24.0 * NP * n + (L1 ? 0.0 : io + 16.0 * NP * n));
    h.begin(h.ctx, "ntt_col_inv");
    // the pipelined inverse column kernel: one workgroup per CU, each
    // working through several tiles
    constexpr bool PIPE = MODE == MODE_WIDE && col_inv_pipe<LOGN, V>();
    const size_t inv_lds = PIPE ? col_inv_pipe_lds<LOGN, V>() : col_lds;
    const int inv_grid = PIPE && device_cus() < ntiles ? device_cus() : ntiles;
    hipLaunchKernelGGL((k_col_inv<LOGN, MODE, V>), dim3(inv_grid, 1, nl), dim3(COLINV_THREADS), inv_lds, stream, L);
    h.end(h.ctx, "ntt_col_inv", 8.0 * NP * n);
    return hipGetLastError();
}

template <int LOGN, int MODE, int V>
static hipError_t set_limits() {
    constexpr int lim = 156 * 1024;   // dynamic LDS limit; leaves room for static LDS
    hipError_t e = hipSuccess;
    if (plan_l1(LOGN, V) > 0)
        e = hipFuncSetAttribute((const void *)k_col_fwd<LOGN, MODE, V>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)k_row<LOGN, MODE, V>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)k_col_inv<LOGN, MODE, V>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    return e;
}

// Launch / LDS-limit functions of one plan variant over the wide-mode sizes
// that have variants (2^12 .. 2^26 points).
#define PG_NTT_VARIANT_SIZES(X) X(12) X(13) X(14) X(15) X(16) X(17) X(18) X(19) X(20) X(21) X(22) X(23) \
    X(24) X(25) X(26)
template <int N, int V>
static hipError_t launch_variant(const NttLaunch &L, hipStream_t stream, const StageHook &hook) {
    if constexpr (plan_distinct(N, V)) return launch_size<N, MODE_WIDE, V>(L, stream, hook);
    else return hipErrorInvalidValue;
}
template <int N, int V>
static hipError_t limits_variant() {
    if constexpr (plan_distinct(N, V)) return set_limits<N, MODE_WIDE, V>();
    else return hipSuccess;
}
#define PG_NTT_VARIANT_FUNCS(V)                                                                          \
    hipError_t launch_ntt_wide_v##V(const NttLaunch &L, hipStream_t stream, const StageHook &hook) {    \
        switch (L.logn) {                                                                                \
            PG_NTT_VARIANT_SIZES(PG_VCASE)                                                               \
        default: return hipErrorInvalidValue;                                                            \
        }                                                                                                \
    }                                                                                                    \
    hipError_t ntt_limits_wide_v##V() {                                                                  \
        hipError_t e = hipSuccess;                                                                       \
        PG_NTT_VARIANT_SIZES(PG_VLIM)                                                                    \
        return e;                                                                                        \
    }

#define PG_NTT_VARIANT_DECL(V)                                                                          \
    hipError_t launch_ntt_wide_v##V(const NttLaunch &L, hipStream_t stream, const StageHook &hook);     \
    hipError_t ntt_limits_wide_v##V();
PG_NTT_VARIANT_DECL(1) PG_NTT_VARIANT_DECL(2) PG_NTT_VARIANT_DECL(3) PG_NTT_VARIANT_DECL(4)
PG_NTT_VARIANT_DECL(5) PG_NTT_VARIANT_DECL(6) PG_NTT_VARIANT_DECL(7) PG_NTT_VARIANT_DECL(8)
PG_NTT_VARIANT_DECL(9) PG_NTT_VARIANT_DECL(10) PG_NTT_VARIANT_DECL(11) PG_NTT_VARIANT_DECL(12)
PG_NTT_VARIANT_DECL(13) PG_NTT_VARIANT_DECL(14) PG_NTT_VARIANT_DECL(15)

}  // namespace pg
