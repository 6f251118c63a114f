// pg_kernels.h -- launch descriptors shared by the host runtime and kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include "pg_modarith.h"

namespace pg {

// Global-memory view of a device pointer read from a descriptor: accesses
// through it are global (not flat) instructions.  A flat access also counts
// against the LDS counter, so LDS reads after it wait for it.
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *gptr(const T *p) {
    return (const __attribute__((address_space(1))) T *)p;
}
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gptr_mut(T *p) {
    return (__attribute__((address_space(1))) T *)p;
}

constexpr int NTT_THREADS = 1024;           // largest forward column workgroup (col_threads)
#ifndef PG_COLINV_THREADS
#define PG_COLINV_THREADS 1024
#endif
constexpr int COLINV_THREADS = PG_COLINV_THREADS;   // workgroup size of the inverse column kernel
constexpr int ROW_MAX_ELEMS = 1 << 15;      // longest in-LDS row transform (N2)
constexpr int COL_MAX_ELEMS = 1 << 14;      // largest column tile (N1 * T)
constexpr int MAX_PASS = 6;

// LDS words for an n-element NTT tile: the layout permutes words within each
// 64-word row (lp() in pg_ntt_impl.h), so it needs no padding.
constexpr size_t lds_words(size_t n) { return n; }

// ---- transform plan, shared by host and kernels ---------------------------
// n = 2^logn = N1 * N2: rows of N2 = 2^plan_l2 points are transformed in LDS
// by the row kernel; columns of N1 points in tiles of T = 2^plan_lt columns.
constexpr int NTT_MIN_LOG = 8, NTT_MAX_LOG = 26;
constexpr int cmin(int a, int b) { return a < b ? a : b; }
constexpr int cmax(int a, int b) { return a > b ? a : b; }
// log2 of a power of two (-1 otherwise)
constexpr int clog2(int x) { return x <= 0 || (x & (x - 1)) ? -1 : x == 1 ? 0 : 1 + clog2(x >> 1); }
#ifndef PG_ROW_LOG_MAX
#define PG_ROW_LOG_MAX 15
#endif
constexpr int ROW_LOG_MAX = PG_ROW_LOG_MAX;
constexpr int plan_l2_default(int logn) {
    return logn <= 11 ? logn : cmin(ROW_LOG_MAX, cmax((logn + 1) / 2, logn - 10));
}
#ifndef PG_RADIX_BITS
#define PG_RADIX_BITS 4
#endif
constexpr int RADIX_BITS = PG_RADIX_BITS;   // largest register pass of the default plans: radix 16
// An in-LDS transform of 2^l points as ceil(l / rb) radix passes of near-equal
// size (largest first).
constexpr int split_np(int l, int rb = RADIX_BITS) { return l ? (l + rb - 1) / rb : 0; }
constexpr int split_lr(int l, int i, int rb = RADIX_BITS) {
    return split_np(l, rb) == 0 ? 0
           : (i < split_np(l, rb) ? l / split_np(l, rb) + (i < l % split_np(l, rb) ? 1 : 0) : 0);
}

// Plan variants the autotuner chooses from (wide mode, logn >= 12).  Variant
// v = geometry (v mod 8) + 8 * radix choice.  Geometry g shifts the default
// row length and column-tile width by (VAR_DL2[g], VAR_DLT[g]):
//   0: default;  1: tiles half as wide;  2: rows one level shorter;
//   3: rows one level longer;  4: shorter rows, half tiles;  5: longer rows,
//   half tiles;  6: rows two levels shorter;  7: tiles a quarter as wide.
// Radix choice: passes of at most radix 2^RADIX_BITS (v < 8) or radix 8.
// A geometry that does not fit (rows outside 2^5 .. 2^ROW_LOG_MAX, columns
// outside 2^2 .. 2^12, tiles under 4 columns) keeps the default geometry.
constexpr int NTT_GEOMS = 8, NTT_VARIANTS = 16, NTT_VARIANT_MIN_LOG = 12;
constexpr int VAR_DL2[NTT_GEOMS] = {0, 0, -1, 1, -1, 1, -2, 0};
constexpr int VAR_DLT[NTT_GEOMS] = {0, -1, 0, 0, -1, -1, 0, -2};
// Single-row plans (l1 = 0) have no column transform: their "tiles" only feed
// the inverse column kernel's CRT, so they are as wide as its 1024 threads
// (64-column tiles left 15 of every 16 threads idle).
constexpr int lt_for(int l1, int l2) { return l1 == 0 ? cmin(l2, 10) : cmin(cmin(6, 14 - l1), l2); }
constexpr bool geom_fits(int logn, int g) {
    return g > 0 && g < NTT_GEOMS && logn >= NTT_VARIANT_MIN_LOG &&
           plan_l2_default(logn) + VAR_DL2[g] >= 5 && plan_l2_default(logn) + VAR_DL2[g] <= ROW_LOG_MAX &&
           logn - plan_l2_default(logn) - VAR_DL2[g] >= 2 && logn - plan_l2_default(logn) - VAR_DL2[g] <= 12 &&
           lt_for(logn - plan_l2_default(logn) - VAR_DL2[g], plan_l2_default(logn) + VAR_DL2[g]) + VAR_DLT[g] >= 2;
}
constexpr int plan_l2(int logn, int v = 0) {
    return geom_fits(logn, v % NTT_GEOMS) ? plan_l2_default(logn) + VAR_DL2[v % NTT_GEOMS] : plan_l2_default(logn);
}
constexpr int plan_l1(int logn, int v = 0) { return logn - plan_l2(logn, v); }
constexpr int plan_lt(int logn, int v = 0) {
    return lt_for(plan_l1(logn, v), plan_l2(logn, v)) + (geom_fits(logn, v % NTT_GEOMS) ? VAR_DLT[v % NTT_GEOMS] : 0);
}
// Radix bits of the variant's register passes.
constexpr int plan_rb(int logn, int v = 0) { return (v >= NTT_GEOMS && logn >= NTT_VARIANT_MIN_LOG) ? 3 : RADIX_BITS; }
// Forward spectra are stored in groups of 2^plan_gl rows: a group's T columns
// of one tile are one block of 2^plan_gl x T words, 64 contiguous bytes when
// 2^plan_gl T = 16 (spec_off in pg_ntt_impl.h).  The forward column kernel
// then writes 64-byte pieces instead of T-word pieces of separate rows, and
// the row kernel reads its row in 64-byte pieces.  At 2^24 points (tiles of 8
// columns) pairs of rows measured 280 -> 249 us forward columns and 320 ->
// 334 us rows; groups of 4 rows (whole 128-byte lines) 243 and 352 us.
constexpr int SPEC_GROUP_LOG_WORDS = 4;
constexpr int plan_gl(int logn, int v = 0) {
    return cmin(plan_l1(logn, v), cmax(0, SPEC_GROUP_LOG_WORDS - plan_lt(logn, v)));
}
// Whether variant v differs from the default at this size.
constexpr bool plan_distinct(int logn, int v) {
    return v == 0 || plan_l2(logn, v) != plan_l2(logn) || plan_lt(logn, v) != plan_lt(logn) ||
           plan_rb(logn, v) != plan_rb(logn);
}
// Row workgroup: 512 threads; 1024 for 2^15-point rows (32 kept values each)
// and for radix-8 plans of 2^13 points and up (8 kept values each and every
// thread busy in every pass, in at most 64 VGPRs: two workgroups per CU).
// Rows shorter than 2^12 points get N2 / 8 threads (one per radix-8 column,
// at least one wave): with 512, most threads of a short row idled in every
// pass, and a CU held few rows (the 4096-product batch's row launches 6.96 ->
// 3.61 ms).
constexpr int row_threads(int logn, int v = 0) {
    return (plan_l2(logn, v) >= 15 || (plan_rb(logn, v) == 3 && plan_l2(logn, v) >= 13)) ? 1024
           : cmax(64, cmin(512, (1 << plan_l2(logn, v)) >> 3));
}
// Forward column workgroup: one thread per radix-8 column of a tile
// (N1 T / 8), at least one wave and at least T (each thread keeps one column
// of the tile), at most NTT_THREADS.  With 1024 threads for every tile, the
// passes of a 2^12..2^20-point transform left half to seven eighths of the
// threads idle (the 4096-product batch's forward column launches 4.26 ->
// 3.16 ms); tiles of 2^13 elements and more keep 1024.
constexpr int col_threads(int logn, int v = 0) {
    return cmax(cmax(64, 1 << plan_lt(logn, v)), cmin(NTT_THREADS, (1 << (plan_l1(logn, v) + plan_lt(logn, v))) >> 3));
}
// Row kernel with both operands' rows in LDS (2 N2 words): for rows of up to
// 2^13 points (64 KB, two workgroups per CU).
#ifndef PG_ROW_DUAL
#define PG_ROW_DUAL 1
#endif
constexpr bool row_dual(int logn, int v = 0) { return PG_ROW_DUAL && plan_l2(logn, v) <= 13; }
// Minimum waves per SIMD asked of the compiler for the row kernel: 8 (64
// VGPRs, two 1024-thread workgroups per CU) for radix-8 rows of 8 values per
// thread; otherwise no request beyond what the workgroup size needs.
constexpr int row_waves(int logn, int v = 0) {
    return (row_threads(logn, v) == 1024 && plan_rb(logn, v) == 3 && (1 << plan_l2(logn, v)) <= 8 * 1024) ? 8 : 1;
}

// Digit modes of the NTT product.
//  MODE_REF  : preclib's own decomposition (mul_ntt.cpp:331-342, 509-513):
//              base-2^16 digits, primes p1, p2, two-prime Garner CRT.
//  MODE_WIDE : base-2^32 digits over all three primes of the same set
//              (mul_ntt.cpp:29-31).  A coefficient is below
//              2^25 * (2^32 - 1)^2 < 2^89 < p1 p2 p3 (~2^90.47) for every size
//              the NTT path accepts (<= 2^24 limbs), so the three-prime CRT is
//              exact; the transforms are half as long with 3/2 as many primes.
enum { MODE_REF = 0, MODE_WIDE = 1 };
// Forward column kernel grids: (tile, operand x prime), (tile x prime,
// operand) with a tile's primes adjacent on one XCD, (tile x operand, prime)
// with a tile's operands adjacent, and the same with primes last to first.
enum { COLF_OP_PRIME = 0, COLF_TILE_PRIME = 1, COLF_TILE_OP = 2, COLF_TILE_OP_REV = 3, ROW_PRIME_REV = 4 };
constexpr int mode_primes(int mode) { return mode == MODE_WIDE ? 3 : 2; }
// Product dispatch (mul_into, mul_basic.cpp:3-5, 238-284): schoolbook when a
// side is <= 24 limbs or hi <= 192, blocked when hi > 2 lo, else a transform.
// preclib's CPU rule sends hi in 193..6144 (and every square) to mul_fft and
// larger products to mul_ntt.  On gfx950 the NTT takes every transform-sized
// product (squares too) within its caps, and the FP64 FFT runs on request
// (PG_PATH_FFT) and beyond the NTT caps: single products of 200..6144 limbs
// measured 16-28 us on the NTT against 41-61 us on the FFT, and batches of
// such products run 25% faster.  The product is exact on either path.
constexpr size_t DISPATCH_SMALL_SIDE = 24, DISPATCH_SMALL_HI = 192;
// Index of the prime whose transforms run on lazy residues in [0, 2p): the
// set's first prime, 469762049 = 7 * 2^26 + 1, the one below 2^30 (the host
// runtime checks this against its prime table).
constexpr int LAZY_PRIME = 0;
constexpr int mode_digit_bits(int mode) { return mode == MODE_WIDE ? 32 : 16; }

struct CrtK {
    // two-prime Garner (mul_ntt.cpp:405-412): t = (r2 - r1) * p1^-1 mod p2
    uint32_t inv12, inv12q;
    // three-prime Garner: x12 = r1 + p1 t2;  t3 = (r3 - x12) * (p1 p2)^-1 mod p3
    uint32_t p1_mod3, p1_mod3q;       // p1 mod p3 (Shoup)
    uint32_t inv123, inv123q;         // (p1 p2)^-1 mod p3 (Shoup)
    uint32_t red[3];                  // floor(2^32 / p_k): Shoup constants of x * 1 mod p_k
};

// One product ("leaf") of an NTT launch.
struct NttLeaf {
    const uint64_t *a, *b;      // operand limbs (device)
    size_t an, bn;              // limb counts
    uint32_t *spec;             // forward spectra [2 operands][nprimes][n]
    // Operand k's forward column spectra [nprimes][n] when they come from
    // elsewhere (null: computed into spec by this leaf): the same operand
    // (pointer and length) transformed by another leaf of the same launch, or
    // by this leaf's other operand (a square).  The forward column kernel
    // skips such an operand; the row kernel reads its rows from here.
    const uint32_t *ext[2];
    uint32_t *prod;             // product spectra / residues [nprimes][n]
    uint64_t *slo, *shi;        // 128-bit limb sums
    size_t sum_len;             // sums written: limbs [0, sum_len)
    // Wide mode, 2^pack_log limbs per piece (sum_pack_log: half a row of a
    // column tile, at most 8 limbs; 0: plain sums; sum_len then a multiple of
    // the piece).  Packed, limb j's sum S_j = lo_j + 2^64 hi_j is written as
    // slo[j] = lo_j + hi_(j-1) mod 2^64 when j is not the first limb of its
    // piece (slo[j] = lo_j when it is), and shi[j >> pack_log] = hi of the
    // piece's last limb (< 2^57) with bit 56 + q holding the carry out of
    // limb q's addition (q = 1 .. 2^pack_log - 1): the carry then reads
    // about 10 bytes per limb instead of 16.
    int pack_log;
};

// Packed limb sums of a transform with column tiles of 2^lt columns
// (NttLeaf::pack_log): pieces of min(2^lt / 2, 8) limbs in wide mode (the
// carry bits of a piece's limbs 1 .. 7 sit in bits 57 .. 63 of its high
// word), none in reference mode or for tiles of fewer than 4 columns.
constexpr int sum_pack_log(int mode, int lt) { return (mode == MODE_WIDE && lt >= 2) ? cmin(lt, 4) - 1 : 0; }

// A launch covers one leaf (`one`) or, when `leaves` is set, nleaves leaves of
// the same transform length; blockIdx.z selects the leaf.
struct NttLaunch {
    NttLeaf one;
    const NttLeaf *leaves;
    int nleaves;
    double io_bytes;            // operand + product limb bytes of all leaves (timing model)
    int logn, lobits, nprimes, mode;
    int variant;                // plan variant (wide mode), chosen by the autotuner table
    int order;                  // workgroup order: bits 0-1 forward column kernel (COLF_*), bit 2
                                // row kernel primes last to first
    PrimeK pk[3];
    CrtK crt;
    const uint2 *col_tw[3][MAX_PASS], *col_itw[3][MAX_PASS];
    const uint2 *row_tw[3][MAX_PASS], *row_itw[3][MAX_PASS];
    const uint32_t *glo[3], *ghi[3], *ilo[3], *ihi[3];
};

// Enqueues the three transform kernels of the plan for L.logn on `stream`;
// per-stage timing hooks are called around each launch.
struct StageHook {
    void (*begin)(void *ctx, const char *name);
    void (*end)(void *ctx, const char *name, double bytes);
    void *ctx;
};
hipError_t launch_ntt(const NttLaunch &L, hipStream_t stream, const StageHook &hook);
hipError_t ntt_set_lds_limits();

// ---- FP64 FFT product (pg_fft.hip) -----------------------------------------
// Complex points n = 2^logn.  Rows of N2 = 2^fft_l2 points live in LDS (one
// row of complex doubles); columns in tiles of T = 2^fft_lt.
constexpr int FFT_MIN_LOG = 6, FFT_MAX_LOG = 24, FFT_ROW_LOG_MAX = 13;
constexpr int fft_l2(int logn) {
    return logn <= 10 ? logn : cmin(FFT_ROW_LOG_MAX, cmax((logn + 1) / 2, logn - 11));
}
constexpr int fft_l1(int logn) { return logn - fft_l2(logn); }
constexpr int fft_lt(int logn) { return fft_l1(logn) == 0 ? cmin(logn, 6) : cmin(4, 13 - fft_l1(logn)); }
// radix passes of at most 2^4 points (16 complex doubles in registers)
constexpr int fsplit_np(int l) { return l ? (l + 3) / 4 : 0; }
constexpr int fsplit_lr(int l, int i) {
    return fsplit_np(l) == 0 ? 0 : (i < fsplit_np(l) ? l / fsplit_np(l) + (i < l % fsplit_np(l) ? 1 : 0) : 0);
}
constexpr int FFT_THREADS = 512;

struct FftK {
    double2 w16[8], iw16[8];    // exp(+-2 pi i k / 16), k < 8
};

struct FftLeaf {
    const uint64_t *a, *b;      // operand limbs (b may equal a)
    size_t an, bn;
    double2 *spec;              // [n] complex, natural row-major
    uint64_t *slo, *shi;        // limb sums (or half-limb parts)
    size_t sum_len;             // sums written: [0, sum_len)
    uint64_t *maxerr;           // bit pattern of the largest rounding distance (atomicMax)
};

struct FftLaunch {
    FftLeaf one;
    const FftLeaf *leaves;
    int nleaves;
    double io_bytes;
    int logn, bits, lobits;     // digit width 16 or 8
    FftK k;
    const double2 *col_tw[MAX_PASS], *col_itw[MAX_PASS];
    const double2 *row_tw[MAX_PASS], *row_itw[MAX_PASS];
    const double2 *glo, *ghi;   // omega_n^e = glo[e mod 2^lobits] * ghi[e >> lobits]
};
hipError_t launch_fft(const FftLaunch &L, hipStream_t stream, const StageHook &hook);
hipError_t fft_set_lds_limits();

// ---- small products (pg_small.hip) --------------------------------------------
// Schoolbook products below the transform threshold (mul_basic.cpp:131-185,
// mul_kara.cpp; sizes <= 192 limbs, or a short side <= 24 limbs) as limb
// sums.  Column k of x*y is sum_i x_i y_{k-i}; a workgroup computes
// SMALL_COLS + 2 consecutive columns (3 words each) and emits the limb sums
// of SMALL_COLS of them.
constexpr int SMALL_THREADS = 256;
constexpr int SMALL_COLS = SMALL_THREADS - 2;
constexpr int SMALL_MAX_SHORT = 256;    // longest short side held in LDS

struct SmallLeaf {
    const uint64_t *a, *b;      // any order; the kernel iterates over the shorter
    size_t an, bn;
    uint64_t *slo, *shi;
    size_t sum_len;
    size_t blk0;                // first workgroup of this leaf in the launch
};
struct SmallLaunch {
    SmallLeaf one;
    const SmallLeaf *leaves;
    const uint32_t *blk_leaf;   // leaf of every workgroup
};
hipError_t launch_small(const SmallLaunch &L, size_t nblocks, hipStream_t stream, const StageHook &hook,
                        double io_bytes);
inline size_t small_blocks(size_t an, size_t bn) { return (an + bn + SMALL_COLS - 1) / SMALL_COLS; }

// ---- carry propagation (pg_carry.hip) ----------------------------------------
constexpr int CARRY_THREADS = 256;
constexpr int CARRY_PER_THREAD = 8;
constexpr int CARRY_CHUNK = CARRY_THREADS * CARRY_PER_THREAD;

// One value to normalize.  Kinds:
//  plain   : limb j's sum is (slo[j], shi[j]); parts == 2 when a limb is two
//            half-limb parts (FFT tiles narrower than one limb);
//  blocked : limb j's sum is the sum of the block products covering j (at
//            most two: blocks are B = 2 lo limbs apart and at most blen =
//            3 lo long; mul_basic.cpp:207-236).  Block k's sums lie in layer
//            k & 1 (slo/shi, slo1/shi1: arrays of m limbs) at limb k B, so
//            limb j's sum is the layers' words at j where a block covers j;
//            elsewhere the layers are not written.
// The result goes to out[0, out_len) as limbs; with out_hi set it is written
// as sums (out_hi zeroed) for a parent blocked node.
//  linear  : value = xmul * X' + 2^pow2 + (yneg ? ~Y' : Y') + cin over m
//            limbs, where X' / Y' are x / y shifted by xsh / ysh bits (> 0
//            right, < 0 left), limbs outside [0, xn) / [0, yn) read as 0 and
//            ~Y' is taken over the m limbs.  With yneg and cin = 1 this is
//            X' - Y' in two's complement; limb m of the exact sum (1 when the
//            difference is >= 0, the carry-out for an addition) goes to *flag.
//            When cond is set, the Y term and cin apply only if *cond ==
//            cond_eq (otherwise the value is xmul * X' + 2^pow2).
enum { NODE_PLAIN = 0, NODE_BLOCKED = 1, NODE_LIN = 2 };
struct CarryNode {
    int kind, parts;
    const uint64_t *slo, *shi;
    const uint64_t *slo1, *shi1;   // blocked: the odd blocks' layer
    int nblocks;
    size_t B, blen;
    size_t m;                   // limbs covered by the sums
    uint64_t *out, *out_hi;
    size_t out_len;
    size_t chunk0;              // first chunk of this node in the launch
    int pack_log;               // plain, whole-limb: sums packed as NttLeaf::pack_log describes
    // NODE_LIN
    const uint64_t *x, *y;
    size_t xn, yn;
    int64_t xsh, ysh, pow2;     // pow2 < 0: no power-of-two term
    uint64_t xmul;              // < 2^59
    int yneg, cin;
    const uint64_t *cond;
    uint64_t cond_eq;
    uint64_t *flag;
};

struct CarryLaunch {
    CarryNode one;              // used when nodes == nullptr
    const CarryNode *nodes;
    const uint32_t *chunk_node; // node of every chunk
    size_t nchunks;
    uint8_t *agg;               // per-chunk G | P << 1 | used carry-in << 2 | resolved << 3
    size_t nnodes;              // nodes of the launch (1 for `one`; 0: not stated)
    int node_kind;              // CARRY_ANY, or every node linear / plain whole-limb
};
enum { CARRY_ANY = 0, CARRY_LIN = 1, CARRY_PLAIN = 2 };

// Enqueues the carry of C on `stream`: the speculative pass and the fix-up
// pass (pg_carry.hip).  When every chunk is the first of its node
// (C.nchunks == C.nnodes), every carry-in is known in the first pass and the
// fix-up is not launched.
hipError_t launch_carry_kernels(const CarryLaunch &C, hipStream_t stream);

}  // namespace pg
