// pg_kernels.h -- launch descriptors shared by the host runtime and kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include "pg_modarith.h"

namespace pg {

constexpr int NTT_THREADS = 1024;           // workgroup size of the NTT kernels
constexpr int ROW_MAX_ELEMS = 1 << 15;      // longest in-LDS row transform (N2)
constexpr int COL_MAX_ELEMS = 1 << 14;      // largest column tile (N1 * T)
constexpr int MAX_PASS = 4;

// A power-of-two transform done in LDS as radix-2^lr[i] passes.
struct SubPlan {
    int logN;
    int npass;
    int lr[MAX_PASS];
};

struct NttLaunch {
    const uint64_t *a, *b;      // operand limbs (device)
    size_t an, bn;              // limb counts
    uint32_t *spec;             // forward spectra [2 operands][nprimes][n]
    uint32_t *prod;             // product spectra / residues [nprimes][n]
    uint64_t *slo, *shi;        // 128-bit limb sums [n/4]
    int logn, logN2, logT, lobits, nprimes;
    SubPlan col, row;
    PrimeK pk[2];
    const uint2 *col_tw[2][MAX_PASS], *col_itw[2][MAX_PASS];
    const uint2 *row_tw[2][MAX_PASS], *row_itw[2][MAX_PASS];
    const uint32_t *glo[2], *ghi[2], *ilo[2], *ihi[2];
    uint32_t crt_inv, crt_invq;   // p1^-1 mod p2 and its Shoup quotient
};

__global__ void k_col_fwd(const NttLaunch L);
__global__ void k_row(const NttLaunch L);
__global__ void k_col_inv(const NttLaunch L);

// ---- carry propagation (pg_carry.hip) ----------------------------------------
constexpr int CARRY_THREADS = 256;
constexpr int CARRY_PER_THREAD = 8;
constexpr int CARRY_CHUNK = CARRY_THREADS * CARRY_PER_THREAD;

struct CarryLaunch {
    const uint64_t *slo, *shi;  // limb sums, m entries
    size_t m;                   // number of limb sums
    uint64_t *out;              // normalized limbs
    size_t out_len;             // limbs to write (>= m is allowed: zero fill)
    uint8_t *agg;               // per-chunk (G | P << 1)
    uint8_t *cin;               // per-chunk carry-in
    size_t nchunks;
};

__global__ void k_carry_reduce(const CarryLaunch C);
__global__ void k_carry_scan(const CarryLaunch C);
__global__ void k_carry_apply(const CarryLaunch C);

}  // namespace pg
