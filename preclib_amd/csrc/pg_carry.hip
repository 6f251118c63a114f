// pg_carry.hip -- parallel carry-lookahead normalization of limb sums.
//
// Replaces the serial carry loop of preclib's ntt_from_residues2
// (src/mul_ntt.cpp:446-456) and fft_from_digits (src/mul_fft.cpp:311-338).
// Input: limb sums S_j = lo_j + 2^64 hi_j with hi_j < 2^50, value
// V = sum_j S_j 2^(64 j).  With t_j = lo_j + hi_{j-1}:
//   g_j = carry out of that 64-bit add,  p_j = (t_j mod 2^64 == 2^64 - 1),
//   x_{j+1} = g_j | (p_j & x_j),        limb_j = t_j + x_j  (mod 2^64).
// g_j = 1 forces t_j < 2^50, so g and p never hold together and x is one bit.
// (g, p) pairs compose associatively; a wave resolves 64 of them at once as
// the carries of the 64-bit sum X + Y + cin with X = ballot(g|p), Y = ballot(g)
// (carry into lane i = bit i of (X + Y + cin) ^ X ^ Y).  Three launches:
// per-chunk reduce, one-workgroup scan of chunk aggregates, apply.
#include "pg_kernels.h"

namespace pg {

struct GP {
    uint32_t g, p;
};

__device__ __forceinline__ void limb_tgp(const CarryLaunch &C, size_t j, uint64_t &t, uint32_t &g,
                                         uint32_t &p) {
    if (j < C.m) {
        const uint64_t lo = C.slo[j];
        const uint64_t h = j ? C.shi[j - 1] : 0;
        t = lo + h;
        g = t < lo;
        p = t == ~0ull;
    } else {
        t = (j == C.m && j) ? C.shi[j - 1] : 0;   // top carry word of the sums
        g = 0;
        p = 0;
    }
}

// Thread aggregate over its CARRY_PER_THREAD limbs (low to high).
__device__ __forceinline__ GP thread_gp(const CarryLaunch &C, size_t base, uint64_t *t, uint32_t *g,
                                        uint32_t *p) {
    GP a{0, 1};
#pragma unroll
    for (int i = 0; i < CARRY_PER_THREAD; ++i) {
        limb_tgp(C, base + i, t[i], g[i], p[i]);
        a.g = g[i] | (p[i] & a.g);
        a.p &= p[i];
    }
    return a;
}

// Wave-level: carry into each lane given the wave's carry-in, and the wave
// aggregate.
__device__ __forceinline__ uint32_t wave_carry(GP a, uint32_t wave_cin, GP *wave_agg) {
    const uint64_t X = __ballot(a.g | a.p);
    const uint64_t Y = __ballot(a.g);
    const uint64_t s0 = X + Y;
    const uint64_t c0 = s0 < X;
    const uint64_t s = s0 + wave_cin;
    const uint64_t c1 = s < s0;
    const uint64_t carries = s ^ X ^ Y;
    if (wave_agg) {
        wave_agg->g = (uint32_t)(c0);            // carry out with cin = 0
        wave_agg->p = (uint32_t)((X ^ Y) == ~0ull);
    }
    (void)c1;
    const int lane = threadIdx.x & 63;
    return (uint32_t)((carries >> lane) & 1);
}

__global__ void __launch_bounds__(CARRY_THREADS) k_carry_reduce(const CarryLaunch C) {
    __shared__ GP waves[CARRY_THREADS / 64];
    const size_t base = (size_t)blockIdx.x * CARRY_CHUNK + (size_t)threadIdx.x * CARRY_PER_THREAD;
    uint64_t t[CARRY_PER_THREAD];
    uint32_t g[CARRY_PER_THREAD], p[CARRY_PER_THREAD];
    GP a = thread_gp(C, base, t, g, p);
    GP w;
    wave_carry(a, 0, &w);
    if ((threadIdx.x & 63) == 0) waves[threadIdx.x >> 6] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        GP b{0, 1};
        for (int i = 0; i < CARRY_THREADS / 64; ++i) {
            b.g = waves[i].g | (waves[i].p & b.g);
            b.p &= waves[i].p;
        }
        C.agg[blockIdx.x] = (uint8_t)(b.g | (b.p << 1));
    }
}

// One workgroup: exclusive scan of chunk aggregates into chunk carry-ins.
__global__ void __launch_bounds__(1024) k_carry_scan(const CarryLaunch C) {
    __shared__ GP waves[16];
    __shared__ uint32_t wave_in[16];
    const size_t per = (C.nchunks + 1023) / 1024;
    const size_t lo = (size_t)threadIdx.x * per;
    GP a{0, 1};
    for (size_t i = lo; i < lo + per && i < C.nchunks; ++i) {
        const uint32_t v = C.agg[i];
        a.g = (v & 1) | ((v >> 1) & a.g);
        a.p &= (v >> 1);
    }
    GP w;
    wave_carry(a, 0, &w);
    if ((threadIdx.x & 63) == 0) waves[threadIdx.x >> 6] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t x = 0;
        for (int i = 0; i < 16; ++i) {
            wave_in[i] = x;
            x = waves[i].g | (waves[i].p & x);
        }
    }
    __syncthreads();
    uint32_t x = wave_carry(a, wave_in[threadIdx.x >> 6], nullptr);
    for (size_t i = lo; i < lo + per && i < C.nchunks; ++i) {
        C.cin[i] = (uint8_t)x;
        const uint32_t v = C.agg[i];
        x = (v & 1) | ((v >> 1) & x);
    }
}

__global__ void __launch_bounds__(CARRY_THREADS) k_carry_apply(const CarryLaunch C) {
    __shared__ GP waves[CARRY_THREADS / 64];
    __shared__ uint32_t wave_in[CARRY_THREADS / 64];
    const size_t base = (size_t)blockIdx.x * CARRY_CHUNK + (size_t)threadIdx.x * CARRY_PER_THREAD;
    uint64_t t[CARRY_PER_THREAD];
    uint32_t g[CARRY_PER_THREAD], p[CARRY_PER_THREAD];
    GP a = thread_gp(C, base, t, g, p);
    GP w;
    wave_carry(a, 0, &w);
    if ((threadIdx.x & 63) == 0) waves[threadIdx.x >> 6] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t x = C.cin[blockIdx.x];
        for (int i = 0; i < CARRY_THREADS / 64; ++i) {
            wave_in[i] = x;
            x = waves[i].g | (waves[i].p & x);
        }
    }
    __syncthreads();
    uint32_t x = wave_carry(a, wave_in[threadIdx.x >> 6], nullptr);
#pragma unroll
    for (int i = 0; i < CARRY_PER_THREAD; ++i) {
        const size_t j = base + i;
        if (j < C.out_len) C.out[j] = t[i] + x;
        x = g[i] | (p[i] & x);
    }
}

}  // namespace pg
