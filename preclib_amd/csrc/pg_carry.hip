// pg_carry.hip -- parallel carry-lookahead normalization of limb sums.
//
// Replaces the serial carry loop of preclib's ntt_from_residues2
// (src/mul_ntt.cpp:446-456) and fft_from_digits (src/mul_fft.cpp:311-338).
// Input: limb sums S_j = lo_j + 2^64 hi_j with hi_j < 2^60, value
// V = sum_j S_j 2^(64 j).  With t_j = lo_j + hi_{j-1}:
//   g_j = carry out of that 64-bit add,  p_j = (t_j mod 2^64 == 2^64 - 1),
//   x_{j+1} = g_j | (p_j & x_j),        limb_j = t_j + x_j  (mod 2^64).
// g_j = 1 forces t_j < 2^60, so g and p never hold together and x is one bit.
//
// A wave resolves 64 consecutive limbs at once: with X = ballot(g | p) and
// Y = ballot(g), the carry into lane i is bit i of (X + Y + cin) ^ X ^ Y and
// the carry out of the group is the overflow of that 64-bit sum.  Each wave
// walks CARRY_ITERS such groups (coalesced loads, the wave carry chained in
// registers); chunks of CARRY_CHUNK limbs are joined by a per-chunk
// aggregate and a one-workgroup scan.  Three launches: reduce, scan, apply.
// Linear nodes reuse the same scan for device additions, subtractions,
// shifts and small-scalar multiples (the add/sub/shift/mul_u32 steps that
// div_newton.cpp, sqrt.cpp and pi_chudnovsky.cpp run between products).
#include "pg_kernels.h"

namespace pg {

constexpr int CARRY_WAVES = CARRY_THREADS / 64;
constexpr int CARRY_ITERS = CARRY_CHUNK / CARRY_THREADS;

struct Group {
    uint64_t X, Y;   // ballots of (g | p) and g over 64 consecutive limbs
};

__device__ __forceinline__ const CarryNode &node_of(const CarryLaunch &C, size_t chunk) {
    return C.nodes ? C.nodes[C.chunk_node[chunk]] : C.one;
}

// Limb j of v (n limbs, zero beyond) shifted right by sh bits (left for sh < 0).
__device__ __forceinline__ uint64_t shifted_limb(const uint64_t *v, size_t n, int64_t sh, size_t j) {
    if (!v) return 0;
    if (sh >= 0) {
        const size_t i = j + ((size_t)sh >> 6);
        const int r = (int)(sh & 63);
        uint64_t w = i < n ? v[i] : 0;
        if (r) w = (w >> r) | ((i + 1 < n ? v[i + 1] : 0) << (64 - r));
        return w;
    }
    const size_t q = (size_t)(-sh) >> 6;
    const int r = (int)((-sh) & 63);
    uint64_t w = (j >= q && j - q < n) ? v[j - q] : 0;
    if (r) w = (w << r) | ((j >= q + 1 && j - q - 1 < n) ? v[j - q - 1] : 0) >> (64 - r);
    return w;
}

__device__ __forceinline__ bool lin_cond(const CarryNode &N) { return !N.cond || *N.cond == N.cond_eq; }

// Limb sum j of a node as (lo, hi).
__device__ __forceinline__ void limb_sum(const CarryNode &N, size_t j, uint64_t &lo, uint64_t &hi) {
    if (N.kind == NODE_LIN) {
        const uint64_t xv = shifted_limb(N.x, N.xn, N.xsh, j);
        lo = xv * N.xmul;
        hi = __umul64hi(xv, N.xmul);
        if (N.pow2 >= 0 && j == ((size_t)N.pow2 >> 6)) {
            const uint64_t b = 1ull << (N.pow2 & 63);
            lo += b;
            hi += lo < b;
        }
        if (lin_cond(N)) {
            uint64_t yv = shifted_limb(N.y, N.yn, N.ysh, j);
            if (N.yneg) yv = ~yv;
            lo += yv;
            hi += lo < yv;
        }
        return;
    }
    if (N.kind == NODE_PLAIN) {
        if (N.parts == 1) {
            lo = N.slo[j];
            hi = N.shi[j];
        } else {   // two parts starting at bit 0 and bit 32 of the limb
            const uint64_t l0 = N.slo[2 * j], h0 = N.shi[2 * j];
            const uint64_t l1 = N.slo[2 * j + 1], h1 = N.shi[2 * j + 1];
            lo = l0 + (l1 << 32);
            hi = h0 + (l1 >> 32) + (h1 << 32) + (lo < l0);
        }
        return;
    }
    lo = 0;
    hi = 0;
    const size_t k = j / N.B;
    for (int d = 0; d < 2; ++d) {
        if (k < (size_t)d) break;
        const size_t kb = k - d;
        if (kb >= (size_t)N.nblocks) continue;
        const BlockRef &blk = N.blocks[kb];
        const size_t local = j - kb * N.B;
        if (local < blk.len) {
            const uint64_t s = lo + blk.slo[local];
            hi += blk.shi[local] + (s < lo);
            lo = s;
        }
    }
}

__device__ __forceinline__ uint64_t limb_t(const CarryNode &N, size_t j, uint32_t &g, uint32_t &p) {
    uint64_t t, lo, hi, hprev = 0;
    if (j && j <= N.m) limb_sum(N, j - 1, lo, hprev);
    if (j < N.m) {
        limb_sum(N, j, lo, hi);
        t = lo + hprev;
        g = t < lo;
        p = t == ~0ull;
    } else {
        t = j == N.m ? hprev : 0;   // top word of the last sum
        g = 0;
        p = 0;
    }
    return t;
}

// Carry out of a group given carry-in cin (0/1).
__device__ __forceinline__ uint32_t group_out(const Group &G, uint32_t cin) {
    const uint64_t s0 = G.X + G.Y;
    const uint64_t s = s0 + cin;
    return (uint32_t)((s0 < G.X) | (s < s0));
}

// (G, P) aggregate of a sequence of groups, as a function of carry-in.
struct Agg {
    uint32_t g, p;
};
__device__ __forceinline__ Agg group_agg(const Group &G) {
    return Agg{group_out(G, 0), (uint32_t)((G.X ^ G.Y) == ~0ull)};
}
__device__ __forceinline__ Agg compose(Agg lo, Agg hi) {   // lo first, then hi
    return Agg{hi.g | (hi.p & lo.g), hi.p & lo.p};
}

__global__ void __launch_bounds__(CARRY_THREADS) k_carry_reduce(const CarryLaunch C) {
    __shared__ Agg waves[CARRY_WAVES];
    const CarryNode &N = node_of(C, blockIdx.x);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const size_t base = (blockIdx.x - N.chunk0) * CARRY_CHUNK + (size_t)wave * (64 * CARRY_ITERS) + lane;
    Agg a{0, 1};
#pragma unroll
    for (int it = 0; it < CARRY_ITERS; ++it) {
        uint32_t g, p;
        limb_t(N, base + 64 * it, g, p);
        const Group G{__ballot(g | p), __ballot(g)};
        a = compose(a, group_agg(G));
    }
    if (lane == 0) waves[wave] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        Agg b{0, 1};
        for (int i = 0; i < CARRY_WAVES; ++i) b = compose(b, waves[i]);
        C.agg[blockIdx.x] = (uint8_t)(b.g | (b.p << 1));
    }
}

// One workgroup: exclusive scan of chunk aggregates into chunk carry-ins.
// The first chunk of every node starts with carry 0 (a "kill" before it).
__device__ __forceinline__ bool node_start(const CarryLaunch &C, size_t i) {
    return C.nodes ? C.nodes[C.chunk_node[i]].chunk0 == i : i == 0;
}
// Carry into the lowest limb of the node that starts at chunk i.
__device__ __forceinline__ uint32_t node_cin(const CarryLaunch &C, size_t i) {
    const CarryNode &N = node_of(C, i);
    return (N.kind == NODE_LIN && N.cin && lin_cond(N)) ? 1u : 0u;
}

__global__ void __launch_bounds__(1024) k_carry_scan(const CarryLaunch C) {
    __shared__ Agg waves[16];
    __shared__ uint32_t wave_in[16];
    const size_t per = (C.nchunks + 1023) / 1024;
    const size_t lo = (size_t)threadIdx.x * per;
    Agg a{0, 1};
    for (size_t i = lo; i < lo + per && i < C.nchunks; ++i) {
        const uint32_t v = C.agg[i];
        if (node_start(C, i)) a = Agg{node_cin(C, i), 0};
        a = compose(a, Agg{v & 1, v >> 1});
    }
    // thread-level (G, P) pairs -> carries via the wave sum trick
    const Group G{__ballot(a.g | a.p), __ballot(a.g)};
    if ((threadIdx.x & 63) == 0) waves[threadIdx.x >> 6] = group_agg(G);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t x = 0;
        for (int i = 0; i < 16; ++i) {
            wave_in[i] = x;
            x = waves[i].g | (waves[i].p & x);
        }
    }
    __syncthreads();
    const uint64_t sum = G.X + G.Y + wave_in[threadIdx.x >> 6];
    uint32_t x = (uint32_t)(((sum ^ G.X ^ G.Y) >> (threadIdx.x & 63)) & 1);
    for (size_t i = lo; i < lo + per && i < C.nchunks; ++i) {
        if (node_start(C, i)) x = node_cin(C, i);
        C.cin[i] = (uint8_t)x;
        const uint32_t v = C.agg[i];
        x = (v & 1) | ((v >> 1) & x);
    }
}

__global__ void __launch_bounds__(CARRY_THREADS) k_carry_apply(const CarryLaunch C) {
    __shared__ Agg waves[CARRY_WAVES];
    __shared__ uint32_t wave_in[CARRY_WAVES];
    const CarryNode &N = node_of(C, blockIdx.x);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const size_t base = (blockIdx.x - N.chunk0) * CARRY_CHUNK + (size_t)wave * (64 * CARRY_ITERS) + lane;
    uint64_t t[CARRY_ITERS];
    Group G[CARRY_ITERS];
    Agg a{0, 1};
#pragma unroll
    for (int it = 0; it < CARRY_ITERS; ++it) {
        uint32_t g, p;
        t[it] = limb_t(N, base + 64 * it, g, p);
        G[it] = Group{__ballot(g | p), __ballot(g)};
        a = compose(a, group_agg(G[it]));
    }
    if (lane == 0) waves[wave] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t x = C.cin[blockIdx.x];
        for (int i = 0; i < CARRY_WAVES; ++i) {
            wave_in[i] = x;
            x = waves[i].g | (waves[i].p & x);
        }
    }
    __syncthreads();
    uint32_t cin = wave_in[wave];
#pragma unroll
    for (int it = 0; it < CARRY_ITERS; ++it) {
        const uint64_t sum = G[it].X + G[it].Y + cin;
        const uint32_t x = (uint32_t)(((sum ^ G[it].X ^ G[it].Y) >> lane) & 1);
        const size_t j = base + 64 * it;
        if (j < N.out_len) {
            N.out[j] = t[it] + x;
            if (N.out_hi) N.out_hi[j] = 0;
        }
        if (j == N.m && N.flag) *N.flag = t[it] + x;
        cin = group_out(G[it], cin);
    }
}

}  // namespace pg
