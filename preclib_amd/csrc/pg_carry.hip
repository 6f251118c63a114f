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
// registers); chunks of CARRY_CHUNK limbs take their carry-in from the limbs
// just below them and are, where that window propagates throughout, fixed
// up from the chunk aggregates (see k_carry_spec).
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
__device__ __forceinline__ uint64_t shifted_limb(const uint64_t *vp, size_t n, int64_t sh, size_t j) {
    if (!vp) return 0;
    const auto v = gptr(vp);
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

__device__ __forceinline__ bool lin_cond(const CarryNode &N) { return !N.cond || *gptr(N.cond) == N.cond_eq; }

// Which layers of a blocked node hold a block product covering limb j
// (0 <= j < m): block k = j / B (layer k & 1) whenever it exists, and block
// k - 1 while j is below its end (k - 1 the last block: always, as j < m;
// otherwise while j - k B < blen - B).
__device__ __forceinline__ void blocked_cover(const CarryNode &N, size_t j, bool &c0, bool &c1) {
    const size_t k = j / N.B, r = j - k * N.B;
    const bool ck = k < (size_t)N.nblocks;
    const bool cp = k >= 1 && (!ck || r < N.blen - N.B);
    c0 = (k & 1) ? cp : ck;
    c1 = (k & 1) ? ck : cp;
}

// Limb sum j (0 <= j < m) of a blocked node: the covering layers' words.
__device__ __forceinline__ void blocked_sum(const CarryNode &N, size_t j, uint64_t &lo, uint64_t &hi) {
    bool c0, c1;
    blocked_cover(N, j, c0, c1);
    const uint64_t l0 = c0 ? gptr(N.slo)[j] : 0, h0 = c0 ? gptr(N.shi)[j] : 0;
    const uint64_t l1 = c1 ? gptr(N.slo1)[j] : 0, h1 = c1 ? gptr(N.shi1)[j] : 0;
    lo = l0 + l1;
    hi = h0 + h1 + (lo < l0);
}

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
        const auto slo = gptr(N.slo), shi = gptr(N.shi);
        if (N.parts == 1) {
            lo = slo[j];
            hi = shi[j];
        } else {   // two parts starting at bit 0 and bit 32 of the limb
            const uint64_t l0 = slo[2 * j], h0 = shi[2 * j];
            const uint64_t l1 = slo[2 * j + 1], h1 = shi[2 * j + 1];
            lo = l0 + (l1 << 32);
            hi = h0 + (l1 >> 32) + (h1 << 32) + (lo < l0);
        }
        return;
    }
    blocked_sum(N, j, lo, hi);
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

// The first chunk of every node starts with its node's carry-in.
__device__ __forceinline__ bool node_start(const CarryLaunch &C, size_t i) {
    return C.nodes ? C.nodes[C.chunk_node[i]].chunk0 == i : i == 0;
}
// Carry into the lowest limb of the node that starts at chunk i.
__device__ __forceinline__ uint32_t node_cin(const CarryLaunch &C, size_t i) {
    const CarryNode &N = node_of(C, i);
    return (N.kind == NODE_LIN && N.cin && lin_cond(N)) ? 1u : 0u;
}

// ---- speculative form: carry-in from the limbs just below each chunk --------
//
// The carry into a chunk is the carry out of the 64 limbs below it unless all
// 64 propagate: then the nearest limb below that generates or kills decides.
// k_carry_spec computes that window with the chunk itself and applies the
// carry it gives, so every chunk is written in one streaming pass with no
// dependence on other workgroups.  A chunk whose window propagates
// throughout is written with carry-in 0 and marked unresolved; k_carry_fix
// then finds each unresolved chunk's true carry-in by walking back over the
// chunk aggregates to the nearest chunk whose carry-out is known (generate,
// kill, or resolved with its carry-in), and adds the missing 1 where it is 1
// (a +1 that ripples through the chunk's propagating prefix).
//
// agg[i] = g | p << 1 | used carry-in << 2 | resolved << 3.

// Limbs t_j, j = base + 64 it, of a single plain node with whole-limb sums
// (the product path).  Every load is issued before the first is used: the
// indices are clamped into [0, m) and the out-of-range words discarded after.
// Packed sums (NttLeaf::pack_log): slo[j] is t_j itself except at the first
// limb of a piece, which adds the high word of the piece before; the carry
// bits of a piece's other limbs sit above that high word.
template <int IT>
__device__ __forceinline__ void limbs_t_packed(const CarryNode &N, long base, uint64_t (&t)[IT],
                                               uint32_t (&g)[IT], uint32_t (&p)[IT]) {
    const auto tv = gptr(N.slo), aux = gptr(N.shi);
    const long m = (long)N.m;
    const int pl = N.pack_log;
    constexpr uint64_t HI = (1ull << 57) - 1;   // high words are below 2^57; carry bits are 57 and up
    uint64_t x[IT], a[IT];
    if (m) {
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const long j = base + 64 * it;
            const long jl = j < 0 ? 0 : (j < m ? j : m - 1);
            const long pc = jl >> pl, q = jl & ((1l << pl) - 1);
            x[it] = tv[jl];
            a[it] = aux[q || pc == 0 ? pc : pc - 1];
        }
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const long j = base + 64 * it;
        const long q = j & ((1l << pl) - 1);
        if (j >= 0 && j < m) {
            if (q) {
                t[it] = x[it];
                g[it] = (uint32_t)((a[it] >> (56 + q)) & 1);
            } else {
                const uint64_t h = j ? (a[it] & HI) : 0;
                t[it] = x[it] + h;
                g[it] = t[it] < x[it];
            }
            p[it] = t[it] == ~0ull;
        } else {
            t[it] = (j >= 0 && j == m) ? (a[it] & HI) : 0;   // top word: the last high word
            g[it] = 0;
            p[it] = 0;
        }
    }
}

template <int IT>
__device__ __forceinline__ void limbs_t_plain(const CarryNode &N, long base, uint64_t (&t)[IT],
                                              uint32_t (&g)[IT], uint32_t (&p)[IT]) {
    if (N.pack_log) {
        limbs_t_packed<IT>(N, base, t, g, p);
        return;
    }
    const auto slo = gptr(N.slo), shi = gptr(N.shi);
    const long m = (long)N.m;
    uint64_t lo[IT], hp[IT];
    if (m) {
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const long j = base + 64 * it;
            const long jl = j < 0 ? 0 : (j < m ? j : m - 1);
            const long jh = j < 1 ? 0 : (j - 1 < m ? j - 1 : m - 1);
            lo[it] = slo[jl];
            hp[it] = shi[jh];
        }
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const long j = base + 64 * it;
        const uint64_t hprev = (j >= 1 && j <= m) ? hp[it] : 0;
        if (j >= 0 && j < m) {
            t[it] = lo[it] + hprev;
            g[it] = t[it] < lo[it];
            p[it] = t[it] == ~0ull;
        } else {
            t[it] = (j >= 0 && j == m) ? hprev : 0;
            g[it] = 0;
            p[it] = 0;
        }
    }
}

// 64-bit value of lane (this lane - 1); lane 0 reads 0.
__device__ __forceinline__ uint64_t lane_up1(uint64_t v) {
    const int lane = threadIdx.x & 63;
    const int src = (lane ? lane - 1 : 0) << 2;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)(uint32_t)(v >> 32));
    return lane ? ((uint64_t)hi << 32 | lo) : 0;
}

// Lane 63's value of v (uniform).
__device__ __forceinline__ uint64_t lane63(uint64_t v) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 63);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), 63);
    return (uint64_t)hi << 32 | lo;
}

// t_j, g_j, p_j of limbs j = base + 64 it (lane l of the wave at base0 + l)
// from their sums (lo, hi; 0 for j outside [0, m)): the high word of limb
// j - 1 comes from the lane below (lane 0: lane 63 of the previous iteration,
// or for the first, `below`, the high word of limb base0 - 1).
template <int IT>
__device__ __forceinline__ void lanes_t(const uint64_t (&lo)[IT], const uint64_t (&hi)[IT], uint64_t below, long base,
                                        long m, uint64_t (&t)[IT], uint32_t (&g)[IT], uint32_t (&p)[IT]) {
    const int lane = threadIdx.x & 63;
    uint64_t hprev_last = below;
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const long j = base + 64 * it;
        const uint64_t up = lane_up1(hi[it]);
        const uint64_t hprev = lane ? up : hprev_last;   // limb j - 1 (0 beyond m)
        hprev_last = lane63(hi[it]);
        if (j < m) {
            t[it] = lo[it] + hprev;
            g[it] = t[it] < lo[it];
            p[it] = t[it] == ~0ull;
        } else {
            t[it] = j == m ? hprev : 0;
            g[it] = 0;
            p[it] = 0;
        }
    }
}

// Limbs t_j, j = base + 64 it (base >= 0, lane l of a wave at base0 + l), of
// a blocked node.  Every lane loads both layers' words at its j (indices
// clamped into [0, m), words no block covers discarded after), so the loads
// of all iterations are in flight together; the high word of limb j - 1
// comes from the lane below (lane 0: lane 63 of the previous iteration, or
// for the first, the limb below the wave, the same for every lane).
template <int IT>
__device__ __forceinline__ void limbs_t_blocked(const CarryNode &N, long base, uint64_t (&t)[IT],
                                                uint32_t (&g)[IT], uint32_t (&p)[IT]) {
    const long m = (long)N.m;
    const auto a_lo = gptr(N.slo), a_hi = gptr(N.shi), b_lo = gptr(N.slo1), b_hi = gptr(N.shi1);
    const int lane = threadIdx.x & 63;
    uint64_t l0[IT], h0[IT], l1[IT], h1[IT];
    if (m) {
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const long j = base + 64 * it;
            const long jl = j < m ? j : m - 1;
            l0[it] = a_lo[jl];
            h0[it] = a_hi[jl];
            l1[it] = b_lo[jl];
            h1[it] = b_hi[jl];
        }
    }
    uint64_t below = 0;   // high word of the limb below the wave's first
    {
        const long jb = base - lane - 1;
        if (jb >= 0 && jb < m) {
            uint64_t lo;
            blocked_sum(N, (size_t)jb, lo, below);
        }
    }
    uint64_t hprev_last = below;
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const long j = base + 64 * it;
        uint64_t lo = 0, hi = 0;
        if (j < m) {
            bool c0, c1;
            blocked_cover(N, (size_t)j, c0, c1);
            const uint64_t x0 = c0 ? l0[it] : 0, x1 = c1 ? l1[it] : 0;
            lo = x0 + x1;
            hi = (c0 ? h0[it] : 0) + (c1 ? h1[it] : 0) + (lo < x0);
        }
        const uint64_t up = lane_up1(hi);
        const uint64_t hprev = lane ? up : hprev_last;   // limb j - 1 (0 beyond m)
        hprev_last = lane63(hi);
        if (j < m) {
            t[it] = lo + hprev;
            g[it] = t[it] < lo;
            p[it] = t[it] == ~0ull;
        } else {
            t[it] = j == m ? hprev : 0;
            g[it] = 0;
            p[it] = 0;
        }
    }
}

// The two words shifted_limb(v, n, sh, j) combines (a: the word shifted
// toward bit 0's side, b: its neighbour), as indices into v; -1 where the
// word lies outside [0, n).
__device__ __forceinline__ void shift_words(long n, int64_t sh, long j, long &ia, long &ib) {
    if (sh >= 0) {
        ia = j + (long)((uint64_t)sh >> 6);
        ib = ia + 1;
    } else {
        ia = j - (long)((uint64_t)(-sh) >> 6);
        ib = ia - 1;
    }
    if (ia < 0 || ia >= n) ia = -1;
    if (ib < 0 || ib >= n) ib = -1;
}

__device__ __forceinline__ uint64_t shift_combine(uint64_t a, uint64_t b, int64_t sh) {
    if (sh >= 0) {
        const int r = (int)(sh & 63);
        return r ? (a >> r) | (b << (64 - r)) : a;
    }
    const int r = (int)((-sh) & 63);
    return r ? (a << r) | (b >> (64 - r)) : a;
}

// shifted_limb(v, n, sh, j) for limbs j = base + 64 it: both words of every
// iteration loaded before any is used (indices clamped, words outside
// [0, n) read as 0), then combined.
template <int IT>
__device__ __forceinline__ void shifted_limbs(const uint64_t *vp, long n, int64_t sh, long base, uint64_t (&w)[IT]) {
    if (!vp || !n) {
#pragma unroll
        for (int it = 0; it < IT; ++it) w[it] = 0;
        return;
    }
    const auto v = gptr(vp);
    uint64_t a[IT], b[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        long ia, ib;
        shift_words(n, sh, base + 64 * it, ia, ib);
        a[it] = v[ia < 0 ? 0 : ia];
        b[it] = v[ib < 0 ? 0 : ib];
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        long ia, ib;
        shift_words(n, sh, base + 64 * it, ia, ib);
        w[it] = shift_combine(ia < 0 ? 0 : a[it], ib < 0 ? 0 : b[it], sh);
    }
}

// Limbs t_j, j = base + 64 it (base >= 0, lane l of a wave at base0 + l), of
// a linear node (limb_sum's arithmetic): both operands' words for every
// iteration are loaded before any is used.
template <int IT>
__device__ __forceinline__ void limbs_t_lin(const CarryNode &N, long base, uint64_t (&t)[IT], uint32_t (&g)[IT],
                                            uint32_t (&p)[IT]) {
    const long m = (long)N.m;
    const int lane = threadIdx.x & 63;
    const bool ycond = lin_cond(N);   // (with y null the term is 0, or ~0 when yneg)
    uint64_t xv[IT], yv[IT];
    shifted_limbs<IT>(N.x, (long)N.xn, N.xsh, base, xv);
    shifted_limbs<IT>(ycond ? N.y : nullptr, (long)N.yn, N.ysh, base, yv);
    uint64_t below = 0;   // high word of the limb below the wave's first
    {
        const long jb = base - lane - 1;
        if (jb >= 0 && jb < m) {
            uint64_t lo;
            limb_sum(N, (size_t)jb, lo, below);
        }
    }
    uint64_t lo[IT], hi[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const long j = base + 64 * it;
        lo[it] = 0;
        hi[it] = 0;
        if (j < m) {
            uint64_t l = xv[it] * N.xmul, h = __umul64hi(xv[it], N.xmul);
            if (N.pow2 >= 0 && j == (long)((uint64_t)N.pow2 >> 6)) {
                const uint64_t bit = 1ull << (N.pow2 & 63);
                l += bit;
                h += l < bit;
            }
            if (ycond) {
                const uint64_t y = N.yneg ? ~yv[it] : yv[it];
                l += y;
                h += l < y;
            }
            lo[it] = l;
            hi[it] = h;
        }
    }
    lanes_t<IT>(lo, hi, below, base, m, t, g, p);
}

// Limbs t_j, j = base + 64 it, of node N.  Launches with a node table take
// the product form for plain whole-limb nodes, the layer form for blocked ones
// and the shifted-word form for linear ones (every load of a thread issued
// before the first is used; blocked and linear nodes take limb j - 1's high
// word from the lane below); the node is uniform across the workgroup.
template <bool PLAIN, int NK, int IT>
__device__ __forceinline__ void chunk_limbs(const CarryNode &N, long base, uint64_t (&t)[IT], uint32_t (&g)[IT],
                                            uint32_t (&p)[IT]) {
    if (NK == CARRY_LIN && base - (long)(threadIdx.x & 63) >= 0) {   // linear nodes only (uniform per wave)
        limbs_t_lin<IT>(N, base, t, g, p);
    } else if (PLAIN || NK == CARRY_PLAIN || (N.kind == NODE_PLAIN && N.parts == 1)) {
        limbs_t_plain<IT>(N, base, t, g, p);
    } else if (N.kind == NODE_BLOCKED && base - (long)(threadIdx.x & 63) >= 0) {   // (uniform per wave)
        limbs_t_blocked<IT>(N, base, t, g, p);
    } else {
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const long j = base + 64 * it;
            if (j >= 0) {
                t[it] = limb_t(N, (size_t)j, g[it], p[it]);
            } else {
                t[it] = 0;
                g[it] = 0;
                p[it] = 0;
            }
        }
    }
}

// PLAIN: C.one is a plain node with whole-limb sums and there is no node table.
template <bool PLAIN, int NK = CARRY_ANY>
__global__ void __launch_bounds__(CARRY_THREADS) k_carry_spec(const CarryLaunch C) {
    __shared__ Agg waves[CARRY_WAVES];
    __shared__ uint32_t wave_in[CARRY_WAVES];
    const size_t chunk = blockIdx.x;
    const CarryNode &N = PLAIN ? C.one : node_of(C, chunk);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long cbase = (long)(chunk - N.chunk0) * CARRY_CHUNK;
    const long base = cbase + (long)wave * (64 * CARRY_ITERS) + lane;
    const bool start = PLAIN ? chunk == 0 : node_start(C, chunk);
    // wave 0 also reads the window: the 64 limbs below the chunk
    constexpr int IT1 = CARRY_ITERS + 1;
    uint64_t t[IT1];
    uint32_t g[IT1], p[IT1];
    if (wave == 0 && !start) {
        chunk_limbs<PLAIN, NK, IT1>(N, base - 64, t, g, p);
    } else {
        uint64_t t0[CARRY_ITERS];
        uint32_t g0[CARRY_ITERS], p0[CARRY_ITERS];
        chunk_limbs<PLAIN, NK, CARRY_ITERS>(N, base, t0, g0, p0);
        t[0] = 0; g[0] = 0; p[0] = 0;
#pragma unroll
        for (int it = 0; it < CARRY_ITERS; ++it) {
            t[it + 1] = t0[it];
            g[it + 1] = g0[it];
            p[it + 1] = p0[it];
        }
    }
    Group G[CARRY_ITERS];
    Agg a{0, 1};
#pragma unroll
    for (int it = 0; it < CARRY_ITERS; ++it) {
        G[it] = Group{__ballot(g[it + 1] | p[it + 1]), __ballot(g[it + 1])};
        a = compose(a, group_agg(G[it]));
    }
    if (lane == 0) waves[wave] = a;
    __syncthreads();
    if (threadIdx.x < 64) {
        uint32_t cin, resolved = 1;
        if (start) {
            cin = PLAIN ? 0u : node_cin(C, chunk);
        } else {
            // window: carry out of the 64 limbs below, if some limb there does
            // not propagate (the highest such limb decides)
            const uint64_t np = __ballot(!p[0]), gb = __ballot(g[0]);
            if (np) {
                const int k = 63 - __clzll((long long)np);
                cin = (uint32_t)((gb >> k) & 1);
            } else {
                cin = 0;
                resolved = 0;
            }
        }
        if (threadIdx.x == 0) {
            Agg b{0, 1};
            uint32_t x = cin;
            for (int i = 0; i < CARRY_WAVES; ++i) {
                wave_in[i] = x;
                x = waves[i].g | (waves[i].p & x);
                b = compose(b, waves[i]);
            }
            C.agg[chunk] = (uint8_t)(b.g | (b.p << 1) | (cin << 2) | (resolved << 3));
        }
    }
    __syncthreads();
    uint32_t cin = wave_in[wave];
#pragma unroll
    for (int it = 0; it < CARRY_ITERS; ++it) {
        const uint64_t sum = G[it].X + G[it].Y + cin;
        const uint32_t x = (uint32_t)(((sum ^ G[it].X ^ G[it].Y) >> lane) & 1);
        const long j = base + 64 * it;
        if ((size_t)j < N.out_len) {
            gptr_mut(N.out)[j] = t[it + 1] + x;
            if (N.out_hi) gptr_mut(N.out_hi)[j] = 0;
        }
        if ((size_t)j == N.m && N.flag) *gptr_mut(N.flag) = t[it + 1] + x;
        cin = group_out(G[it], cin);
    }
}

// True carry-in of an unresolved chunk, by one wave: the carry out of the
// nearest predecessor whose carry-out does not depend on its own carry-in
// (it generates, it kills, or it was resolved and so used its exact carry-in)
// passes unchanged through the propagating chunks between.  The walk ends at
// the latest at the first chunk of the node, which is always resolved.
__device__ uint32_t fix_carry_in(const CarryLaunch &C, size_t chunk) {
    const int lane = threadIdx.x & 63;
    long top = (long)chunk - 1;
    for (;;) {
        const long i = top - lane;
        uint32_t d = 1, val = 0;
        if (i >= 0) {
            const uint32_t a = C.agg[i];
            const uint32_t g = a & 1, p = (a >> 1) & 1, cu = (a >> 2) & 1, res = (a >> 3) & 1;
            d = g | (p ^ 1) | res;
            val = g ? 1u : (p ? cu : 0u);
        }
        const uint64_t dec = __ballot(d);
        if (dec) return (uint32_t)__shfl((int)val, __ffsll((unsigned long long)dec) - 1);
        top -= 64;
    }
}

// One wave per chunk: an unresolved chunk whose true carry-in is 1 was
// written with 0; add 1 at its lowest limb, rippling through the limbs that
// propagate (written as 2^64 - 1, now 0) into the first that does not
// (written as t, now t + 1).  Limbs past out_len are not stored but still
// carry the ripple; limb m is the flag word.  A chunk past out_len that does
// not hold the flag word stores nothing and returns at once (the look-backs of
// other chunks read only the first pass's aggregates): a difference y - p
// whose top limbs cancel, computed for its flag over all of y but stored in
// fewer limbs, otherwise walked back and rippled through every chunk of its
// propagating top part.
template <bool PLAIN, int NK = CARRY_ANY>
__global__ void __launch_bounds__(CARRY_THREADS) k_carry_fix(const CarryLaunch C) {
    const size_t chunk = (size_t)blockIdx.x * CARRY_WAVES + (threadIdx.x >> 6);
    if (chunk >= C.nchunks) return;
    const uint32_t v = C.agg[chunk];
    if (v & 8) return;
    const CarryNode &N = PLAIN ? C.one : node_of(C, chunk);
    const long cbase = (long)(chunk - N.chunk0) * CARRY_CHUNK;
    if (cbase >= (long)N.out_len && !(N.flag && (long)N.m >= cbase && (long)N.m < cbase + CARRY_CHUNK)) return;
    if (!fix_carry_in(C, chunk)) return;
    const int lane = threadIdx.x & 63;
    if (v & 2) {
        // every limb of the chunk propagates (t = 2^64 - 1; limb m, whose p is
        // 0, is not among them): with the carry in, each becomes 0 - stored
        // without reading the limbs again
        const long lim = cbase + CARRY_CHUNK < (long)N.out_len ? cbase + CARRY_CHUNK : (long)N.out_len;
        for (long j = cbase + lane; j < lim; j += 64) gptr_mut(N.out)[j] = 0;
        return;
    }
    for (int grp = 0; grp < CARRY_CHUNK / 64; ++grp) {
        const long j = cbase + 64 * grp + lane;
        uint64_t t[1];
        uint32_t g[1], p[1];
        chunk_limbs<PLAIN, NK, 1>(N, j, t, g, p);
        const uint64_t stop = __ballot(!p[0]);
        const int k = stop ? __ffsll((unsigned long long)stop) - 1 : 64;
        if (lane <= k) {   // k == 64: the whole group propagates
            const uint64_t val = t[0] + 1;
            if ((size_t)j < N.out_len) gptr_mut(N.out)[j] = val;
            if ((size_t)j == N.m && N.flag) *gptr_mut(N.flag) = val;
        }
        if (stop) return;
    }
}

hipError_t launch_carry_kernels(const CarryLaunch &C, hipStream_t stream) {
    if (!C.nchunks) return hipSuccess;
    if (!C.agg) return hipErrorInvalidValue;
    const unsigned nch = (unsigned)C.nchunks, nfix = (unsigned)((C.nchunks + CARRY_WAVES - 1) / CARRY_WAVES);
    // every chunk a node start: all carry-ins are exact in the first pass
    const bool one_pass = C.nnodes == C.nchunks;
    if (!C.nodes && C.one.kind == NODE_PLAIN && C.one.parts == 1) {
        hipLaunchKernelGGL(k_carry_spec<true>, dim3(nch), dim3(CARRY_THREADS), 0, stream, C);
        if (!one_pass)
            hipLaunchKernelGGL(k_carry_fix<true>, dim3(nfix), dim3(CARRY_THREADS), 0, stream, C);
    } else if (C.node_kind == CARRY_LIN) {   // instances of their own: each loader's registers stay
        hipLaunchKernelGGL((k_carry_spec<false, CARRY_LIN>), dim3(nch), dim3(CARRY_THREADS), 0, stream, C);
        if (!one_pass)
            hipLaunchKernelGGL((k_carry_fix<false, CARRY_LIN>), dim3(nfix), dim3(CARRY_THREADS), 0, stream, C);
    } else if (C.node_kind == CARRY_PLAIN) {   // out of the other node kinds' launches
        hipLaunchKernelGGL((k_carry_spec<false, CARRY_PLAIN>), dim3(nch), dim3(CARRY_THREADS), 0, stream, C);
        if (!one_pass)
            hipLaunchKernelGGL((k_carry_fix<false, CARRY_PLAIN>), dim3(nfix), dim3(CARRY_THREADS), 0, stream, C);
    } else {
        hipLaunchKernelGGL(k_carry_spec<false>, dim3(nch), dim3(CARRY_THREADS), 0, stream, C);
        if (!one_pass)
            hipLaunchKernelGGL(k_carry_fix<false>, dim3(nfix), dim3(CARRY_THREADS), 0, stream, C);
    }
    return hipGetLastError();
}

}  // namespace pg
