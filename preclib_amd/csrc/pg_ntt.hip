// pg_ntt.hip -- four-step NTT product kernels for gfx950.
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
#include "pg_modarith.h"
#include "pg_kernels.h"

namespace pg {

// ---- in-LDS radix passes --------------------------------------------------

template <int R>
__device__ __forceinline__ void pass_fwd(uint32_t *lds, int logN, int logT, int logB,
                                         const uint2 *__restrict__ tw, const PrimeK &c) {
    constexpr int LR = (R == 2) ? 1 : (R == 4) ? 2 : (R == 8) ? 3 : (R == 16) ? 4 : 5;
    const int logM = logB - LR;
    const int M = 1 << logM;
    const int ncols = 1 << (logN - LR + logT);
    const int T = 1 << logT;
    for (int col = threadIdx.x; col < ncols; col += blockDim.x) {
        const int cc = col & (T - 1);
        const int r = col >> logT;
        const int blk = r >> logM;
        const int j2 = r & (M - 1);
        const int rowbase = (blk << logB) + j2;
        uint32_t x[R];
#pragma unroll
        for (int q = 0; q < R; ++q) x[q] = lds[((rowbase + (q << logM)) << logT) + cc];
        dft_dif<R>(x, c);
        if (M > 1) {
#pragma unroll
            for (int q = 1; q < R; ++q) {
                const uint2 w = tw[(bitrev<R>(q) << logM) + j2];
                x[q] = mul_shoup(x[q], w.x, w.y, c.p);
            }
        }
#pragma unroll
        for (int q = 0; q < R; ++q) lds[((rowbase + (bitrev<R>(q) << logM)) << logT) + cc] = x[q];
    }
}

template <int R>
__device__ __forceinline__ void pass_inv(uint32_t *lds, int logN, int logT, int logB,
                                         const uint2 *__restrict__ itw, const PrimeK &c) {
    constexpr int LR = (R == 2) ? 1 : (R == 4) ? 2 : (R == 8) ? 3 : (R == 16) ? 4 : 5;
    const int logM = logB - LR;
    const int M = 1 << logM;
    const int ncols = 1 << (logN - LR + logT);
    const int T = 1 << logT;
    for (int col = threadIdx.x; col < ncols; col += blockDim.x) {
        const int cc = col & (T - 1);
        const int r = col >> logT;
        const int blk = r >> logM;
        const int j2 = r & (M - 1);
        const int rowbase = (blk << logB) + j2;
        uint32_t x[R];
#pragma unroll
        for (int q = 0; q < R; ++q) x[q] = lds[((rowbase + (bitrev<R>(q) << logM)) << logT) + cc];
        if (M > 1) {
#pragma unroll
            for (int q = 1; q < R; ++q) {
                const uint2 w = itw[(bitrev<R>(q) << logM) + j2];
                x[q] = mul_shoup(x[q], w.x, w.y, c.p);
            }
        }
        dft_dit_inv<R>(x, c);
#pragma unroll
        for (int q = 0; q < R; ++q) lds[((rowbase + (q << logM)) << logT) + cc] = x[q];
    }
}

__device__ __forceinline__ void run_pass(bool fwd, int lr, uint32_t *lds, int logN, int logT,
                                         int logB, const uint2 *tw, const PrimeK &c) {
    switch (lr) {
    case 1: fwd ? pass_fwd<2>(lds, logN, logT, logB, tw, c) : pass_inv<2>(lds, logN, logT, logB, tw, c); break;
    case 2: fwd ? pass_fwd<4>(lds, logN, logT, logB, tw, c) : pass_inv<4>(lds, logN, logT, logB, tw, c); break;
    case 3: fwd ? pass_fwd<8>(lds, logN, logT, logB, tw, c) : pass_inv<8>(lds, logN, logT, logB, tw, c); break;
    case 4: fwd ? pass_fwd<16>(lds, logN, logT, logB, tw, c) : pass_inv<16>(lds, logN, logT, logB, tw, c); break;
    default: fwd ? pass_fwd<32>(lds, logN, logT, logB, tw, c) : pass_inv<32>(lds, logN, logT, logB, tw, c); break;
    }
}

// Forward N-point transform of every column of the LDS tile.
__device__ void lds_forward(uint32_t *lds, int logT, const SubPlan &sp, const uint2 *const *tw,
                            const PrimeK &c) {
    int logB = sp.logN;
    for (int i = 0; i < sp.npass; ++i) {
        run_pass(true, sp.lr[i], lds, sp.logN, logT, logB, tw[i], c);
        __syncthreads();
        logB -= sp.lr[i];
    }
}

// Inverse (unnormalized) N-point transform, the forward network reversed.
__device__ void lds_inverse(uint32_t *lds, int logT, const SubPlan &sp, const uint2 *const *itw,
                            const PrimeK &c) {
    int logB = sp.lr[sp.npass - 1];
    for (int i = sp.npass - 1; i >= 0; --i) {
        run_pass(false, sp.lr[i], lds, sp.logN, logT, logB, itw[i], c);
        __syncthreads();
        if (i > 0) logB += sp.lr[i - 1];
    }
}

// Frequency index held at position rho after the forward passes of sp
// (mixed-radix digit reversal).
__device__ __forceinline__ uint32_t freq_of_pos(uint32_t rho, const SubPlan &sp) {
    uint32_t k = 0, w = 1;
    int logB = sp.logN;
    for (int i = 0; i < sp.npass; ++i) {
        logB -= sp.lr[i];
        uint32_t d = (rho >> logB) & ((1u << sp.lr[i]) - 1);
        k += d * w;
        w <<= sp.lr[i];
    }
    return k;
}

__device__ __forceinline__ uint32_t twiddle_mont(const uint32_t *__restrict__ lo,
                                                 const uint32_t *__restrict__ hi, uint32_t e,
                                                 int lobits, const PrimeK &c) {
    return mont_mul(lo[e & ((1u << lobits) - 1)], hi[e >> lobits], c.p, c.ninv);
}

// ---- kernel 1: forward column pass ------------------------------------------

__global__ void __launch_bounds__(NTT_THREADS)
k_col_fwd(const NttLaunch L) {
    extern __shared__ uint32_t lds[];
    const int tile = blockIdx.x, op = blockIdx.y, pr = blockIdx.z;
    const PrimeK &c = L.pk[pr];
    const int logT = L.logT, T = 1 << logT;
    const uint32_t c0 = (uint32_t)tile << logT;
    const uint64_t *src = op ? L.b : L.a;
    const size_t nl = op ? L.bn : L.an;
    const int N1 = 1 << L.col.logN;
    // digits of the tile: row j1 holds digits j1*N2 + c0 + [0, T)
    const int lpr = T >> 2;   // limbs per tile row
    for (int e = threadIdx.x; e < N1 * lpr; e += blockDim.x) {
        const int j1 = e / lpr, m = e - j1 * lpr;
        const size_t li = (((size_t)j1 << L.logN2) + c0 + 4 * m) >> 2;
        const uint64_t v = li < nl ? src[li] : 0;
        uint32_t *d = lds + j1 * T + 4 * m;
        d[0] = (uint32_t)(v & 0xFFFF);
        d[1] = (uint32_t)((v >> 16) & 0xFFFF);
        d[2] = (uint32_t)((v >> 32) & 0xFFFF);
        d[3] = (uint32_t)(v >> 48);
    }
    __syncthreads();
    lds_forward(lds, logT, L.col, L.col_tw[pr], c);
    uint32_t *out = L.spec + ((size_t)(op * L.nprimes + pr) << L.logn);
    const uint32_t nmask = (1u << L.logn) - 1;
    for (int e = threadIdx.x; e < N1 * T; e += blockDim.x) {
        const int rho = e >> logT, cc = e & (T - 1);
        const uint32_t k1 = freq_of_pos(rho, L.col);
        const uint32_t ex = ((c0 + cc) * k1) & nmask;
        uint32_t x = lds[e];
        if (ex) x = mont_mul(x, twiddle_mont(L.glo[pr], L.ghi[pr], ex, L.lobits, c), c.p, c.ninv);
        out[((size_t)rho << L.logN2) + c0 + cc] = x;
    }
}

// ---- kernel 2: row pass (forward A, forward B, pointwise, inverse) -------------

__device__ void load_row(uint32_t *lds, const NttLaunch &L, int op, int pr, int row) {
    const int N2 = 1 << L.logN2;
    if (L.col.logN == 0) {
        // single-row plan: the row is the digit vector itself
        const uint64_t *src = op ? L.b : L.a;
        const size_t nl = op ? L.bn : L.an;
        for (int m = threadIdx.x; m < (N2 >> 2); m += blockDim.x) {
            const uint64_t v = (size_t)m < nl ? src[m] : 0;
            uint32_t *d = lds + 4 * m;
            d[0] = (uint32_t)(v & 0xFFFF);
            d[1] = (uint32_t)((v >> 16) & 0xFFFF);
            d[2] = (uint32_t)((v >> 32) & 0xFFFF);
            d[3] = (uint32_t)(v >> 48);
        }
    } else {
        const uint4 *in = (const uint4 *)(L.spec + ((size_t)(op * L.nprimes + pr) << L.logn) +
                                          ((size_t)row << L.logN2));
        for (int m = threadIdx.x; m < (N2 >> 2); m += blockDim.x) ((uint4 *)lds)[m] = in[m];
    }
}

__global__ void __launch_bounds__(NTT_THREADS)
k_row(const NttLaunch L) {
    extern __shared__ uint32_t lds[];
    const int row = blockIdx.x, pr = blockIdx.y;
    const PrimeK &c = L.pk[pr];
    const int N2 = 1 << L.logN2;
    constexpr int EMAX = ROW_MAX_ELEMS / NTT_THREADS;
    uint32_t keep[EMAX];

    load_row(lds, L, 0, pr, row);
    __syncthreads();
    lds_forward(lds, 0, L.row, L.row_tw[pr], c);
#pragma unroll
    for (int k = 0; k < EMAX; ++k) {
        const int i = threadIdx.x + k * NTT_THREADS;
        if (i < N2) keep[k] = lds[i];
    }
    __syncthreads();
    load_row(lds, L, 1, pr, row);
    __syncthreads();
    lds_forward(lds, 0, L.row, L.row_tw[pr], c);
#pragma unroll
    for (int k = 0; k < EMAX; ++k) {
        const int i = threadIdx.x + k * NTT_THREADS;
        if (i < N2) lds[i] = mont_mul(lds[i], keep[k], c.p, c.ninv);
    }
    __syncthreads();
    lds_inverse(lds, 0, L.row, L.row_itw[pr], c);
    uint4 *out = (uint4 *)(L.prod + ((size_t)pr << L.logn) + ((size_t)row << L.logN2));
    for (int m = threadIdx.x; m < (N2 >> 2); m += blockDim.x) out[m] = ((const uint4 *)lds)[m];
}

// ---- kernel 3: inverse column pass, CRT, limb sums ------------------------------

__global__ void __launch_bounds__(NTT_THREADS)
k_col_inv(const NttLaunch L) {
    extern __shared__ uint32_t lds[];
    const int tile = blockIdx.x;
    const int logT = L.logT, T = 1 << logT;
    const uint32_t c0 = (uint32_t)tile << logT;
    const int N1 = 1 << L.col.logN;
    const int NE = N1 * T;
    constexpr int EMAX = COL_MAX_ELEMS / NTT_THREADS;
    uint32_t res0[EMAX];
    uint64_t coef[EMAX];
    const uint32_t nmask = (1u << L.logn) - 1;

    for (int pr = 0; pr < L.nprimes; ++pr) {
        const PrimeK &c = L.pk[pr];
        const uint32_t *in = L.prod + ((size_t)pr << L.logn);
        for (int e = threadIdx.x; e < NE; e += blockDim.x) {
            const int rho = e >> logT, cc = e & (T - 1);
            const uint32_t k1 = freq_of_pos(rho, L.col);
            const uint32_t ex = ((c0 + cc) * k1) & nmask;   // ilo/ihi hold omega_n^-e
            const uint32_t x = in[((size_t)rho << L.logN2) + c0 + cc];
            lds[e] = mont_mul(x, twiddle_mont(L.ilo[pr], L.ihi[pr], ex, L.lobits, c), c.p, c.ninv);
        }
        __syncthreads();
        if (L.col.npass) lds_inverse(lds, logT, L.col, L.col_itw[pr], c);
#pragma unroll
        for (int k = 0; k < EMAX; ++k) {
            const int i = threadIdx.x + k * NTT_THREADS;
            if (i < NE) {
                const uint32_t r = lds[i];
                if (pr == 0) {
                    res0[k] = r;
                } else {
                    // Garner (mul_ntt.cpp:405-412): x = r1 + p1 * ((r2 - r1) * p1^-1 mod p2)
                    const uint32_t d = mod_sub(r, res0[k], c.p);   // r1 < p1 < p2
                    const uint32_t t = mul_shoup(d, L.crt_inv, L.crt_invq, c.p);
                    coef[k] = (uint64_t)res0[k] + (uint64_t)L.pk[0].p * t;
                }
            }
        }
        __syncthreads();
    }
    // Limb sums: limb (rho, m) collects columns 4m..4m+3 of row rho.
    const int lpr = T >> 2;
    const int NL = N1 * lpr;
    constexpr int LMAX = (COL_MAX_ELEMS / 4 + NTT_THREADS - 1) / NTT_THREADS;
    uint64_t slo[LMAX], shi[LMAX];
#pragma unroll
    for (int k = 0; k < LMAX; ++k) { slo[k] = 0; shi[k] = 0; }
    for (int half = 0; half < 2; ++half) {
#pragma unroll
        for (int k = 0; k < EMAX; ++k) {
            const int i = threadIdx.x + k * NTT_THREADS;
            if (i < NE) lds[i] = half ? (uint32_t)(coef[k] >> 32) : (uint32_t)coef[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < LMAX; ++k) {
            const int e = threadIdx.x + k * NTT_THREADS;
            if (e < NL) {
                const int rho = e / lpr, m = e - rho * lpr;
                const uint32_t *w = lds + rho * T + 4 * m;
                // sum_t w[t] * 2^(16t + 32*half), accumulated in 128 bits
                uint64_t lo = slo[k], hi = shi[k];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int sh = 16 * t + 32 * half;   // 0..112
                    const uint64_t v = w[t];
                    uint64_t plo, phi;
                    if (sh == 0) { plo = v; phi = 0; }
                    else if (sh < 64) { plo = v << sh; phi = v >> (64 - sh); }
                    else { plo = 0; phi = v << (sh - 64); }
                    const uint64_t s = lo + plo;
                    hi += phi + (s < lo);
                    lo = s;
                }
                slo[k] = lo; shi[k] = hi;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < LMAX; ++k) {
        const int e = threadIdx.x + k * NTT_THREADS;
        if (e < NL) {
            const int rho = e / lpr, m = e - rho * lpr;
            const size_t li = (((size_t)rho << L.logN2) + c0 + 4 * m) >> 2;
            L.slo[li] = slo[k];
            L.shi[li] = shi[k];
        }
    }
}

}  // namespace pg
