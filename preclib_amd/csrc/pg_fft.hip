// pg_fft.hip -- FP64 FFT product kernels for gfx950 (preclib's mul_fft path).
//
// Follows mul_fft (src/mul_fft.cpp:340-394): digits of `bits` bits (16 up to
// 1024 limbs, 8 above, 390-393), a's digits in the real part and b's in the
// imaginary part of one complex sequence (234-246, 350-351), one forward
// transform with omega = exp(+2 pi i / len) (94), the pair convolution
// A_k B_k with A_k = (X_k + conj X_{n-k}) / 2, B_k = (X_k - conj X_{n-k}) / 2i
// (248-305), an inverse transform with exp(-2 pi i / len) and 1/n (96, 219),
// rounding half away from zero (307-309), the largest rounding distance
// |v - nearbyint(v)| (359-382) and a carry in base 2^bits (311-338).
//
// Layout is the NTT's four-step (pg_ntt.hip) with complex doubles:
//   k_fft_col_fwd : digits -> columns of N1 points, twiddle omega_n^(c k1)
//   k_fft_row     : a pair of rows {k1, N1 - k1}: forward row DFTs, pair
//                   convolution across the two rows (out_{n-k} = conj out_k),
//                   inverse row DFTs.  One row lives in LDS, the other in
//                   registers.
//   k_fft_col_inv : inverse twiddle, inverse column DFTs, 1/n, rounding,
//                   rounding-distance max, limb sums (or half-limb parts)
//   carry         : pg_carry.hip
// Twiddles are accurate to the last bit (host long double), so the rounding
// distance is the transform's own error, not the reference's value.
#include "pg_kernels.h"

namespace pg {

__device__ __forceinline__ int fp(int i) { return i + (i >> 5); }   // padded LDS index
constexpr size_t fft_lds_words(size_t n) { return n + (n >> 5); }   // LDS slots for fp(0 .. n-1)

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 w) {
    return make_double2(fma(a.x, w.x, -a.y * w.y), fma(a.x, w.y, a.y * w.x));
}

// R-point DFT in registers (DIF, natural in, x[q] = X[bitrev(q)] out),
// omega_R^k = w16[k * 16 / R].
template <int R>
__device__ __forceinline__ void cdft_dif(double2 (&x)[R], const double2 *w16) {
#pragma unroll
    for (int h = R / 2; h >= 1; h >>= 1) {
#pragma unroll
        for (int b = 0; b < R; b += 2 * h) {
#pragma unroll
            for (int j = 0; j < h; ++j) {
                const double2 u = x[b + j], v = x[b + j + h];
                x[b + j] = cadd(u, v);
                const double2 d = csub(u, v);
                x[b + j + h] = j == 0 ? d : cmul(d, w16[j * (8 / h)]);
            }
        }
    }
}

// R-point inverse DFT (DIT, x[q] = Y[bitrev(q)] in, natural out, unscaled).
template <int R>
__device__ __forceinline__ void cdft_dit_inv(double2 (&x)[R], const double2 *iw16) {
#pragma unroll
    for (int h = 1; h <= R / 2; h <<= 1) {
#pragma unroll
        for (int b = 0; b < R; b += 2 * h) {
#pragma unroll
            for (int j = 0; j < h; ++j) {
                const double2 u = x[b + j];
                const double2 v = j == 0 ? x[b + j + h] : cmul(x[b + j + h], iw16[j * (8 / h)]);
                x[b + j] = cadd(u, v);
                x[b + j + h] = csub(u, v);
            }
        }
    }
}

template <int R>
__device__ __forceinline__ constexpr int cbitrev(int q) {
    int r = 0;
    for (int b = 1, c = R >> 1; b < R; b <<= 1, c >>= 1)
        if (q & b) r |= c;
    return r;
}

template <int LR, int LOGN, int LOGT, int LOGB, int NT>
__device__ __forceinline__ void cpass_fwd(double2 *lds, const double2 *__restrict__ tw, const FftK &k) {
    constexpr int R = 1 << LR, LOGM = LOGB - LR, M = 1 << LOGM, T = 1 << LOGT;
    constexpr int NCOLS = 1 << (LOGN - LR + LOGT), S = M * T;
#pragma unroll 1
    for (int col = threadIdx.x; col < NCOLS; col += NT) {
        const int cc = col & (T - 1), r = col >> LOGT;
        const int blk = r >> LOGM, j2 = r & (M - 1);
        const int base = (((blk << LOGB) + j2) << LOGT) + cc;
        double2 x[R];
#pragma unroll
        for (int q = 0; q < R; ++q) x[q] = lds[fp(base + q * S)];
        cdft_dif<R>(x, k.w16);
        if constexpr (M > 1) {
#pragma unroll
            for (int q = 1; q < R; ++q) x[q] = cmul(x[q], tw[cbitrev<R>(q) * M + j2]);
        }
#pragma unroll
        for (int q = 0; q < R; ++q) lds[fp(base + cbitrev<R>(q) * S)] = x[q];
    }
}

template <int LR, int LOGN, int LOGT, int LOGB, int NT>
__device__ __forceinline__ void cpass_inv(double2 *lds, const double2 *__restrict__ itw, const FftK &k) {
    constexpr int R = 1 << LR, LOGM = LOGB - LR, M = 1 << LOGM, T = 1 << LOGT;
    constexpr int NCOLS = 1 << (LOGN - LR + LOGT), S = M * T;
#pragma unroll 1
    for (int col = threadIdx.x; col < NCOLS; col += NT) {
        const int cc = col & (T - 1), r = col >> LOGT;
        const int blk = r >> LOGM, j2 = r & (M - 1);
        const int base = (((blk << LOGB) + j2) << LOGT) + cc;
        double2 x[R];
#pragma unroll
        for (int q = 0; q < R; ++q) x[q] = lds[fp(base + cbitrev<R>(q) * S)];
        if constexpr (M > 1) {
#pragma unroll
            for (int q = 1; q < R; ++q) x[q] = cmul(x[q], itw[cbitrev<R>(q) * M + j2]);
        }
        cdft_dit_inv<R>(x, k.iw16);
#pragma unroll
        for (int q = 0; q < R; ++q) lds[fp(base + q * S)] = x[q];
    }
}

template <int LOGN, int LOGT, int NT, int LOGB, int I>
__device__ __forceinline__ void cfwd_from(double2 *lds, const double2 *const *tw, const FftK &k) {
    constexpr int LR = fsplit_lr(LOGN, I);
    if constexpr (LR > 0) {
        cpass_fwd<LR, LOGN, LOGT, LOGB, NT>(lds, tw[I], k);
        __syncthreads();
        cfwd_from<LOGN, LOGT, NT, LOGB - LR, I + 1>(lds, tw, k);
    }
}

template <int LOGN, int LOGT, int NT, int LOGB, int I>
__device__ __forceinline__ void cinv_from(double2 *lds, const double2 *const *itw, const FftK &k) {
    constexpr int LR = fsplit_lr(LOGN, I);
    if constexpr (LR > 0) {
        cinv_from<LOGN, LOGT, NT, LOGB - LR, I + 1>(lds, itw, k);
        cpass_inv<LR, LOGN, LOGT, LOGB, NT>(lds, itw[I], k);
        __syncthreads();
    }
}

// Frequency at position rho after the forward passes, and its inverse.
template <int LOGN>
__device__ __forceinline__ uint32_t cfreq(uint32_t rho) {
    uint32_t f = 0;
    int w = 0, logB = LOGN;
#pragma unroll
    for (int i = 0; i < fsplit_np(LOGN); ++i) {
        const int lr = fsplit_lr(LOGN, i);
        logB -= lr;
        f |= ((rho >> logB) & ((1u << lr) - 1)) << w;
        w += lr;
    }
    return f;
}
template <int LOGN>
__device__ __forceinline__ uint32_t cpos(uint32_t f) {
    uint32_t rho = 0;
    int w = 0, logB = LOGN;
#pragma unroll
    for (int i = 0; i < fsplit_np(LOGN); ++i) {
        const int lr = fsplit_lr(LOGN, i);
        logB -= lr;
        rho |= ((f >> w) & ((1u << lr) - 1)) << logB;
        w += lr;
    }
    return rho;
}

__device__ __forceinline__ double2 gtw(const FftLaunch &L, uint32_t e) {
    return cmul(L.glo[e & ((1u << L.lobits) - 1)], L.ghi[e >> L.lobits]);
}

__device__ __forceinline__ int fxcd_tile(int i, int ntiles) {
    return (ntiles & 7) ? i : (i & 7) * (ntiles >> 3) + (i >> 3);
}

// Digit j (of `bits` bits) of a limb array, 0 beyond the array.
__device__ __forceinline__ double digit_of(const uint64_t *x, size_t xn, size_t j, int bits) {
    const int dpl = 64 / bits;
    const size_t li = j / dpl;
    if (li >= xn) return 0.0;
    return (double)((x[li] >> ((j % dpl) * bits)) & ((1ull << bits) - 1));
}

// ---- kernel 1: digits and forward column pass -----------------------------

template <int LOGN>
__global__ void __launch_bounds__(FFT_THREADS) k_fft_col_fwd(const FftLaunch L) {
    constexpr int L1 = fft_l1(LOGN), L2 = fft_l2(LOGN), LT = fft_lt(LOGN);
    constexpr int N1 = 1 << L1, T = 1 << LT, NE = N1 * T;
    constexpr uint32_t NMASK = (1u << LOGN) - 1;
    extern __shared__ double2 clds[];
    const FftLeaf &D = L.leaves ? L.leaves[blockIdx.z] : L.one;
    const int tile = fxcd_tile(blockIdx.x, gridDim.x);
    const uint32_t c0 = (uint32_t)tile << LT;
    for (int e = threadIdx.x; e < NE; e += FFT_THREADS) {
        const int j1 = e >> LT, cc = e & (T - 1);
        const size_t j = ((size_t)j1 << L2) + c0 + cc;
        clds[fp(e)] = make_double2(digit_of(D.a, D.an, j, L.bits), digit_of(D.b, D.bn, j, L.bits));
    }
    __syncthreads();
    cfwd_from<L1, LT, FFT_THREADS, L1, 0>(clds, L.col_tw, L.k);
    for (int e = threadIdx.x; e < NE; e += FFT_THREADS) {
        const int rho = e >> LT, cc = e & (T - 1);
        const uint32_t ex = ((c0 + cc) * cfreq<L1>(rho)) & NMASK;
        double2 x = clds[fp(e)];
        if (ex) x = cmul(x, gtw(L, ex));
        D.spec[((size_t)rho << L2) + c0 + cc] = x;
    }
}

// ---- kernel 2: row pairs, pair convolution ---------------------------------------

// out_k from X_k = x and X_{n-k} = y (mul_fft.cpp:290-304).
__device__ __forceinline__ double2 pair_mul(double2 x, double2 y) {
    const double ar = (x.x + y.x) * 0.5, ai = (x.y - y.y) * 0.5;
    const double br = (x.y + y.y) * 0.5, bi = (y.x - x.x) * 0.5;
    return make_double2(ar * br - ai * bi, ar * bi + ai * br);
}

template <int LOGN>
__device__ __forceinline__ void load_crow(double2 *lds, const FftLaunch &L, const FftLeaf &D, int rho) {
    constexpr int L1 = fft_l1(LOGN), L2 = fft_l2(LOGN), N2 = 1 << L2;
    if constexpr (L1 == 0) {
        for (int j = threadIdx.x; j < N2; j += FFT_THREADS)
            lds[fp(j)] = make_double2(digit_of(D.a, D.an, j, L.bits), digit_of(D.b, D.bn, j, L.bits));
    } else {
        const double2 *in = D.spec + ((size_t)rho << L2);
        for (int j = threadIdx.x; j < N2; j += FFT_THREADS) lds[fp(j)] = in[j];
    }
}

template <int LOGN>
__device__ __forceinline__ void store_crow(const double2 *lds, const FftLeaf &D, int rho) {
    constexpr int L2 = fft_l2(LOGN), N2 = 1 << L2;
    double2 *out = D.spec + ((size_t)rho << L2);
    for (int j = threadIdx.x; j < N2; j += FFT_THREADS) out[j] = lds[fp(j)];
}

template <int LOGN>
__global__ void __launch_bounds__(FFT_THREADS) k_fft_row(const FftLaunch L) {
    constexpr int L1 = fft_l1(LOGN), L2 = fft_l2(LOGN), N1 = 1 << L1, N2 = 1 << L2;
    constexpr int E = (N2 + FFT_THREADS - 1) / FFT_THREADS;
    extern __shared__ double2 clds[];
    const FftLeaf &D = L.leaves ? L.leaves[blockIdx.z] : L.one;
    if (blockIdx.x == 0 && threadIdx.x == 0) *D.maxerr = 0;   // read only by k_fft_col_inv
    const uint32_t k1 = blockIdx.x;                 // 0 .. N1/2
    const uint32_t k1p = (N1 - k1) & (N1 - 1);      // partner row frequency
    const int rho = L1 ? (int)cpos<L1>(k1) : 0;
    const int rhop = L1 ? (int)cpos<L1>(k1p) : 0;
    // partner of row frequency k2: n - (k1 + N1 k2) = k1p + N1 * k2p
    auto partner = [&](uint32_t k2) -> uint32_t {
        return k1 == 0 ? ((N2 - k2) & (N2 - 1)) : (N2 - 1 - k2);
    };
    if (k1 == k1p) {
        // self-paired row (k1 = 0 or N1/2): the whole pair convolution is in LDS
        load_crow<LOGN>(clds, L, D, rho);
        __syncthreads();
        cfwd_from<L2, 0, FFT_THREADS, L2, 0>(clds, L.row_tw, L.k);
        for (int i = threadIdx.x; i < N2; i += FFT_THREADS) {
            const uint32_t k2 = cfreq<L2>(i), k2p = partner(k2);
            if (k2 > k2p) continue;   // each unordered pair once
            const int ip = cpos<L2>(k2p);
            const double2 o = pair_mul(clds[fp(i)], clds[fp(ip)]);
            clds[fp(i)] = o;
            clds[fp(ip)] = make_double2(o.x, -o.y);
        }
        __syncthreads();
        cinv_from<L2, 0, FFT_THREADS, L2, 0>(clds, L.row_itw, L.k);
        store_crow<LOGN>(clds, D, rho);
        return;
    }
    // row k1 -> registers, row k1p in LDS
    double2 keep[E];
    load_crow<LOGN>(clds, L, D, rho);
    __syncthreads();
    cfwd_from<L2, 0, FFT_THREADS, L2, 0>(clds, L.row_tw, L.k);
#pragma unroll
    for (int q = 0; q < E; ++q) {
        const int i = threadIdx.x + q * FFT_THREADS;
        if (i < N2) keep[q] = clds[fp(i)];
    }
    __syncthreads();
    load_crow<LOGN>(clds, L, D, rhop);
    __syncthreads();
    cfwd_from<L2, 0, FFT_THREADS, L2, 0>(clds, L.row_tw, L.k);
#pragma unroll
    for (int q = 0; q < E; ++q) {
        const int i = threadIdx.x + q * FFT_THREADS;
        if (i < N2) {
            const int ip = cpos<L2>(partner(cfreq<L2>(i)));
            const double2 o = pair_mul(keep[q], clds[fp(ip)]);
            keep[q] = o;
            clds[fp(ip)] = make_double2(o.x, -o.y);   // out_{n-k} = conj(out_k)
        }
    }
    __syncthreads();
    cinv_from<L2, 0, FFT_THREADS, L2, 0>(clds, L.row_itw, L.k);
    store_crow<LOGN>(clds, D, rhop);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < E; ++q) {
        const int i = threadIdx.x + q * FFT_THREADS;
        if (i < N2) clds[fp(i)] = keep[q];
    }
    __syncthreads();
    cinv_from<L2, 0, FFT_THREADS, L2, 0>(clds, L.row_itw, L.k);
    store_crow<LOGN>(clds, D, rho);
}

// ---- kernel 3: inverse column pass, rounding, limb sums --------------------------

// (lo, hi) += v * 2^sh for a signed 64-bit v, 0 <= sh < 64 (sign-extended 128-bit add).
__device__ __forceinline__ void sacc128(uint64_t &lo, uint64_t &hi, int64_t v, int sh) {
    const uint64_t ext = v < 0 ? ~0ull : 0ull;
    const uint64_t plo = sh ? ((uint64_t)v << sh) : (uint64_t)v;
    const uint64_t phi = sh ? (((uint64_t)v >> (64 - sh)) | (ext << sh)) : ext;
    const uint64_t s = lo + plo;
    hi += phi + (s < lo);
    lo = s;
}

template <int LOGN>
__global__ void __launch_bounds__(FFT_THREADS) k_fft_col_inv(const FftLaunch L) {
    constexpr int L1 = fft_l1(LOGN), L2 = fft_l2(LOGN), LT = fft_lt(LOGN);
    constexpr int N1 = 1 << L1, T = 1 << LT, NE = N1 * T;
    constexpr uint32_t NMASK = (1u << LOGN) - 1;
    extern __shared__ double2 clds[];
    const FftLeaf &D = L.leaves ? L.leaves[blockIdx.z] : L.one;
    const int tile = fxcd_tile(blockIdx.x, gridDim.x);
    const uint32_t c0 = (uint32_t)tile << LT;
    for (int e = threadIdx.x; e < NE; e += FFT_THREADS) {
        const int rho = e >> LT, cc = e & (T - 1);
        const uint32_t ex = ((c0 + cc) * cfreq<L1>(rho)) & NMASK;
        double2 x = D.spec[((size_t)rho << L2) + c0 + cc];
        if (ex) {
            const double2 w = gtw(L, ex);
            x = cmul(x, make_double2(w.x, -w.y));   // omega_n^-e
        }
        clds[fp(e)] = x;
    }
    __syncthreads();
    cinv_from<L1, LT, FFT_THREADS, L1, 0>(clds, L.col_itw, L.k);
    // round the real parts (mul_fft.cpp:364-376) into int64 coefficients, in place
    const double scale = 1.0 / (double)(1u << LOGN);
    double emax = 0.0;
    int64_t *coef = (int64_t *)clds;   // reuse: coefficient e at coef[2 * fp(e)]
    for (int e = threadIdx.x; e < NE; e += FFT_THREADS) {
        const double v = clds[fp(e)].x * scale;
        const double err = fabs(v - rint(v));
        emax = fmax(emax, err);
        coef[2 * fp(e)] = (int64_t)(v >= 0.0 ? v + 0.5 : v - 0.5);
    }
    // wave max, workgroup max, then one atomic per workgroup (non-negative
    // doubles order as their bit patterns)
    __shared__ double wmax[FFT_THREADS / 64];
    for (int off = 32; off > 0; off >>= 1) emax = fmax(emax, __shfl_xor(emax, off));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = emax;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < FFT_THREADS / 64; ++w) emax = fmax(emax, wmax[w]);
        atomicMax((unsigned long long *)D.maxerr, (unsigned long long)__double_as_longlong(emax));
    }
    // limb sums: DPL digits per limb; a tile row holds T digits, so one sum
    // covers min(T, DPL) digits.  PARTS = 2 when a limb spans two tiles.
    const int bits = L.bits, dpl = 64 / bits;
    const int seg = T < dpl ? T : dpl;            // digits per sum
    const int nsum = NE / seg;
    const int parts = dpl / seg;                  // 1 or 2
    for (int s = threadIdx.x; s < nsum; s += FFT_THREADS) {
        const int e0 = s * seg;
        const int rho = e0 >> LT, cc = e0 & (T - 1);
        uint64_t lo = 0, hi = 0;
        for (int t = 0; t < seg; ++t) sacc128(lo, hi, coef[2 * fp(e0 + t)], bits * t);
        const size_t d0 = ((size_t)rho << L2) + c0 + cc;   // first digit of the sum
        const size_t idx = parts == 1 ? d0 / dpl : d0 / seg;
        if (idx < D.sum_len) {
            D.slo[idx] = lo;
            D.shi[idx] = hi;
        }
    }
}

// ---- host-side launch table ----------------------------------------------------

template <int LOGN>
static hipError_t launch_fft_size(const FftLaunch &L, hipStream_t stream, const StageHook &h) {
    constexpr int L1 = fft_l1(LOGN), L2 = fft_l2(LOGN), LT = fft_lt(LOGN);
    const int ntiles = (1 << L2) >> LT;
    const size_t col_lds = fft_lds_words((size_t)1 << (L1 + LT)) * 16, row_lds = fft_lds_words((size_t)1 << L2) * 16;
    const int nl = L.leaves ? L.nleaves : 1;
    const double n = (double)(1u << LOGN) * nl, io = L.io_bytes;
    if constexpr (L1 > 0) {
        h.begin(h.ctx, "fft_col_fwd");
        hipLaunchKernelGGL(k_fft_col_fwd<LOGN>, dim3(ntiles, 1, nl), dim3(FFT_THREADS), col_lds, stream, L);
        h.end(h.ctx, "fft_col_fwd", io + 32.0 * n);
    }
    h.begin(h.ctx, "fft_row");
    hipLaunchKernelGGL(k_fft_row<LOGN>, dim3((1 << L1) / 2 + 1, 1, nl), dim3(FFT_THREADS), row_lds, stream, L);
    h.end(h.ctx, "fft_row", 64.0 * n);
    h.begin(h.ctx, "fft_col_inv");
    hipLaunchKernelGGL(k_fft_col_inv<LOGN>, dim3(ntiles, 1, nl), dim3(FFT_THREADS), col_lds, stream, L);
    h.end(h.ctx, "fft_col_inv", 32.0 * n);
    return hipGetLastError();
}

template <int LOGN>
static hipError_t fft_limits() {
    constexpr int lim = 156 * 1024;   // dynamic LDS limit; leaves room for static LDS
    hipError_t e = hipSuccess;
    if (fft_l1(LOGN) > 0)
        e = hipFuncSetAttribute((const void *)k_fft_col_fwd<LOGN>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)k_fft_row<LOGN>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)k_fft_col_inv<LOGN>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    return e;
}

#define PG_FFT_SIZES(X) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18) \
    X(19) X(20) X(21) X(22) X(23) X(24)

hipError_t launch_fft(const FftLaunch &L, hipStream_t stream, const StageHook &hook) {
    switch (L.logn) {
#define PG_CASE(N) case N: return launch_fft_size<N>(L, stream, hook);
        PG_FFT_SIZES(PG_CASE)
#undef PG_CASE
    default: return hipErrorInvalidValue;
    }
}

hipError_t fft_set_lds_limits() {
    hipError_t e = hipSuccess;
#define PG_LIM(N) if (e == hipSuccess) e = fft_limits<N>();
    PG_FFT_SIZES(PG_LIM)
#undef PG_LIM
    return e;
}

}  // namespace pg
