// pg_modarith.h -- 32-bit modular arithmetic for the NTT primes on gfx950.
//
// The primes are preclib's NTT set (src/mul_ntt.cpp:29-31), all below 2^31,
// so a sum of two residues fits in 32 bits and every reduction below is a
// single unsigned min().  Two multiplication forms are used:
//  * Shoup's precomputed-quotient product x*w mod p for fixed twiddles
//    (wq = floor(w * 2^32 / p)): 1 mul_hi, 1 mul_lo and 1 mad_u64_u32,
//    result in [0, 2p), then one min() to [0, p).
//  * Montgomery's product a*b*2^-32 mod p (the reference's mont_reduce,
//    mul_ntt.cpp:43-48) for the pointwise spectrum product and for twiddles
//    formed on the fly, which have no precomputed quotient.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pg {

struct PrimeK {
    uint32_t p;      // modulus
    uint32_t ninv;   // -p^-1 mod 2^32
    uint32_t np;     // 2^32 - p
    uint32_t w32[16], w32q[16];     // omega_32^k, Shoup quotients (forward)
    uint32_t iw32[16], iw32q[16];   // omega_32^-k, Shoup quotients (inverse)
};

__device__ __forceinline__ uint32_t mod_add(uint32_t a, uint32_t b, uint32_t p) {
    uint32_t s = a + b;
    return min(s, s - p);
}

__device__ __forceinline__ uint32_t mod_sub(uint32_t a, uint32_t b, uint32_t p) {
    uint32_t d = a - b;
    return min(d, d + p);
}

// Lazy residues (LZ): for a prime below 2^30, values are kept in [0, 2p)
// instead of [0, p) through a whole transform (D. Harvey, "Faster arithmetic
// for number-theoretic transforms", J. Symbolic Comput. 60 (2014)).  Sums and
// differences are reduced against 2p, and products by a fixed twiddle skip
// Shoup's last correction.  Every operand stays below 4p < 2^32.
constexpr uint32_t LAZY_PRIME_LIMIT = 1u << 30;
#ifndef PG_DIF_RAW
#define PG_DIF_RAW 1
#endif

// x * w mod c.p with wq = floor(w * 2^32 / p); x < 2^32, w < p.  With
// q = floor(x wq / 2^32), x w - q p lies in [0, 2p); one min() brings it to
// [0, p) (LZ: left in [0, 2p)).  x w - q p is formed as x w + q (2^32 - p)
// with the negated modulus read from the prime's constants: the compiler
// then emits one v_mad_u64_u32 (full rate on gfx950) for the multiply-add
// instead of a mul_lo and a subtract.
template <bool LZ = false>
__device__ __forceinline__ uint32_t mul_shoup(uint32_t x, uint32_t w, uint32_t wq, const PrimeK &c) {
    const uint32_t r = x * w + __umulhi(x, wq) * c.np;   // [0, 2p)
    if constexpr (LZ) return r;
    else return min(r, r - c.p);
}

// a * b * 2^-32 mod p for a * b < p * 2^32.
__device__ __forceinline__ uint32_t mont_mul(uint32_t a, uint32_t b, uint32_t p, uint32_t ninv) {
    uint64_t t = (uint64_t)a * b;
    uint32_t m = (uint32_t)t * ninv;
    uint64_t u = (t + (uint64_t)m * p) >> 32;
    uint32_t r = (uint32_t)u;
    return min(r, r - p);
}

template <int R>
__device__ __forceinline__ constexpr int bitrev(int q) {
    int r = 0;
    for (int b = 1, c = R >> 1; b < R; b <<= 1, c >>= 1)
        if (q & b) r |= c;
    return r;
}

// R-point forward DFT in registers, Gentleman-Sande DIF: natural order in,
// x[q] = X[bitrev(q)] out.  Twiddles omega_R^k = omega_32^(k * 32 / R).
// HALFZ: x[R/2 .. R) are zero on entry, so the first stage is only the
// twiddle products of the lower half.  LZ: residues in [0, 2p) in and out.
// RAW: outputs x[1 .. R) feed Shoup products (which take any x < 2^32), so
// the last stage leaves them unreduced, below 2 * (reduction bound) <= 2^32.
template <int R, bool HALFZ = false, bool LZ = false, bool RAW = false>
__device__ __forceinline__ void dft_dif(uint32_t (&x)[R], const PrimeK &c) {
    const uint32_t p = c.p, pr = LZ ? 2 * c.p : c.p;   // reduction bound
    if constexpr (HALFZ) {
#pragma unroll
        for (int j = 0; j < R / 2; ++j)
            x[j + R / 2] = j == 0 ? x[0] : mul_shoup<LZ>(x[j], c.w32[j * (32 / R)], c.w32q[j * (32 / R)], c);
    }
#pragma unroll
    for (int h = HALFZ ? R / 4 : R / 2; h >= 1; h >>= 1) {
#pragma unroll
        for (int b = 0; b < R; b += 2 * h) {
#pragma unroll
            for (int j = 0; j < h; ++j) {
                uint32_t u = x[b + j], v = x[b + j + h];
                if (PG_DIF_RAW && RAW && h == 1) {
                    x[b] = b == 0 ? mod_add(u, v, pr) : u + v;
                    x[b + 1] = u - v + pr;
                    continue;
                }
                x[b + j] = mod_add(u, v, pr);
                if (j == 0) {
                    x[b + j + h] = mod_sub(u, v, pr);
                } else {
                    // u - v + pr lies in [1, 2 pr) < 2^32: Shoup's product
                    // takes it unreduced
                    const int k = j * (16 / h);   // omega_{2h}^j = omega_32^(j*16/h)
                    x[b + j + h] = mul_shoup<LZ>(u - v + pr, c.w32[k], c.w32q[k], c);
                }
            }
        }
    }
}

// R-point inverse DFT (unnormalized), Cooley-Tukey DIT: x[q] = Y[bitrev(q)]
// in, natural order out.  LZ: residues in [0, 2p) in and out.
template <int R, bool LZ = false>
__device__ __forceinline__ void dft_dit_inv(uint32_t (&x)[R], const PrimeK &c) {
    const uint32_t p = c.p, pr = LZ ? 2 * c.p : c.p;
#pragma unroll
    for (int h = 1; h <= R / 2; h <<= 1) {
#pragma unroll
        for (int b = 0; b < R; b += 2 * h) {
#pragma unroll
            for (int j = 0; j < h; ++j) {
                uint32_t u = x[b + j], v = x[b + j + h];
                if (j != 0) {
                    const int k = j * (16 / h);
                    v = mul_shoup<LZ>(v, c.iw32[k], c.iw32q[k], c);
                }
                x[b + j] = mod_add(u, v, pr);
                x[b + j + h] = mod_sub(u, v, pr);
            }
        }
    }
}

}  // namespace pg
