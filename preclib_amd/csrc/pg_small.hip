// pg_small.hip -- schoolbook products below the transform threshold, as limb
// sums for the carry kernels.
//
// preclib multiplies operands of <= 192 limbs (and any product with a side of
// <= 24 limbs) by schoolbook or Karatsuba (mul_basic.cpp:131-185, 255-258,
// 271-275; mul_kara.cpp:31-49).  On the device they appear as the leaves of
// blocked multiplies and batches.  Each column sum c_k = sum_i x_i y_{k-i}
// (at most 256 terms of 128 bits) is kept in three words (w0, w1, w2); limb j
// of the product receives w0_j + w1_{j-1} + w2_{j-2}, whose overflow (<= 2)
// becomes the high word of the limb sum.
#include "pg_kernels.h"

namespace pg {

__device__ __forceinline__ void mac128(uint64_t &w0, uint64_t &w1, uint64_t &w2, uint64_t x, uint64_t y) {
    const uint64_t lo = x * y, hi = __umul64hi(x, y);
    w0 += lo;
    const uint64_t c0 = w0 < lo;
    const uint64_t h = hi + c0;            // hi <= 2^64 - 2, no overflow
    w1 += h;
    w2 += w1 < h;
}

__global__ void __launch_bounds__(SMALL_THREADS) k_small(const SmallLaunch L) {
    __shared__ uint64_t xs[SMALL_MAX_SHORT];
    __shared__ uint64_t w[3][SMALL_THREADS];
    const SmallLeaf &D = L.leaves ? L.leaves[L.blk_leaf[blockIdx.x]] : L.one;
    const bool swap = D.an > D.bn;
    const uint64_t *x = swap ? D.b : D.a, *y = swap ? D.a : D.b;
    const long nx = (long)(swap ? D.bn : D.an), ny = (long)(swap ? D.an : D.bn);
    for (int i = threadIdx.x; i < nx; i += SMALL_THREADS) xs[i] = x[i];
    __syncthreads();
    const long c0 = (long)(blockIdx.x - D.blk0) * SMALL_COLS;
    const long k = c0 - 2 + threadIdx.x;       // this thread's column
    uint64_t w0 = 0, w1 = 0, w2 = 0;
    if (k >= 0) {
        const long ilo = k - ny + 1 > 0 ? k - ny + 1 : 0;
        const long ihi = k < nx - 1 ? k : nx - 1;
        for (long i = ilo; i <= ihi; ++i) mac128(w0, w1, w2, xs[i], y[k - i]);
    }
    w[0][threadIdx.x] = w0;
    w[1][threadIdx.x] = w1;
    w[2][threadIdx.x] = w2;
    __syncthreads();
    if (threadIdx.x < 2) return;
    const size_t j = (size_t)k;
    if (j >= D.sum_len) return;
    // limb j = w0_j + w1_{j-1} + w2_{j-2}
    const uint64_t a0 = w[0][threadIdx.x], a1 = w[1][threadIdx.x - 1], a2 = w[2][threadIdx.x - 2];
    const uint64_t s1 = a0 + a1;
    const uint64_t s2 = s1 + a2;
    D.slo[j] = s2;
    D.shi[j] = (s1 < a0) + (s2 < s1);
}

hipError_t launch_small(const SmallLaunch &L, size_t nblocks, hipStream_t stream, const StageHook &h, double io_bytes) {
    h.begin(h.ctx, "small");
    hipLaunchKernelGGL(k_small, dim3(nblocks), dim3(SMALL_THREADS), 0, stream, L);
    h.end(h.ctx, "small", io_bytes);
    return hipGetLastError();
}

}  // namespace pg
