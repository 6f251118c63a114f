// pg_ntt.hip -- NTT launch table: the default plan of every transform length
// in both digit modes, and the dispatch to the autotuner's plan variants
// (pg_ntt_v1.hip .. pg_ntt_v15.hip).  Kernels: pg_ntt_impl.h.
#include "pg_ntt_impl.h"

namespace pg {

#define PG_NTT_SIZES(X) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18) X(19) \
    X(20) X(21) X(22) X(23) X(24) X(25) X(26)

hipError_t launch_ntt(const NttLaunch &L, hipStream_t stream, const StageHook &hook) {
    if (L.mode == MODE_WIDE && L.logn >= NTT_VARIANT_MIN_LOG && plan_distinct(L.logn, L.variant)) {
        switch (L.variant) {
        case 1: return launch_ntt_wide_v1(L, stream, hook);
        case 2: return launch_ntt_wide_v2(L, stream, hook);
        case 3: return launch_ntt_wide_v3(L, stream, hook);
        case 4: return launch_ntt_wide_v4(L, stream, hook);
        case 5: return launch_ntt_wide_v5(L, stream, hook);
        case 6: return launch_ntt_wide_v6(L, stream, hook);
        case 7: return launch_ntt_wide_v7(L, stream, hook);
        case 8: return launch_ntt_wide_v8(L, stream, hook);
        case 9: return launch_ntt_wide_v9(L, stream, hook);
        case 10: return launch_ntt_wide_v10(L, stream, hook);
        case 11: return launch_ntt_wide_v11(L, stream, hook);
        case 12: return launch_ntt_wide_v12(L, stream, hook);
        case 13: return launch_ntt_wide_v13(L, stream, hook);
        case 14: return launch_ntt_wide_v14(L, stream, hook);
        case 15: return launch_ntt_wide_v15(L, stream, hook);
        default: break;
        }
    }
    if (L.mode == MODE_WIDE) {
        switch (L.logn) {
#define PG_CASE(N) case N: return launch_size<N, MODE_WIDE, 0>(L, stream, hook);
            PG_NTT_SIZES(PG_CASE)
#undef PG_CASE
        default: return hipErrorInvalidValue;
        }
    }
    switch (L.logn) {
#define PG_CASE(N) case N: return launch_size<N, MODE_REF, 0>(L, stream, hook);
        PG_NTT_SIZES(PG_CASE)
#undef PG_CASE
    default: return hipErrorInvalidValue;
    }
}

hipError_t ntt_set_lds_limits() {
    hipError_t e = hipSuccess;
#define PG_LIM(N) if (e == hipSuccess) e = set_limits<N, MODE_REF, 0>(); if (e == hipSuccess) e = set_limits<N, MODE_WIDE, 0>();
    PG_NTT_SIZES(PG_LIM)
#undef PG_LIM
    if (e == hipSuccess) e = ntt_limits_wide_v1();
    if (e == hipSuccess) e = ntt_limits_wide_v2();
    if (e == hipSuccess) e = ntt_limits_wide_v3();
    if (e == hipSuccess) e = ntt_limits_wide_v4();
    if (e == hipSuccess) e = ntt_limits_wide_v5();
    if (e == hipSuccess) e = ntt_limits_wide_v6();
    if (e == hipSuccess) e = ntt_limits_wide_v7();
    if (e == hipSuccess) e = ntt_limits_wide_v8();
    if (e == hipSuccess) e = ntt_limits_wide_v9();
    if (e == hipSuccess) e = ntt_limits_wide_v10();
    if (e == hipSuccess) e = ntt_limits_wide_v11();
    if (e == hipSuccess) e = ntt_limits_wide_v12();
    if (e == hipSuccess) e = ntt_limits_wide_v13();
    if (e == hipSuccess) e = ntt_limits_wide_v14();
    if (e == hipSuccess) e = ntt_limits_wide_v15();
    return e;
}

}  // namespace pg
