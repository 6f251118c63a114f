// Prints the NTT plan constants the column kernels' indexing depends on (the
// inverse kernel's pipelined form, the forward kernel's workgroup), one line
// per (transform length, plan variant), from the same constexpr functions the
// kernels use (tests/test_colinv_pipe_bounds.py).
#include <cstdio>
#include "pg_ntt_impl.h"

using namespace pg;

template <int LOGN, int V>
static void row() {
    using TW = ColTw<plan_l1(LOGN, V), plan_lt(LOGN, V), plan_rb(LOGN, V), COLINV_THREADS>;
    using FW = ColTw<plan_l1(LOGN, V), plan_lt(LOGN, V), plan_rb(LOGN, V), col_threads(LOGN, V)>;
    std::printf("%d %d %d %d %d %d %d %d %d %d %zu %d %d %d %d\n", LOGN, V, plan_l1(LOGN, V), plan_l2(LOGN, V),
                plan_lt(LOGN, V), plan_rb(LOGN, V), (int)TW::STEP, TW::D, col_inv_waves<LOGN, V>(),
                (int)col_inv_pipe<LOGN, V>(), col_inv_pipe_lds<LOGN, V>(), col_inv_tw_words<LOGN, V>(),
                col_threads(LOGN, V), (int)FW::STEP, FW::D);
}

template <int LOGN, int V>
static void rows() {
    row<LOGN, V>();
    if constexpr (V + 1 < NTT_VARIANTS) rows<LOGN, V + 1>();
}

template <int LOGN>
static void sizes() {
    rows<LOGN, 0>();
    if constexpr (LOGN < NTT_MAX_LOG) sizes<LOGN + 1>();
}

int main() {
    std::printf("# logn v l1 l2 lt rb step d waves pipe lds tw_words fwd_threads fwd_step fwd_d threads=%d\n", COLINV_THREADS);
    sizes<NTT_MIN_LOG>();
    return 0;
}
