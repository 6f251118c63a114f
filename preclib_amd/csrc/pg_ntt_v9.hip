// pg_ntt_v9.hip -- NTT plan variant 9 (see plan_l2 / plan_lt / plan_rb in pg_kernels.h)
// for the wide digit mode, in its own translation unit so the variants
// compile in parallel.
#include "pg_ntt_impl.h"

namespace pg {
#define PG_VCASE(N) case N: return launch_variant<N, 9>(L, stream, hook);
#define PG_VLIM(N) if (e == hipSuccess) e = limits_variant<N, 9>();
PG_NTT_VARIANT_FUNCS(9)
}  // namespace pg
