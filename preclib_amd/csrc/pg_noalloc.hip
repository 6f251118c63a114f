// pg_noalloc.hip -- pg_mul_dev_static: the device-resident NTT product with
// every scratch buffer (spectra, product residues, limb sums, carry chunk
// aggregates) carved from one static device array instead of the library's
// pool, for callers that cannot allocate: code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc.  It runs the same
// kernels as pg_mul_dev with PG_PATH_NTT (pg_ntt_impl.h, pg_carry.hip).
//
// The array is `scratch` below, STATIC_SCRATCH_BYTES (48 MiB) in the code
// object: every GPU the library runs on holds its own copy until the process
// ends.  Its device address is looked up once per device at set-up
// (static_scratch_setup, from pg_init) and kept in g_scratch[device]; host
// code uses only that address, never the array's name.  Set-up also builds
// the twiddle tables of every wide-mode plan up to 2^STATIC_MAX_LOG points,
// so the entry point finds them instead of creating them.
//
// Every call shares the array and leaves its contents behind; no word of it
// needs clearing, since every word a call reads it has written first.  Calls
// are ordered on the library's stream, so two calls never use the array at
// once; a caller that runs its own work on another stream must order it
// after the product (pg_sync) before reading r.
#include <algorithm>
#include "pg_batch.h"
#include "pg_runtime.h"
#include "pgpu.h"

namespace pg {

constexpr int STATIC_MAX_LOG = 20;   // longest wide-mode transform: 2^20 points
constexpr size_t STATIC_MAX_LIMBS = (size_t)1 << (STATIC_MAX_LOG - 1);   // an + bn <= 2^19
constexpr size_t STATIC_SCRATCH_BYTES = (size_t)48 << 20;
constexpr int STATIC_MAX_DEVICES = 64;

static __device__ __align__(256) uint8_t scratch[STATIC_SCRATCH_BYTES];
static uint8_t *g_scratch[STATIC_MAX_DEVICES];   // device address of `scratch` per device

static size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes of the carve-out for a wide-mode transform of 2^logn points and a
// product of out_len limbs (same buffers as ntt_mul_device takes from the pool).
static size_t carve_bytes(int logn, size_t out_len) {
    const size_t n = (size_t)1 << logn, m = n / 2;
    const size_t nchunks = (std::max(m, out_len) + CARRY_CHUNK - 1) / CARRY_CHUNK;
    return up256(2 * 3 * n * 4) + up256(3 * n * 4) + 2 * up256(m * 8) + up256(nchunks);
}
// two base-2^32 digits per limb, planned as if the top limbs were full
static_assert(STATIC_MAX_LIMBS * 2 == ((size_t)1 << STATIC_MAX_LOG), "limb limit matches the transform limit");

int static_scratch_setup(int device) {
    if (device < 0 || device >= STATIC_MAX_DEVICES) return set_error(PG_EINVAL, "pg_init: device index too large");
    if (carve_bytes(STATIC_MAX_LOG, STATIC_MAX_LIMBS) > STATIC_SCRATCH_BYTES)
        return set_error(PG_EINTERNAL, "static scratch smaller than its largest product");
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return set_error(PG_EDEVICE, std::string("pg_init: static scratch address: ") + hipGetErrorString(e));
    g_scratch[device] = (uint8_t *)p;
    // twiddle tables of every plan variant of every length the entry point takes
    for (int logn = NTT_MIN_LOG; logn <= STATIC_MAX_LOG; ++logn)
        for (int v = 0; v < NTT_VARIANTS; ++v) {
            if (logn < NTT_VARIANT_MIN_LOG && v) break;
            if (!plan_distinct(logn, v)) continue;
            NttLaunch L{};
            L.logn = logn;
            L.mode = MODE_WIDE;
            L.variant = v;
            if (const int st = fill_ntt_tables(L, true)) return st;
        }
    return PG_OK;
}

static void no_hook(void *, const char *) {}
static void no_hook_end(void *, const char *, double) {}

}  // namespace pg

using namespace pg;

extern "C" int pg_mul_dev_static(uint64_t *r, const uint64_t *a, size_t an, const uint64_t *b, size_t bn) try {
    std::lock_guard<std::recursive_mutex> lk(ctx().mu);
    Context &C = ctx();
    if (!C.stream || C.device < 0 || C.device >= STATIC_MAX_DEVICES || !g_scratch[C.device])
        return set_error(PG_ENOINIT, "pg_init not called");
    if (!r || !a || !b || !an || !bn) return set_error(PG_EINVAL, "pg_mul_dev_static: null pointer or empty operand");
    if (std::max(an, bn) <= DISPATCH_SMALL_HI)
        return set_error(PG_ERANGE, "pg_mul_dev_static: operands of at most 192 limbs are below the NTT range");
    if (an + bn > STATIC_MAX_LIMBS)
        return set_error(PG_ERANGE, "pg_mul_dev_static: an + bn above 2^19 limbs (the static scratch holds "
                                    "transforms of at most 2^20 points)");
    // planned with full top limbs, as pg_mul_dev does (no device read)
    NttLaunch L;
    if (const int st = ntt_plan(L, a, an, ~0ull, b, bn, ~0ull, MODE_WIDE, false)) return st;
    const size_t n = (size_t)1 << L.logn, m = n / 2;
    if (L.logn > STATIC_MAX_LOG || carve_bytes(L.logn, an + bn) > STATIC_SCRATCH_BYTES)
        return set_error(PG_ERANGE, "pg_mul_dev_static: transform longer than the static scratch holds");
    uint8_t *p = g_scratch[C.device];
    auto take = [&](size_t bytes) {
        uint8_t *q = p;
        p += up256(bytes);
        return q;
    };
    if (plan_l1(L.logn, L.variant)) {
        L.one.spec = (uint32_t *)take(2 * 3 * n * 4);
        if (a == b && an == bn) L.one.ext[1] = L.one.spec;   // a square: one forward transform
    }
    L.one.prod = (uint32_t *)take(3 * n * 4);
    L.one.slo = (uint64_t *)take(m * 8);
    L.one.shi = (uint64_t *)take(m * 8);
    L.one.sum_len = m;
    CarryLaunch cl{};
    cl.one.kind = NODE_PLAIN;
    cl.one.parts = 1;
    cl.one.slo = L.one.slo;
    cl.one.shi = L.one.shi;
    cl.one.m = m;
    cl.one.out = r;
    cl.one.out_len = an + bn;
    cl.nchunks = (std::max(m, an + bn) + CARRY_CHUNK - 1) / CARRY_CHUNK;
    cl.nnodes = 1;
    cl.agg = take(cl.nchunks);
    HIP_TRY(launch_ntt(L, C.stream, StageHook{no_hook, no_hook_end, nullptr}));
    HIP_TRY(launch_carry_kernels(cl, C.stream));
    return PG_OK;
} catch (...) {
    return pg::exception_status();
}
