// pg_runtime.h -- host-side device context of libpgpu.so.
#pragma once
#include <hip/hip_runtime.h>
#include <map>
#include <cstdlib>
#include <memory>
#include <new>
#include <utility>
#include <mutex>
#include <string>
#include <vector>
#include "pg_kernels.h"

namespace pg {

// Page-locked host memory (hipHostMalloc) for descriptor images: an upload
// from it is a plain DMA that returns at once, where one from pageable memory
// is staged and can hold the host until the stream drains.  Value
// construction leaves the bytes as they are, so resizing an image does not
// zero it (every byte that matters is written by its user).
template <class T>
struct PinnedAlloc {
    using value_type = T;
    PinnedAlloc() = default;
    template <class U>
    PinnedAlloc(const PinnedAlloc<U> &) noexcept {}
    // A 16-byte header records where the block came from: page-locked memory,
    // or (when the runtime refuses more of it) the ordinary heap.
    T *allocate(size_t n) {
        const size_t bytes = n * sizeof(T) + 16;
        void *p = nullptr;
        uint64_t from = 1;
        if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess || !p) {
            (void)hipGetLastError();   // not a device error: clear it
            from = 2;
            p = std::malloc(bytes);
            if (!p) throw std::bad_alloc();
        }
        *static_cast<uint64_t *>(p) = from;
        return reinterpret_cast<T *>(static_cast<char *>(p) + 16);
    }
    void deallocate(T *q, size_t) noexcept {
        char *p = reinterpret_cast<char *>(q) - 16;
        if (*reinterpret_cast<uint64_t *>(p) == 1) (void)hipHostFree(p);
        else std::free(p);
    }
    template <class U>
    void construct(U *p) noexcept {
        ::new ((void *)p) U;
    }
    template <class U, class... A>
    void construct(U *p, A &&...a) {
        ::new ((void *)p) U(std::forward<A>(a)...);
    }
    template <class U>
    bool operator==(const PinnedAlloc<U> &) const noexcept { return true; }
    template <class U>
    bool operator!=(const PinnedAlloc<U> &) const noexcept { return false; }
};
using HostImg = std::vector<char, PinnedAlloc<char>>;

// NTT primes and generators (src/mul_ntt.cpp:29-34).
constexpr uint32_t PRIME[3] = {469762049u, 1811939329u, 2013265921u};
constexpr uint32_t GEN[3] = {3u, 13u, 31u};
static_assert(PRIME[LAZY_PRIME] < LAZY_PRIME_LIMIT, "lazy residues need 4p < 2^32");

uint32_t powmod(uint32_t b, uint64_t e, uint32_t p);
uint32_t shoup_q(uint32_t w, uint32_t p);
uint32_t to_mont(uint32_t x, uint32_t p);
PrimeK make_prime(int idx);

struct Status {
    int code;
    std::string msg;
};

// Caching device allocator with a byte budget.  Blocks are kept in free
// lists keyed by their rounded size; nothing is returned to HIP until
// shutdown, so steady-state products never call hipMalloc.
class Pool {
public:
    void set_budget(size_t b) { budget_ = b; }
    void set_stream(hipStream_t s) { stream_ = s; }
    void *get(size_t bytes);
    void put(void *p);
    void release_all();
    size_t in_use() const { return in_use_; }
    size_t reserved() const { return reserved_; }
    size_t budget() const { return budget_; }

private:
    static size_t round(size_t b);
    std::multimap<size_t, void *> free_;
    std::map<void *, size_t> live_;
    size_t budget_ = 0, in_use_ = 0, reserved_ = 0;
    hipStream_t stream_ = nullptr;
};

// Device twiddle tables, built once per (prime, block size, radix) and per
// (prime, n) for the four-step tables.
struct TwTables {
    const uint2 *fwd = nullptr, *inv = nullptr;
};
struct GlobalTw {
    const uint32_t *glo = nullptr, *ghi = nullptr, *ilo = nullptr, *ihi = nullptr;
    int lobits = 0;
};

struct FftTw {
    const double2 *fwd = nullptr, *inv = nullptr;
};
struct FftGlobal {
    const double2 *glo = nullptr, *ghi = nullptr;
    int lobits = 0;
};

struct StageStat {
    uint64_t launches = 0;
    double ms = 0.0, bytes = 0.0;
};

class Context {
public:
    int device = -1;
    hipStream_t stream = nullptr;
    Pool pool;
    std::recursive_mutex mu;
    PrimeK pk[3];
    CrtK crt{};

    int init(int dev, size_t budget);
    void shutdown();

    const TwTables &tw(int prime, int logB, int lr);
    const GlobalTw &gtw(int prime, int logn);
    // the same tables if they exist, else nullptr (never allocates)
    const TwTables *tw_find(int prime, int logB, int lr) const;
    const GlobalTw *gtw_find(int prime, int logn) const;
    const FftTw &ftw(int logB, int lr);
    const FftGlobal &fgtw(int logn);
    FftK fk{};
    uint64_t *maxerr = nullptr;   // device word for the FFT rounding distance
    // NTT plan variant per transform length (wide mode), from the autotuner's
    // table (ntt_plans.txt next to libpgpu.so, or $PGPU_NTT_PLANS)
    int ntt_variant_tab[NTT_MAX_LOG + 1] = {};
    int ntt_variant(int logn, int mode) const {
        return (mode == MODE_WIDE && logn >= 0 && logn <= NTT_MAX_LOG) ? ntt_variant_tab[logn] : 0;
    }
    // Plan variant of a batched launch (pg_batch) of `points` transform
    // points in all: launches of at least BATCH_PLAN_MIN_POINTS (one 2^21-point
    // transform's worth, about what fills the device) are throughput-bound and
    // take the batch table ("batch logn variant" lines of the same file), tuned
    // on the 4096-product batch; smaller ones, and sizes the batch table leaves
    // at -1, take the single-product variant.
    static constexpr size_t BATCH_PLAN_MIN_POINTS = (size_t)1 << 21;
    int ntt_batch_tab[NTT_MAX_LOG + 1] = {};
    int ntt_batch_variant(int logn, int mode, size_t points) const {
        if (mode == MODE_WIDE && logn >= 0 && logn <= NTT_MAX_LOG && points >= BATCH_PLAN_MIN_POINTS &&
            ntt_batch_tab[logn] >= 0)
            return ntt_batch_tab[logn];
        return ntt_variant(logn, mode);
    }
    void load_plan_table();
    // workgroup order of the NTT kernels (NttLaunch::order); $PGPU_NTT_ORDER overrides
    int ntt_order = COLF_OP_PRIME | ROW_PRIME_REV;
    static constexpr size_t NWORDS = 4096;
    uint64_t *words = nullptr;    // device flags / small results of linear steps

    // Host descriptor images handed to hipMemcpyAsync stay alive until the
    // next sync() of the stream; their buffers are then kept for reuse, so a
    // later image of similar size needs neither fresh pages nor a zero fill.
    // Every entry point that waits for the stream recycles them (sync() or
    // wait()); a run of asynchronous calls that holds more than
    // HELD_MAX bytes waits for the stream in hold().
    static constexpr size_t HELD_MAX = (size_t)256 << 20;
    std::vector<HostImg> held, spare;
    size_t held_bytes = 0;
    HostImg take_img(size_t bytes);
    void hold(HostImg &&img);
    void recycle_held();
    int sync();
    // hipStreamSynchronize, then recycle_held (without touching the error text)
    hipError_t wait();

    // timing
    bool timing = false;
    std::map<std::string, StageStat> stats;
    struct Pending {
        std::string name;
        hipEvent_t a, b;
        double bytes;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    void stage_begin(hipEvent_t &a);
    void stage_end(const char *name, hipEvent_t a, double bytes);
    void collect_timing();
    hipEvent_t hook_event = nullptr;
    StageHook hook();

private:
    std::map<long, TwTables> tw_;
    std::map<long, GlobalTw> gtw_;
    std::map<long, FftTw> ftw_;
    std::map<long, FftGlobal> fgtw_;
    std::vector<void *> tables_;   // table allocations (freed at shutdown)
    void *table_alloc(size_t bytes);
};

Context &ctx();
int set_error(int code, const std::string &msg);
// Status of the exception being handled: the C ABI lets none escape (host
// allocation failure -> PG_ENOMEM, anything else -> PG_EINTERNAL).
int exception_status();

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return set_error(PG_EDEVICE, std::string(#expr " failed: ") + hipGetErrorString(e_)); \
    } while (0)

// Product entry points on device pointers (stream-ordered on ctx().stream).
// rp must hold an + bn limbs; it may alias ap or bp.  Returns a status code.
int ntt_mul_device(uint64_t *rp, const uint64_t *ap, size_t an, uint64_t atop,
                   const uint64_t *bp, size_t bn, uint64_t btop, int mode);
// The launch of that product without its scratch pointers (L.one.spec, prod,
// slo, shi): transform length, plan variant, table pointers.  create = false
// uses only tables that already exist (PG_EINTERNAL otherwise).
int ntt_plan(NttLaunch &L, const uint64_t *ap, size_t an, uint64_t atop, const uint64_t *bp, size_t bn,
             uint64_t btop, int mode, bool create);
// Sets up the allocation-free product (pg_mul_dev_static) on the current
// device (pg_noalloc.hip); called by Context::init.
int static_scratch_setup(int device);
// mul_fft on device pointers; *max_round (if non-null) receives the largest
// rounding distance of this product (waits for the device).
int fft_mul_device(uint64_t *rp, const uint64_t *ap, size_t an, uint64_t atop,
                   const uint64_t *bp, size_t bn, uint64_t btop, double *max_round);
// Times every plan variant of each wide-mode transform length 2^lo .. 2^hi
// (one product of that length per variant, reps timed runs, identical
// products required) and keeps the fastest in the context's table.  ms gets
// (hi - lo + 1) * NTT_VARIANTS entries (-1 where a variant equals the default).
int autotune_ntt(int lo, int hi, int reps, double *ms, int *chosen);

// mul_into's routing (src/mul_basic.cpp:238-276) on device pointers; atop /
// btop are the top limbs or ~0 when unknown (full top limbs assumed).
int mul_device(uint64_t *rp, const uint64_t *ap, size_t an, uint64_t atop, const uint64_t *bp,
               size_t bn, uint64_t btop, int path, double *fft_max_round);

}  // namespace pg
