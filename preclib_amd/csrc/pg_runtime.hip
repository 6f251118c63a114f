// pg_runtime.hip -- device context, memory pool, plans and twiddle tables of
// libpgpu.so, plus the NTT product driver.
#include "pg_runtime.h"
#include <dlfcn.h>
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <exception>
#include "pgpu.h"

namespace pg {

static thread_local std::string g_err;
int set_error(int code, const std::string &msg) {
    g_err = msg;
    return code;
}
const char *last_error() { return g_err.c_str(); }

int exception_status() {
    try {
        throw;
    } catch (const std::bad_alloc &) {
        return set_error(PG_ENOMEM, "host memory exhausted");
    } catch (const std::exception &e) {
        return set_error(PG_EINTERNAL, std::string("internal error: ") + e.what());
    } catch (...) {
        return set_error(PG_EINTERNAL, "internal error");
    }
}


uint32_t powmod(uint32_t b, uint64_t e, uint32_t p) {
    uint64_t r = 1, x = b % p;
    while (e) {
        if (e & 1) r = r * x % p;
        x = x * x % p;
        e >>= 1;
    }
    return (uint32_t)r;
}
uint32_t shoup_q(uint32_t w, uint32_t p) { return (uint32_t)(((uint64_t)w << 32) / p); }
uint32_t to_mont(uint32_t x, uint32_t p) { return (uint32_t)(((uint64_t)x << 32) % p); }

PrimeK make_prime(int idx) {
    PrimeK k{};
    const uint32_t p = PRIME[idx];
    uint32_t x = 1;
    for (int i = 0; i < 5; ++i) x *= 2u - p * x;   // p^-1 mod 2^32
    k.p = p;
    k.ninv = 0u - x;
    k.np = 0u - p;
    const uint32_t w = powmod(GEN[idx], (p - 1) / 32, p);
    const uint32_t iw = powmod(w, p - 2, p);
    uint32_t a = 1, b = 1;
    for (int i = 0; i < 16; ++i) {
        k.w32[i] = a; k.w32q[i] = shoup_q(a, p);
        k.iw32[i] = b; k.iw32q[i] = shoup_q(b, p);
        a = (uint64_t)a * w % p;
        b = (uint64_t)b * iw % p;
    }
    return k;
}

// ---- pool --------------------------------------------------------------------

size_t Pool::round(size_t b) {
    if (b < 4096) return 4096;
    // power-of-two classes split in eighths: waste below 12.5%
    size_t top = 1;
    while ((top << 1) <= b) top <<= 1;
    const size_t step = std::max<size_t>(top >> 3, 4096);
    return (b + step - 1) / step * step;
}

void *Pool::get(size_t bytes) {
    const size_t r = round(bytes);
    auto it = free_.find(r);
    void *p = nullptr;
    if (it != free_.end()) {
        p = it->second;
        free_.erase(it);
    } else {
        if (reserved_ + r > budget_ && !free_.empty()) {
            // cached blocks may still be named by queued work: drain first
            if (stream_) hipStreamSynchronize(stream_);
            for (auto &f : free_) {
                hipFree(f.second);
                reserved_ -= f.first;
            }
            free_.clear();
        }
        if (reserved_ + r > budget_) return nullptr;
        if (hipMalloc(&p, r) != hipSuccess) return nullptr;
        reserved_ += r;
    }
    live_[p] = r;
    in_use_ += r;
    return p;
}

void Pool::put(void *p) {
    if (!p) return;
    auto it = live_.find(p);
    if (it == live_.end()) return;
    in_use_ -= it->second;
    free_.insert({it->second, p});
    live_.erase(it);
}

void Pool::release_all() {
    for (auto &f : free_) hipFree(f.second);
    for (auto &l : live_) hipFree(l.first);
    free_.clear();
    live_.clear();
    reserved_ = in_use_ = 0;
}

// ---- context --------------------------------------------------------------------

static Context g_ctx;
Context &ctx() { return g_ctx; }

int Context::init(int dev, size_t budget) {
    std::lock_guard<std::recursive_mutex> lk(mu);
    if (device == dev && stream) return PG_OK;
    if (stream) shutdown();
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (dev < 0 || dev >= n) return set_error(PG_EINVAL, "pg_init: no such device");
    HIP_TRY(hipSetDevice(dev));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return set_error(PG_EDEVICE, std::string("pg_init: built for gfx950, device is ") + prop.gcnArchName);
    HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    size_t fr = 0, tot = 0;
    HIP_TRY(hipMemGetInfo(&fr, &tot));
    pool.set_budget(budget ? budget : fr / 10 * 9);
    pool.set_stream(stream);
    for (int i = 0; i < 3; ++i) pk[i] = make_prime(i);
    crt.inv12 = powmod(PRIME[0], PRIME[1] - 2, PRIME[1]);
    crt.inv12q = shoup_q(crt.inv12, PRIME[1]);
    crt.p1_mod3 = PRIME[0] % PRIME[2];
    crt.p1_mod3q = shoup_q(crt.p1_mod3, PRIME[2]);
    crt.inv123 = powmod((uint32_t)((uint64_t)PRIME[0] * PRIME[1] % PRIME[2]), PRIME[2] - 2, PRIME[2]);
    crt.inv123q = shoup_q(crt.inv123, PRIME[2]);
    for (int k = 0; k < 3; ++k) crt.red[k] = (uint32_t)((((uint64_t)1) << 32) / PRIME[k]);
    HIP_TRY(ntt_set_lds_limits());
    HIP_TRY(fft_set_lds_limits());
    HIP_TRY(hipMalloc(&maxerr, 256));
    HIP_TRY(hipMalloc(&words, NWORDS * sizeof(uint64_t)));
    load_plan_table();
    if (const char *e = std::getenv("PGPU_NTT_ORDER")) ntt_order = std::atoi(e) & 7;
    device = dev;
    if (const int st = static_scratch_setup(dev)) {
        const std::string msg = last_error();
        shutdown();
        return set_error(st, msg);
    }
    {
        const long double PI = 3.141592653589793238462643383279502884L;
        for (int k = 0; k < 8; ++k) {
            const long double a = 2.0L * PI * k / 16.0L;
            fk.w16[k] = make_double2((double)cosl(a), (double)sinl(a));
            fk.iw16[k] = make_double2((double)cosl(a), -(double)sinl(a));
        }
    }
    device = dev;
    return PG_OK;
}

void Context::shutdown() {
    std::lock_guard<std::recursive_mutex> lk(mu);
    if (!stream) return;
    hipStreamSynchronize(stream);
    collect_timing();
    for (auto e : event_pool) hipEventDestroy(e);
    event_pool.clear();
    for (void *p : tables_) hipFree(p);
    tables_.clear();
    tw_.clear();
    gtw_.clear();
    ftw_.clear();
    fgtw_.clear();
    if (maxerr) hipFree(maxerr);
    maxerr = nullptr;
    if (words) hipFree(words);
    words = nullptr;
    held.clear();
    held_bytes = 0;
    spare.clear();
    pool.release_all();
    hipStreamDestroy(stream);
    stream = nullptr;
    device = -1;
}

// "logn variant" lines (single products) and "batch logn variant" lines
// (batched launches); unknown or invalid entries keep the default plan, and
// the single-product plan for batches.
void Context::load_plan_table() {
    for (int &v : ntt_variant_tab) v = 0;
    for (int &v : ntt_batch_tab) v = -1;
    std::string path;
    if (const char *e = std::getenv("PGPU_NTT_PLANS")) {
        path = e;
    } else {
        Dl_info info{};
        if (dladdr((const void *)&ctx, &info) && info.dli_fname) {
            path = info.dli_fname;
            const size_t slash = path.rfind('/');
            path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/ntt_plans.txt";
        }
    }
    FILE *f = path.empty() ? nullptr : std::fopen(path.c_str(), "r");
    if (!f) return;
    int logn = 0, v = 0;
    char line[256];
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] == '#') continue;
        const bool batch = std::sscanf(line, "batch %d %d", &logn, &v) == 2;
        if ((batch || std::sscanf(line, "%d %d", &logn, &v) == 2) && logn >= NTT_VARIANT_MIN_LOG &&
            logn <= NTT_MAX_LOG && v >= 0 && v < NTT_VARIANTS)
            (batch ? ntt_batch_tab : ntt_variant_tab)[logn] = v;
    }
    std::fclose(f);
}

int Context::sync() {
    if (hipStreamSynchronize(stream) != hipSuccess) return set_error(PG_EDEVICE, "device error");
    recycle_held();
    return PG_OK;
}

hipError_t Context::wait() {
    const hipError_t e = hipStreamSynchronize(stream);
    if (e == hipSuccess) recycle_held();
    return e;
}

void Context::hold(HostImg &&img) {
    held_bytes += img.capacity();
    held.push_back(std::move(img));
    // the uploads from every held image are already enqueued: once the stream
    // has passed them the buffers can be reused
    if (held_bytes > HELD_MAX && hipStreamSynchronize(stream) == hipSuccess) recycle_held();
}

// The stream has passed every upload from the held images: keep their buffers
// (up to HELD_MAX bytes, largest first) for take_img.
void Context::recycle_held() {
    for (auto &h : held) spare.push_back(std::move(h));
    held.clear();
    held_bytes = 0;
    std::sort(spare.begin(), spare.end(), [](const HostImg &a, const HostImg &b) { return a.capacity() > b.capacity(); });
    size_t kept = 0, n = 0;
    while (n < spare.size() && kept + spare[n].capacity() <= HELD_MAX) kept += spare[n++].capacity();
    spare.resize(n);
}

HostImg Context::take_img(size_t bytes) {
    // the smallest spare buffer that holds `bytes` (spare is sorted by
    // decreasing capacity)
    for (size_t i = spare.size(); i-- > 0;) {
        if (spare[i].capacity() >= bytes) {
            HostImg img = std::move(spare[i]);
            spare.erase(spare.begin() + (long)i);
            img.resize(bytes);
            return img;
        }
    }
    HostImg img;
    img.resize(bytes);
    return img;
}

void *Context::table_alloc(size_t bytes) {
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    tables_.push_back(p);
    return p;
}

// Table of pass (logB, lr): entry [k1][j2] = omega_B^(+-j2*k1) with Shoup quotient.
const TwTables &Context::tw(int prime, int logB, int lr) {
    const long key = ((long)prime << 16) | (logB << 8) | lr;
    auto it = tw_.find(key);
    if (it != tw_.end()) return it->second;
    TwTables t;
    const int R = 1 << lr, M = 1 << (logB - lr);
    if (M > 1) {
        const uint32_t p = PRIME[prime];
        const uint32_t w = powmod(GEN[prime], (p - 1) >> logB, p);
        const uint32_t iw = powmod(w, p - 2, p);
        std::vector<uint2> f((size_t)R * M), v((size_t)R * M);
        for (int k1 = 0; k1 < R; ++k1) {
            const uint32_t wk = powmod(w, k1, p), iwk = powmod(iw, k1, p);
            uint32_t a = 1, b = 1;
            for (int j2 = 0; j2 < M; ++j2) {
                f[(size_t)k1 * M + j2] = make_uint2(a, shoup_q(a, p));
                v[(size_t)k1 * M + j2] = make_uint2(b, shoup_q(b, p));
                a = (uint64_t)a * wk % p;
                b = (uint64_t)b * iwk % p;
            }
        }
        uint2 *df = (uint2 *)table_alloc(f.size() * sizeof(uint2));
        uint2 *dv = (uint2 *)table_alloc(v.size() * sizeof(uint2));
        if (df && dv) {
            hipMemcpy(df, f.data(), f.size() * sizeof(uint2), hipMemcpyHostToDevice);
            hipMemcpy(dv, v.data(), v.size() * sizeof(uint2), hipMemcpyHostToDevice);
        }
        t.fwd = df;
        t.inv = dv;
    }
    return tw_[key] = t;
}

// Four-step twiddles omega_n^e = lo[e mod 2^s] * hi[e >> s] in Montgomery form;
// the inverse low table carries the scale n^-1 * 2^32 that undoes both the
// unnormalized inverse passes and the 2^-32 of the pointwise Montgomery product.
const TwTables *Context::tw_find(int prime, int logB, int lr) const {
    auto it = tw_.find(((long)prime << 16) | (logB << 8) | lr);
    return it == tw_.end() ? nullptr : &it->second;
}

const GlobalTw *Context::gtw_find(int prime, int logn) const {
    auto it = gtw_.find(((long)prime << 16) | logn);
    return it == gtw_.end() ? nullptr : &it->second;
}

const GlobalTw &Context::gtw(int prime, int logn) {
    const long key = ((long)prime << 16) | logn;
    auto it = gtw_.find(key);
    if (it != gtw_.end()) return it->second;
    GlobalTw g;
    const uint32_t p = PRIME[prime];
    const int s = (logn + 1) / 2;
    const size_t nlo = (size_t)1 << s, nhi = (size_t)1 << (logn - s);
    const uint32_t w = powmod(GEN[prime], (p - 1) >> logn, p);
    const uint32_t iw = powmod(w, p - 2, p);
    const uint32_t ninv_n = powmod((uint32_t)(((uint64_t)1 << logn) % p), p - 2, p);
    const uint32_t scale = (uint32_t)((uint64_t)ninv_n * to_mont(1, p) % p);
    std::vector<uint32_t> lo(nlo), hi(nhi), ilo(nlo), ihi(nhi);
    uint32_t a = 1, b = 1;
    for (size_t e = 0; e < nlo; ++e) {
        lo[e] = to_mont(a, p);
        ilo[e] = to_mont((uint32_t)((uint64_t)b * scale % p), p);
        a = (uint64_t)a * w % p;
        b = (uint64_t)b * iw % p;
    }
    const uint32_t ws = powmod(w, nlo, p), iws = powmod(iw, nlo, p);
    a = b = 1;
    for (size_t e = 0; e < nhi; ++e) {
        hi[e] = to_mont(a, p);
        ihi[e] = to_mont(b, p);
        a = (uint64_t)a * ws % p;
        b = (uint64_t)b * iws % p;
    }
    uint32_t *d = (uint32_t *)table_alloc((2 * nlo + 2 * nhi) * sizeof(uint32_t));
    if (d) {
        hipMemcpy(d, lo.data(), nlo * 4, hipMemcpyHostToDevice);
        hipMemcpy(d + nlo, ilo.data(), nlo * 4, hipMemcpyHostToDevice);
        hipMemcpy(d + 2 * nlo, hi.data(), nhi * 4, hipMemcpyHostToDevice);
        hipMemcpy(d + 2 * nlo + nhi, ihi.data(), nhi * 4, hipMemcpyHostToDevice);
        g.glo = d;
        g.ilo = d + nlo;
        g.ghi = d + 2 * nlo;
        g.ihi = d + 2 * nlo + nhi;
    }
    g.lobits = s;
    return gtw_[key] = g;
}

// FFT pass table [k1][j2] = exp(+-2 pi i j2 k1 / B), from long double.
const FftTw &Context::ftw(int logB, int lr) {
    const long key = ((long)logB << 8) | lr;
    auto it = ftw_.find(key);
    if (it != ftw_.end()) return it->second;
    FftTw t;
    const int R = 1 << lr, M = 1 << (logB - lr);
    if (M > 1) {
        const long double PI = 3.141592653589793238462643383279502884L;
        const long double B = (long double)(1L << logB);
        std::vector<double2> f((size_t)R * M), v((size_t)R * M);
        for (int k1 = 0; k1 < R; ++k1)
            for (int j2 = 0; j2 < M; ++j2) {
                const long double a = 2.0L * PI * (long double)((long)j2 * k1) / B;
                const double c = (double)cosl(a), sn = (double)sinl(a);
                f[(size_t)k1 * M + j2] = make_double2(c, sn);
                v[(size_t)k1 * M + j2] = make_double2(c, -sn);
            }
        double2 *df = (double2 *)table_alloc(f.size() * sizeof(double2));
        double2 *dv = (double2 *)table_alloc(v.size() * sizeof(double2));
        if (df && dv) {
            hipMemcpy(df, f.data(), f.size() * sizeof(double2), hipMemcpyHostToDevice);
            hipMemcpy(dv, v.data(), v.size() * sizeof(double2), hipMemcpyHostToDevice);
        }
        t.fwd = df;
        t.inv = dv;
    }
    return ftw_[key] = t;
}

// omega_n^e = glo[e mod 2^s] * ghi[e >> s], omega_n = exp(+2 pi i / n).
const FftGlobal &Context::fgtw(int logn) {
    auto it = fgtw_.find(logn);
    if (it != fgtw_.end()) return it->second;
    FftGlobal g;
    const int s = (logn + 1) / 2;
    const size_t nlo = (size_t)1 << s, nhi = (size_t)1 << (logn - s);
    const long double PI = 3.141592653589793238462643383279502884L;
    const long double n = (long double)(1L << logn);
    std::vector<double2> lo(nlo), hi(nhi);
    for (size_t e = 0; e < nlo; ++e) {
        const long double a = 2.0L * PI * (long double)e / n;
        lo[e] = make_double2((double)cosl(a), (double)sinl(a));
    }
    for (size_t e = 0; e < nhi; ++e) {
        const long double a = 2.0L * PI * (long double)(e * nlo) / n;
        hi[e] = make_double2((double)cosl(a), (double)sinl(a));
    }
    double2 *d = (double2 *)table_alloc((nlo + nhi) * sizeof(double2));
    if (d) {
        hipMemcpy(d, lo.data(), nlo * sizeof(double2), hipMemcpyHostToDevice);
        hipMemcpy(d + nlo, hi.data(), nhi * sizeof(double2), hipMemcpyHostToDevice);
        g.glo = d;
        g.ghi = d + nlo;
    }
    g.lobits = s;
    return fgtw_[logn] = g;
}

// ---- timing -------------------------------------------------------------------

void Context::stage_begin(hipEvent_t &a) {
    if (!timing) return;
    if (event_pool.empty()) {
        hipEvent_t e;
        hipEventCreate(&e);
        event_pool.push_back(e);
    }
    a = event_pool.back();
    event_pool.pop_back();
    hipEventRecord(a, stream);
}

void Context::stage_end(const char *name, hipEvent_t a, double bytes) {
    if (!timing) return;
    if (event_pool.empty()) {
        hipEvent_t e;
        hipEventCreate(&e);
        event_pool.push_back(e);
    }
    hipEvent_t b = event_pool.back();
    event_pool.pop_back();
    hipEventRecord(b, stream);
    pending.push_back({name, a, b, bytes});
    if (pending.size() > 4096) collect_timing();
}

static void hook_begin(void *c, const char *) {
    Context *C = (Context *)c;
    C->stage_begin(C->hook_event);
}
static void hook_end(void *c, const char *name, double bytes) {
    Context *C = (Context *)c;
    C->stage_end(name, C->hook_event, bytes);
}
StageHook Context::hook() { return StageHook{hook_begin, hook_end, this}; }

void Context::collect_timing() {
    for (auto &p : pending) {
        hipEventSynchronize(p.b);
        float ms = 0.f;
        hipEventElapsedTime(&ms, p.a, p.b);
        auto &s = stats[p.name];
        s.launches++;
        s.ms += ms;
        s.bytes += p.bytes;
        event_pool.push_back(p.a);
        event_pool.push_back(p.b);
    }
    pending.clear();
}

// ---- NTT product driver ----------------------------------------------------

static int ndig16(uint64_t top) {
    int d = 0;
    while (top) { ++d; top >>= 16; }
    return d;
}

struct Scratch {
    std::vector<void *> bufs;
    ~Scratch() {
        for (void *p : bufs) ctx().pool.put(p);
    }
    template <class T>
    T *get(size_t n) {
        void *p = ctx().pool.get(n * sizeof(T));
        if (p) bufs.push_back(p);
        return (T *)p;
    }
};

int launch_carry(const uint64_t *slo, const uint64_t *shi, size_t m, uint64_t *out, size_t out_len,
                 Scratch &s, int parts = 1, int pack_log = 0) {
    Context &C = ctx();
    CarryLaunch cl{};
    cl.one.kind = NODE_PLAIN;
    cl.one.parts = parts;
    cl.one.pack_log = pack_log;
    cl.one.slo = slo;
    cl.one.shi = shi;
    cl.one.m = m;
    cl.one.out = out;
    cl.one.out_len = out_len;
    cl.nchunks = (std::max(m, out_len) + CARRY_CHUNK - 1) / CARRY_CHUNK;
    cl.nnodes = 1;
    cl.agg = s.get<uint8_t>(cl.nchunks);
    if (!cl.agg) return set_error(PG_ENOMEM, "carry scratch");
    hipEvent_t e0 = nullptr;
    C.stage_begin(e0);
    HIP_TRY(launch_carry_kernels(cl, C.stream));
    C.stage_end("carry", e0, (double)m * 32.0 + (double)out_len * 8.0);
    return PG_OK;
}

int fill_ntt_tables(NttLaunch &L, bool create) {
    Context &C = ctx();
    const int logn = L.logn, NP = mode_primes(L.mode);
    L.nprimes = NP;
    L.crt = C.crt;
    const int v = L.mode == MODE_WIDE ? L.variant : 0;
    const int l1 = plan_l1(logn, v), l2 = plan_l2(logn, v), rb = plan_rb(logn, v);
    auto pass = [&](int k, int logB, int lr) -> const TwTables * {
        return create ? &C.tw(k, logB, lr) : C.tw_find(k, logB, lr);
    };
    for (int k = 0; k < NP; ++k) {
        L.pk[k] = C.pk[k];
        int logB = l1;
        for (int i = 0; i < split_np(l1, rb); ++i) {
            const TwTables *t = pass(k, logB, split_lr(l1, i, rb));
            if (!t) return set_error(PG_EINTERNAL, "twiddle tables not prepared");
            L.col_tw[k][i] = t->fwd; L.col_itw[k][i] = t->inv;
            logB -= split_lr(l1, i, rb);
        }
        logB = l2;
        for (int i = 0; i < split_np(l2, rb); ++i) {
            const TwTables *t = pass(k, logB, split_lr(l2, i, rb));
            if (!t) return set_error(PG_EINTERNAL, "twiddle tables not prepared");
            L.row_tw[k][i] = t->fwd; L.row_itw[k][i] = t->inv;
            logB -= split_lr(l2, i, rb);
        }
        const GlobalTw *g = create ? &C.gtw(k, logn) : C.gtw_find(k, logn);
        if (!g) return set_error(PG_EINTERNAL, "twiddle tables not prepared");
        if (!g->glo) return set_error(PG_ENOMEM, "twiddle tables");
        L.glo[k] = g->glo; L.ghi[k] = g->ghi; L.ilo[k] = g->ilo; L.ihi[k] = g->ihi; L.lobits = g->lobits;
    }
    return PG_OK;
}

int fill_fft_tables(FftLaunch &L) {
    Context &C = ctx();
    const int logn = L.logn, l1 = fft_l1(logn), l2 = fft_l2(logn);
    L.k = C.fk;
    int logB = l1;
    for (int i = 0; i < fsplit_np(l1); ++i) {
        const FftTw &t = C.ftw(logB, fsplit_lr(l1, i));
        L.col_tw[i] = t.fwd; L.col_itw[i] = t.inv;
        logB -= fsplit_lr(l1, i);
    }
    logB = l2;
    for (int i = 0; i < fsplit_np(l2); ++i) {
        const FftTw &t = C.ftw(logB, fsplit_lr(l2, i));
        L.row_tw[i] = t.fwd; L.row_itw[i] = t.inv;
        logB -= fsplit_lr(l2, i);
    }
    const FftGlobal &g = C.fgtw(logn);
    if (!g.glo) return set_error(PG_ENOMEM, "FFT twiddle tables");
    L.glo = g.glo; L.ghi = g.ghi; L.lobits = g.lobits;
    return PG_OK;
}

int launch_carry_nodes(size_t nnodes, double bytes, const CarryNode *dnodes, const uint32_t *dchunk,
                       size_t nchunks, uint8_t *agg, int node_kind) {
    Context &C = ctx();
    CarryLaunch cl{};
    cl.node_kind = node_kind;
    cl.nodes = dnodes;
    cl.chunk_node = dchunk;
    cl.nchunks = nchunks;
    cl.nnodes = nnodes;
    cl.agg = agg;
    hipEvent_t e0 = nullptr;
    C.stage_begin(e0);
    HIP_TRY(launch_carry_kernels(cl, C.stream));
    C.stage_end("carry", e0, bytes);
    return PG_OK;
}

int ntt_plan(NttLaunch &L, const uint64_t *ap, size_t an, uint64_t atop, const uint64_t *bp, size_t bn,
             uint64_t btop, int mode, bool create) {
    Context &C = ctx();
    if (!C.stream) return set_error(PG_ENOINIT, "pg_init not called");
    if (!an || !bn || !atop || !btop) return set_error(PG_EINVAL, "operands must be trimmed and nonzero");
    // mul_ntt's caps (mul_ntt.cpp:497-506), evaluated on its own base-2^16 digits
    const size_t d16a = 4 * (an - 1) + ndig16(atop), d16b = 4 * (bn - 1) + ndig16(btop);
    if (std::max(an, bn) > ((size_t)1 << 24)) return set_error(PG_ERANGE, "NTT limit 2^24 limbs");
    size_t n16 = 1;
    while (n16 < d16a + d16b) n16 <<= 1;
    if (n16 > ((size_t)1 << 26)) return set_error(PG_ERANGE, "NTT limit 2^26 points");
    size_t da = d16a, db = d16b;
    if (mode == MODE_WIDE) {
        da = 2 * (an - 1) + ((atop >> 32) ? 2 : 1);
        db = 2 * (bn - 1) + ((btop >> 32) ? 2 : 1);
    }
    int logn = 0;
    while (((size_t)1 << logn) < da + db) ++logn;
    if (logn < NTT_MIN_LOG) logn = NTT_MIN_LOG;
    if (logn > NTT_MAX_LOG) return set_error(PG_ERANGE, "transform too long");
    L = NttLaunch{};
    L.one.a = ap; L.one.b = bp; L.one.an = an; L.one.bn = bn;
    L.io_bytes = (double)(an + bn) * 16.0;
    L.logn = logn; L.mode = mode;
    L.variant = C.ntt_variant(logn, mode);
    L.order = C.ntt_order;
    return fill_ntt_tables(L, create);
}

int ntt_mul_device(uint64_t *rp, const uint64_t *ap, size_t an, uint64_t atop, const uint64_t *bp,
                   size_t bn, uint64_t btop, int mode) {
    Context &C = ctx();
    NttLaunch L;
    { int st = ntt_plan(L, ap, an, atop, bp, bn, btop, mode, true); if (st) return st; }
    const size_t n = (size_t)1 << L.logn;
    const int NP = mode_primes(mode), DPL = 64 / mode_digit_bits(mode);
    Scratch s;
    const int l1 = plan_l1(L.logn, mode == MODE_WIDE ? L.variant : 0);
    // a square transforms its operand once: the row kernel reads operand b's
    // spectra from operand a's
    const bool sqr = ap == bp && an == bn;
    if (l1) L.one.spec = s.get<uint32_t>((sqr ? 1 : 2) * NP * n);
    if (l1 && sqr) L.one.ext[1] = L.one.spec;
    L.one.prod = s.get<uint32_t>(NP * n);
    const size_t m = n / DPL;
    // Wide mode: limb sums packed (pieces of min(T / 2, 8) limbs,
    // NttLeaf::pack_log; high words stay below 2^57 because every coefficient
    // is below 2^25 (2^32 - 1)^2 < 2^89).  m is a multiple of the piece.
    L.one.pack_log = sum_pack_log(mode, plan_lt(L.logn, mode == MODE_WIDE ? L.variant : 0));
    L.one.slo = s.get<uint64_t>(m);
    L.one.shi = s.get<uint64_t>(m >> L.one.pack_log);
    L.one.sum_len = m;
    if ((l1 && !L.one.spec) || !L.one.prod || !L.one.slo || !L.one.shi) return set_error(PG_ENOMEM, "NTT scratch");
    HIP_TRY(launch_ntt(L, C.stream, C.hook()));
    return launch_carry(L.one.slo, L.one.shi, m, rp, an + bn, s, 1, L.one.pack_log);
}

}  // namespace pg

namespace pg {

// mul_fft_bits (src/mul_fft.cpp:340-384) on device pointers.
int fft_mul_device(uint64_t *rp, const uint64_t *ap, size_t an, uint64_t atop, const uint64_t *bp, size_t bn,
                   uint64_t btop, double *max_round) {
    Context &C = ctx();
    if (!C.stream) return set_error(PG_ENOINIT, "pg_init not called");
    if (!an || !bn || !atop || !btop) return set_error(PG_EINVAL, "operands must be trimmed and nonzero");
    const int bits = std::max(an, bn) <= 1024 ? 16 : 8;   // mul_fft.cpp:390-393
    const int dpl = 64 / bits;
    auto count = [&](size_t xn, uint64_t top) {            // fft_digit_count, mul_fft.cpp:222-232
        size_t d = (xn - 1) * dpl;
        do { ++d; top >>= bits; } while (top);
        return d;
    };
    const size_t da = count(an, atop), db = count(bn, btop);
    int logn = 0;
    while (((size_t)1 << logn) < da + db) ++logn;
    if (logn < FFT_MIN_LOG) logn = FFT_MIN_LOG;
    if (logn > FFT_MAX_LOG) return set_error(PG_ERANGE, "FFT limit 2^24 complex points");
    const size_t n = (size_t)1 << logn;
    const int l1 = fft_l1(logn), l2 = fft_l2(logn), lt = fft_lt(logn);
    const int seg = std::min(1 << lt, dpl);     // digits per limb sum
    const int parts = dpl / seg;
    Scratch s;
    FftLaunch L{};
    L.one.a = ap; L.one.b = bp; L.one.an = an; L.one.bn = bn;
    L.io_bytes = (double)(an + bn) * 16.0;
    L.logn = logn; L.bits = bits;
    L.one.maxerr = C.maxerr;
    { int st = fill_fft_tables(L); if (st) return st; }
    L.one.spec = s.get<double2>(n);
    const size_t nsum = n / seg;
    L.one.slo = s.get<uint64_t>(nsum);
    L.one.shi = s.get<uint64_t>(nsum);
    L.one.sum_len = nsum;
    if (!L.one.spec || !L.one.slo || !L.one.shi) return set_error(PG_ENOMEM, "FFT scratch");
    HIP_TRY(launch_fft(L, C.stream, C.hook()));
    int st = launch_carry(L.one.slo, L.one.shi, n / dpl, rp, an + bn, s, parts);
    if (st) return st;
    if (max_round) {
        uint64_t bitsv = 0;
        HIP_TRY(hipMemcpyAsync(&bitsv, C.maxerr, 8, hipMemcpyDeviceToHost, C.stream));
        HIP_TRY(hipStreamSynchronize(C.stream));
        double e;
        std::memcpy(&e, &bitsv, 8);
        *max_round = e;
    }
    return PG_OK;
}

// ---- plan autotuner -------------------------------------------------------------------

int autotune_ntt(int lo, int hi, int reps, double *ms, int *chosen) {
    Context &C = ctx();
    if (!C.stream) return set_error(PG_ENOINIT, "pg_init not called");
    if (lo > hi || reps < 1) return set_error(PG_EINVAL, "autotune: bad range");
    int st = PG_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    for (int logn = lo; logn <= hi && !st; ++logn) {
        const int row = logn - lo;
        for (int v = 0; v < NTT_VARIANTS; ++v) ms[row * NTT_VARIANTS + v] = -1.0;
        chosen[row] = 0;
        if (logn < NTT_VARIANT_MIN_LOG || logn > NTT_MAX_LOG) continue;
        // an = bn = 2^(logn-2) full limbs: 2^logn base-2^32 digits in all
        const size_t an = (size_t)1 << (logn - 2), bn = an;
        std::vector<uint64_t> h(an + bn);
        uint64_t x = 0x9E3779B97F4A7C15ull ^ (uint64_t)logn;
        for (auto &w : h) {
            x ^= x << 13; x ^= x >> 7; x ^= x << 17;
            w = x | (1ull << 63);
        }
        Scratch s;
        uint64_t *a = s.get<uint64_t>(an + bn), *r0 = s.get<uint64_t>(an + bn), *r = s.get<uint64_t>(an + bn);
        if (!a || !r0 || !r) { st = set_error(PG_ENOMEM, "autotune buffers"); break; }
        HIP_TRY(hipMemcpyAsync(a, h.data(), (an + bn) * 8, hipMemcpyHostToDevice, C.stream));
        double best = 0.0;
        const int saved = C.ntt_variant_tab[logn];
        std::vector<uint64_t> ref, got;
        for (int v = 0; v < NTT_VARIANTS && !st; ++v) {
            if (!plan_distinct(logn, v)) continue;
            C.ntt_variant_tab[logn] = v;
            uint64_t *out = v ? r : r0;
            st = ntt_mul_device(out, a, an, ~0ull, a + an, bn, ~0ull, MODE_WIDE);   // warm-up, tables
            if (st == PG_ERANGE) {   // beyond mul_ntt's caps: the length is never used
                st = PG_OK;
                break;
            }
            if (!st) st = C.sync();
            if (st) break;
            // The GPU clocks ramp over ~30 ms of load: warm up for at least
            // that long, then time at least 10 ms (and reps products).
            float t = 0.f;
            HIP_TRY(hipEventRecord(e0, C.stream));
            st = ntt_mul_device(out, a, an, ~0ull, a + an, bn, ~0ull, MODE_WIDE);
            HIP_TRY(hipEventRecord(e1, C.stream));
            if (!st) st = C.sync();
            if (st) break;
            HIP_TRY(hipEventElapsedTime(&t, e0, e1));
            const double one = std::max(0.001, (double)t);
            const int warm = (int)std::min(2000.0, std::ceil(30.0 / one));
            const int timed = std::max(reps, (int)std::min(2000.0, std::ceil(10.0 / one)));
            for (int k = 0; k < warm && !st; ++k) st = ntt_mul_device(out, a, an, ~0ull, a + an, bn, ~0ull, MODE_WIDE);
            HIP_TRY(hipEventRecord(e0, C.stream));
            for (int k = 0; k < timed && !st; ++k) st = ntt_mul_device(out, a, an, ~0ull, a + an, bn, ~0ull, MODE_WIDE);
            HIP_TRY(hipEventRecord(e1, C.stream));
            if (!st) st = C.sync();
            if (st) break;
            HIP_TRY(hipEventElapsedTime(&t, e0, e1));
            t /= (float)timed;
            ms[row * NTT_VARIANTS + v] = t;
            if (v) {   // every variant must give the default plan's product
                got.resize(an + bn);
                HIP_TRY(hipMemcpyAsync(got.data(), r, (an + bn) * 8, hipMemcpyDeviceToHost, C.stream));
                if ((st = C.sync())) break;
                if (got != ref) { st = set_error(PG_EINTERNAL, "autotune: plan variant changed the product"); break; }
            } else {
                ref.resize(an + bn);
                HIP_TRY(hipMemcpyAsync(ref.data(), r0, (an + bn) * 8, hipMemcpyDeviceToHost, C.stream));
                if ((st = C.sync())) break;
            }
            if (v == 0 || t < best) {
                best = t;
                chosen[row] = v;
            }
        }
        C.ntt_variant_tab[logn] = st ? saved : chosen[row];
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return st;
}

}  // namespace pg
