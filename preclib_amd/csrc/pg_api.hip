// pg_api.hip -- extern "C" entry points of libpgpu.so (declared in pgpu.h).
#include <algorithm>
#include <cstring>
#include "pg_batch.h"
#include "pg_dint.h"
#include "pg_pi.h"
#include "pg_runtime.h"
#include "pgpu.h"

using namespace pg;

namespace pg {
const char *last_error();
int mul_device(uint64_t *rp, const uint64_t *ap, size_t an, uint64_t atop, const uint64_t *bp,
               size_t bn, uint64_t btop, int path, double *fft_max_round);
}

struct pg_buf {
    uint64_t *d;
    size_t cap;
};

#define LOCK std::lock_guard<std::recursive_mutex> lk_(ctx().mu)
#define NEED_INIT \
    if (!ctx().stream) return set_error(PG_ENOINIT, "pg_init not called")

extern "C" {

int pg_init(int device, size_t hbm_budget) try {
    return ctx().init(device, hbm_budget);
} catch (...) {
    return pg::exception_status();
}

void pg_shutdown(void) { ctx().shutdown(); }

const char *pg_last_error(void) { return pg::last_error(); }

int pg_sync(void) try {
    LOCK;
    NEED_INIT;
    return ctx().sync();
} catch (...) {
    return pg::exception_status();
}

int pg_buf_alloc(pg_buf **h, size_t limbs) try {
    LOCK;
    NEED_INIT;
    if (!h || !limbs) return set_error(PG_EINVAL, "pg_buf_alloc: bad argument");
    void *p = ctx().pool.get(limbs * sizeof(uint64_t));
    if (!p) return set_error(PG_ENOMEM, "pg_buf_alloc: device budget exhausted");
    *h = new pg_buf{(uint64_t *)p, limbs};
    return PG_OK;
} catch (...) {
    return pg::exception_status();
}

int pg_buf_free(pg_buf *h) try {
    LOCK;
    if (!h) return PG_OK;
    ctx().pool.put(h->d);
    delete h;
    return PG_OK;
} catch (...) {
    return pg::exception_status();
}

size_t pg_buf_capacity(const pg_buf *h) { return h ? h->cap : 0; }
uint64_t *pg_buf_ptr(pg_buf *h) { return h ? h->d : nullptr; }

int pg_buf_upload(pg_buf *h, const uint64_t *src, size_t limbs) try {
    LOCK;
    NEED_INIT;
    if (!h || (!src && limbs) || limbs > h->cap) return set_error(PG_EINVAL, "pg_buf_upload: bad argument");
    if (hipMemcpyAsync(h->d, src, limbs * 8, hipMemcpyHostToDevice, ctx().stream) != hipSuccess ||
        ctx().wait() != hipSuccess)
        return set_error(PG_EDEVICE, "pg_buf_upload: copy failed");
    return PG_OK;
} catch (...) {
    return pg::exception_status();
}

int pg_buf_download(const pg_buf *h, uint64_t *dst, size_t limbs) try {
    LOCK;
    NEED_INIT;
    if (!h || (!dst && limbs) || limbs > h->cap) return set_error(PG_EINVAL, "pg_buf_download: bad argument");
    if (hipMemcpyAsync(dst, h->d, limbs * 8, hipMemcpyDeviceToHost, ctx().stream) != hipSuccess ||
        ctx().wait() != hipSuccess)
        return set_error(PG_EDEVICE, "pg_buf_download: copy failed");
    return PG_OK;
} catch (...) {
    return pg::exception_status();
}

static int check_operands(size_t an, size_t bn, size_t r_cap) {
    if (!an || !bn) return set_error(PG_EINVAL, "operands must have at least one limb");
    if (r_cap < an + bn) return set_error(PG_EINVAL, "r_cap < an + bn");
    return PG_OK;
}

int pg_mul(uint64_t *r, size_t r_cap, size_t *r_len, const uint64_t *a, size_t an, const uint64_t *b,
           size_t bn, int path, double *fft_max_round) try {
    LOCK;
    NEED_INIT;
    if (!an || !bn) {   // a zero operand gives the empty product (mul_basic.cpp:239-242)
        if (r_len) *r_len = 0;
        return PG_OK;
    }
    if (!r || !a || !b) return set_error(PG_EINVAL, "pg_mul: null pointer");
    int st = check_operands(an, bn, r_cap);
    if (st) return st;
    if (!a[an - 1] || !b[bn - 1]) return set_error(PG_EINVAL, "pg_mul: operands must be trimmed");
    Context &C = ctx();
    // a square (same host operand) is uploaded once, so the device sees
    // precn_sqr's aliasing (mul_basic.cpp:244-246)
    const bool same = a == b && an == bn;
    uint64_t *da = (uint64_t *)C.pool.get(an * 8);
    uint64_t *db = same ? da : (uint64_t *)C.pool.get(bn * 8);
    uint64_t *dr = (uint64_t *)C.pool.get((an + bn) * 8);
    auto release = [&] {
        C.pool.put(da);
        if (!same) C.pool.put(db);
        C.pool.put(dr);
    };
    if (!da || !db || !dr) {
        release();
        return set_error(PG_ENOMEM, "pg_mul: device budget exhausted");
    }
    st = PG_OK;
    if (hipMemcpyAsync(da, a, an * 8, hipMemcpyHostToDevice, C.stream) != hipSuccess ||
        (!same && hipMemcpyAsync(db, b, bn * 8, hipMemcpyHostToDevice, C.stream) != hipSuccess))
        st = set_error(PG_EDEVICE, "pg_mul: upload failed");
    if (!st) st = mul_device(dr, da, an, a[an - 1], db, bn, b[bn - 1], path, fft_max_round);
    if (!st && hipMemcpyAsync(r, dr, (an + bn) * 8, hipMemcpyDeviceToHost, C.stream) != hipSuccess)
        st = set_error(PG_EDEVICE, "pg_mul: download failed");
    if (C.wait() != hipSuccess && !st) st = set_error(PG_EDEVICE, "pg_mul: device error");
    release();
    if (st) return st;
    size_t len = an + bn;
    while (len && r[len - 1] == 0) --len;
    if (r_len) *r_len = len;
    return PG_OK;
} catch (...) {
    return pg::exception_status();
}

int pg_mul_dev(pg_buf *r, size_t *r_len, const pg_buf *a, size_t an, const pg_buf *b, size_t bn, int path,
               double *fft_max_round) try {
    LOCK;
    NEED_INIT;
    if (!r || !a || !b) return set_error(PG_EINVAL, "pg_mul_dev: null handle");
    if (an > a->cap || bn > b->cap) return set_error(PG_EINVAL, "pg_mul_dev: length beyond buffer");
    int st = check_operands(an, bn, r->cap);
    if (st) return st;
    Context &C = ctx();
    // The top limbs stay on the device: plan with full top limbs (~0).  The
    // transform is then at least as long as the exact digit count needs, so
    // the product is the same; no host round trip per product.
    st = mul_device(r->d, a->d, an, ~0ull, b->d, bn, ~0ull, path, fft_max_round);
    if (st) return st;
    if (r_len) {
        uint64_t top = 0;
        if (hipMemcpyAsync(&top, r->d + an + bn - 1, 8, hipMemcpyDeviceToHost, C.stream) != hipSuccess ||
            C.wait() != hipSuccess)
            return set_error(PG_EDEVICE, "pg_mul_dev: device error");
        *r_len = top ? an + bn : an + bn - 1;
    }
    return PG_OK;
} catch (...) {
    return pg::exception_status();
}

struct pg_batch {
    pg::Batch b;
};

int pg_batch_plan(pg_batch **out, const pg_job *jobs, size_t n) try {
    LOCK;
    NEED_INIT;
    if (!out || (!jobs && n)) return set_error(PG_EINVAL, "pg_batch_plan: bad argument");
    pg_batch *h = new pg_batch;
    const int st = h->b.build(jobs, n);
    if (st) {
        delete h;
        return st;
    }
    *out = h;
    return PG_OK;
} catch (...) {
    return pg::exception_status();
}

int pg_batch_run(pg_batch *h) try {
    LOCK;
    NEED_INIT;
    if (!h) return set_error(PG_EINVAL, "pg_batch_run: null handle");
    return h->b.run();
} catch (...) {
    return pg::exception_status();
}

int pg_batch_errors(pg_batch *h, double *err, size_t n) try {
    LOCK;
    NEED_INIT;
    if (!h || (!err && n)) return set_error(PG_EINVAL, "pg_batch_errors: bad argument");
    std::vector<double> e;
    const int st = h->b.read_errors(e);
    if (st) return st;
    for (size_t i = 0; i < n && i < e.size(); ++i) err[i] = e[i];
    return PG_OK;
} catch (...) {
    return pg::exception_status();
}

int pg_batch_free(pg_batch *h) try {
    LOCK;
    if (h) {
        if (ctx().stream) ctx().wait();
        delete h;
    }
    return PG_OK;
} catch (...) {
    return pg::exception_status();
}

int pg_mul_batch(pg_job *jobs, size_t n) try {
    LOCK;
    NEED_INIT;
    if (!jobs && n) return set_error(PG_EINVAL, "pg_mul_batch: null jobs");
    Context &C = ctx();
    size_t total = 0;
    for (size_t i = 0; i < n; ++i) {
        jobs[i].status = PG_OK;
        jobs[i].r_len = 0;
        jobs[i].fft_max_round = 0.0;
        if (!jobs[i].a || !jobs[i].b || !jobs[i].r || !jobs[i].an || !jobs[i].bn ||
            jobs[i].r_cap < jobs[i].an + jobs[i].bn)
            return set_error(PG_EINVAL, "pg_mul_batch: bad job " + std::to_string(i));
        total += 2 * (jobs[i].an + jobs[i].bn);
    }
    uint64_t *dev = (uint64_t *)C.pool.get(total * 8 + 8);
    if (!dev) return set_error(PG_ENOMEM, "pg_mul_batch: device budget exhausted");
    std::vector<pg_job> dj(jobs, jobs + n);
    size_t off = 0;
    int st = PG_OK;
    for (size_t i = 0; i < n && !st; ++i) {
        uint64_t *da = dev + off, *db = da + jobs[i].an, *dr = db + jobs[i].bn;
        if (hipMemcpyAsync(da, jobs[i].a, jobs[i].an * 8, hipMemcpyHostToDevice, C.stream) != hipSuccess ||
            hipMemcpyAsync(db, jobs[i].b, jobs[i].bn * 8, hipMemcpyHostToDevice, C.stream) != hipSuccess)
            st = set_error(PG_EDEVICE, "pg_mul_batch: upload failed");
        dj[i].a = da; dj[i].b = db; dj[i].r = dr;
        dj[i].r_cap = jobs[i].an + jobs[i].bn;
        off += 2 * (jobs[i].an + jobs[i].bn);
    }
    std::vector<double> errs;
    if (!st) {
        pg::Batch b;
        b.use_graph = false;
        st = b.build(dj.data(), n);
        if (!st) st = b.run();
        if (!st) st = b.read_errors(errs);
        for (size_t i = 0; i < n && !st; ++i)
            if (hipMemcpyAsync(jobs[i].r, dj[i].r, (jobs[i].an + jobs[i].bn) * 8, hipMemcpyDeviceToHost,
                               C.stream) != hipSuccess)
                st = set_error(PG_EDEVICE, "pg_mul_batch: download failed");
        if (C.wait() != hipSuccess && !st) st = set_error(PG_EDEVICE, "pg_mul_batch: device error");
    }
    C.pool.put(dev);
    for (size_t i = 0; i < n; ++i) {
        jobs[i].status = st;
        if (st) continue;
        size_t len = jobs[i].an + jobs[i].bn;
        while (len && jobs[i].r[len - 1] == 0) --len;
        jobs[i].r_len = len;
        jobs[i].fft_max_round = i < errs.size() ? errs[i] : 0.0;
    }
    return st;
} catch (...) {
    return pg::exception_status();
}

// ---- Newton division (div_newton.cpp) -------------------------------------------

static size_t host_trim(const uint64_t *a, size_t n) {
    while (n && !a[n - 1]) --n;
    return n;
}

static int finish_out(const DBuf &src, size_t len, uint64_t *dst, size_t cap, size_t *out_len, bool device) {
    if (out_len) *out_len = len;
    if (!dst) return PG_OK;
    if (len > cap) return set_error(PG_EINVAL, "output capacity too small");
    Context &C = ctx();
    if (len)
        HIP_TRY(hipMemcpyAsync(dst, src.d, len * 8, device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                               C.stream));
    return C.sync();
}

int pg_div_dev(pg_buf *q, size_t q_cap, size_t *q_len, pg_buf *r, size_t r_cap, size_t *r_len, const pg_buf *a,
               size_t an, const pg_buf *b, size_t bn) try {
    LOCK;
    NEED_INIT;
    if (!a || !b || an > a->cap || bn > b->cap || (q && q_cap > q->cap) || (r && r_cap > r->cap))
        return set_error(PG_EINVAL, "pg_div_dev: bad argument");
    size_t at = 0, bt = 0;
    int st = top_len(a->d, an, &at);
    if (!st) st = top_len(b->d, bn, &bt);
    if (st) return st;
    DBuf dq, dr;
    size_t ql = 0, rl = 0;
    if ((st = div_mulinv(dq, ql, r ? &dr : nullptr, &rl, a->d, at, b->d, bt))) return st;
    if ((st = finish_out(dq, ql, q ? q->d : nullptr, q_cap, q_len, true))) return st;
    return finish_out(dr, rl, r ? r->d : nullptr, r_cap, r_len, true);
} catch (...) {
    return pg::exception_status();
}

int pg_div(uint64_t *q, size_t q_cap, size_t *q_len, uint64_t *r, size_t r_cap, size_t *r_len, const uint64_t *a,
           size_t an, const uint64_t *b, size_t bn) try {
    LOCK;
    NEED_INIT;
    if ((an && !a) || (bn && !b)) return set_error(PG_EINVAL, "pg_div: bad argument");
    an = host_trim(a, an);
    bn = host_trim(b, bn);
    Context &C = ctx();
    DBuf da(an), db(bn), dq, dr;
    if (!da.d || !db.d) return set_error(PG_ENOMEM, "pg_div: operands");
    if (an) HIP_TRY(hipMemcpyAsync(da.d, a, an * 8, hipMemcpyHostToDevice, C.stream));
    if (bn) HIP_TRY(hipMemcpyAsync(db.d, b, bn * 8, hipMemcpyHostToDevice, C.stream));
    size_t ql = 0, rl = 0;
    int st = div_mulinv(dq, ql, r ? &dr : nullptr, &rl, da.d, an, db.d, bn);
    if (st) return st;
    if ((st = finish_out(dq, ql, q, q_cap, q_len, false))) return st;
    return finish_out(dr, rl, r, r_cap, r_len, false);
} catch (...) {
    return pg::exception_status();
}

int pg_recip_dev(pg_buf *x, size_t x_cap, size_t *x_len, const pg_buf *b, size_t bn, size_t bits) try {
    LOCK;
    NEED_INIT;
    if (!x || !b || bn > b->cap || x_cap > x->cap) return set_error(PG_EINVAL, "pg_recip_dev: bad argument");
    size_t bt = 0, xl = 0;
    int st = top_len(b->d, bn, &bt);
    if (st) return st;
    DBuf dx;
    if ((st = recip_exact(dx, xl, b->d, bt, bits))) return st;
    return finish_out(dx, xl, x->d, x_cap, x_len, true);
} catch (...) {
    return pg::exception_status();
}

int pg_recip(uint64_t *x, size_t x_cap, size_t *x_len, const uint64_t *b, size_t bn, size_t bits) try {
    LOCK;
    NEED_INIT;
    if (bn && !b) return set_error(PG_EINVAL, "pg_recip: bad argument");
    bn = host_trim(b, bn);
    Context &C = ctx();
    DBuf db(bn), dx;
    if (!db.d) return set_error(PG_ENOMEM, "pg_recip: operand");
    if (bn) HIP_TRY(hipMemcpyAsync(db.d, b, bn * 8, hipMemcpyHostToDevice, C.stream));
    size_t xl = 0;
    int st = recip_exact(dx, xl, db.d, bn, bits);
    if (st) return st;
    return finish_out(dx, xl, x, x_cap, x_len, false);
} catch (...) {
    return pg::exception_status();
}

// ---- square root, decimal conversion, binary splitting, pi -------------------------

int pg_sqrt(uint64_t *x, size_t x_cap, size_t *x_len, const uint64_t *a, size_t an) try {
    LOCK;
    NEED_INIT;
    if (an && !a) return set_error(PG_EINVAL, "pg_sqrt: bad argument");
    an = host_trim(a, an);
    Context &C = ctx();
    DBuf da(an), dx;
    if (!da.d) return set_error(PG_ENOMEM, "pg_sqrt: operand");
    if (an) HIP_TRY(hipMemcpyAsync(da.d, a, an * 8, hipMemcpyHostToDevice, C.stream));
    size_t xl = 0;
    int st = isqrt(dx, xl, da.d, an);
    if (st) return st;
    return finish_out(dx, xl, x, x_cap, x_len, false);
} catch (...) {
    return pg::exception_status();
}

static int put_string(const std::string &s, char *out, size_t cap, size_t *len) {
    if (len) *len = s.size();
    if (!out) return PG_OK;
    if (cap < s.size()) return set_error(PG_EINVAL, "output capacity too small");
    std::memcpy(out, s.data(), s.size());
    return PG_OK;
}

int pg_to_decimal(char *out, size_t cap, size_t *len, const uint64_t *a, size_t an) try {
    LOCK;
    NEED_INIT;
    if (an && !a) return set_error(PG_EINVAL, "pg_to_decimal: bad argument");
    an = host_trim(a, an);
    Context &C = ctx();
    DBuf da(an);
    if (!da.d) return set_error(PG_ENOMEM, "pg_to_decimal: operand");
    if (an) HIP_TRY(hipMemcpyAsync(da.d, a, an * 8, hipMemcpyHostToDevice, C.stream));
    std::string s;
    size_t n = 0;
    int st = to_decimal(da.d, an, s, 0, out, cap, &n);   // with a buffer, straight into it
    if (len) *len = n;
    if (st || out) return st;
    return put_string(s, out, cap, len);
} catch (...) {
    return pg::exception_status();
}

int pg_from_decimal(uint64_t *out, size_t cap, size_t *len, const char *s, size_t n) try {
    LOCK;
    NEED_INIT;
    if ((n && !s) || !len) return set_error(PG_EINVAL, "pg_from_decimal: bad argument");
    // every character outside '0'..'9' is skipped (init_basic.cpp:69-72)
    std::string digits;
    digits.reserve(n);
    for (size_t i = 0; i < n; ++i)
        if (s[i] >= '0' && s[i] <= '9') digits.push_back(s[i]);
    std::vector<uint64_t> v;
    const int st = from_decimal(digits.data(), digits.size(), v);
    if (st) return st;
    *len = v.size();
    if (v.size() > cap || (!out && !v.empty())) return set_error(PG_EINVAL, "pg_from_decimal: output capacity too small");
    if (!v.empty()) std::memcpy(out, v.data(), v.size() * 8);
    return PG_OK;
} catch (...) {
    return pg::exception_status();
}

int pg_chud_split(uint64_t *p, size_t p_cap, size_t *p_len, uint64_t *q, size_t q_cap, size_t *q_len, uint64_t *t,
                  size_t t_cap, size_t *t_len, size_t nterms) try {
    LOCK;
    NEED_INIT;
    std::vector<DBuf> keep;
    DBuf P, Q, T;
    size_t lp = 0, lq = 0, lt = 0;
    int st = chud_split(nterms, P, lp, Q, lq, T, lt, p != nullptr, keep);
    if (st) return st;
    if ((st = finish_out(P, lp, p, p_cap, p_len, false))) return st;
    if ((st = finish_out(Q, lq, q, q_cap, q_len, false))) return st;
    return finish_out(T, lt, t, t_cap, t_len, false);
} catch (...) {
    return pg::exception_status();
}

int pg_pi(char *out, size_t cap, size_t *len, size_t digits, double *phase_ms) try {
    LOCK;
    NEED_INIT;
    std::string s;
    size_t n = 0;
    // with a buffer, the digits are formatted straight into it
    int st = pi_digits(digits, s, phase_ms, out, cap, &n);
    if (len) *len = n;
    if (st || out) return st;
    return put_string(s, out, cap, len);
} catch (...) {
    return pg::exception_status();
}

// ---- plan autotuner ------------------------------------------------------------------

int pg_ntt_autotune(int logn_lo, int logn_hi, int reps, double *ms, int *chosen) try {
    LOCK;
    NEED_INIT;
    if (!ms || !chosen) return set_error(PG_EINVAL, "pg_ntt_autotune: null");
    return autotune_ntt(logn_lo, logn_hi, reps, ms, chosen);
} catch (...) {
    return pg::exception_status();
}

int pg_ntt_plan_get(int *variants, int n) try {
    LOCK;
    NEED_INIT;
    for (int i = 0; i < n; ++i) variants[i] = i <= NTT_MAX_LOG ? ctx().ntt_variant_tab[i] : 0;
    return PG_OK;
} catch (...) {
    return pg::exception_status();
}

int pg_ntt_plan_set(const int *variants, int n) try {
    LOCK;
    NEED_INIT;
    for (int i = 0; i < n && i <= NTT_MAX_LOG; ++i)
        if (variants[i] < 0 || variants[i] >= NTT_VARIANTS) return set_error(PG_EINVAL, "pg_ntt_plan_set: variant");
    for (int i = 0; i < n && i <= NTT_MAX_LOG; ++i) ctx().ntt_variant_tab[i] = i >= NTT_VARIANT_MIN_LOG ? variants[i] : 0;
    return PG_OK;
} catch (...) {
    return pg::exception_status();
}

int pg_ntt_batch_plan_get(int *variants, int n) try {
    LOCK;
    NEED_INIT;
    for (int i = 0; i < n; ++i) variants[i] = i <= NTT_MAX_LOG ? ctx().ntt_batch_tab[i] : -1;
    return PG_OK;
} catch (...) {
    return pg::exception_status();
}

int pg_ntt_batch_plan_set(const int *variants, int n) try {
    LOCK;
    NEED_INIT;
    for (int i = 0; i < n && i <= NTT_MAX_LOG; ++i)
        if (variants[i] < -1 || variants[i] >= NTT_VARIANTS) return set_error(PG_EINVAL, "pg_ntt_batch_plan_set: variant");
    for (int i = 0; i < n && i <= NTT_MAX_LOG; ++i) ctx().ntt_batch_tab[i] = i >= NTT_VARIANT_MIN_LOG ? variants[i] : -1;
    return PG_OK;
} catch (...) {
    return pg::exception_status();
}

int pg_timing_enable(int on) try {
    LOCK;
    ctx().timing = on != 0;
    return PG_OK;
} catch (...) {
    return pg::exception_status();
}

int pg_timing_reset(void) try {
    LOCK;
    if (ctx().stream) ctx().wait();
    ctx().collect_timing();
    ctx().stats.clear();
    return PG_OK;
} catch (...) {
    return pg::exception_status();
}

int pg_debug_perturb(int what, int64_t amount) try {
    LOCK;
    NEED_INIT;
    return pg::debug_perturb(what, amount);
} catch (...) {
    return pg::exception_status();
}

int pg_host_staging(size_t *held_bytes, size_t *spare_bytes) try {
    LOCK;
    size_t sp = 0;
    for (const auto &h : ctx().spare) sp += h.capacity();
    if (held_bytes) *held_bytes = ctx().held_bytes;
    if (spare_bytes) *spare_bytes = sp;
    return PG_OK;
} catch (...) {
    return pg::exception_status();
}

int pg_timing_read(pg_stage_time *out, int cap) try {
    LOCK;
    if (ctx().stream) ctx().wait();
    ctx().collect_timing();
    int i = 0;
    for (auto &kv : ctx().stats) {
        if (out && i < cap) {
            out[i].name = kv.first.c_str();
            out[i].launches = kv.second.launches;
            out[i].total_ms = kv.second.ms;
            out[i].bytes = kv.second.bytes;
        }
        ++i;
    }
    return i;
} catch (...) {
    return pg::exception_status();
}

}  // extern "C"

namespace pg {
// Size-based routing above 192 limbs, following mul_into (src/mul_basic.cpp:
// 259-269) with the NTT for every transform-sized product (DISPATCH_* in
// pg_kernels.h), and mul_ntt's fallbacks (src/mul_ntt.cpp:497-506).
int mul_device(uint64_t *rp, const uint64_t *ap, size_t an, uint64_t atop, const uint64_t *bp, size_t bn,
               uint64_t btop, int path, double *fft_max_round) {
    if (fft_max_round) *fft_max_round = 0.0;
    const size_t lo = std::min(an, bn), hi = std::max(an, bn);
    const bool sqr = ap == bp && an == bn;
    // Below the transform threshold mul_ntt / mul_fft fall back to mul_basic
    // (mul_ntt.cpp:496, mul_fft.cpp:388); mul_into takes the schoolbook for a
    // side <= 24 limbs and the blocked multiply for hi > 2 lo.
    const bool small = hi <= DISPATCH_SMALL_HI || (path == PG_PATH_AUTO && lo <= DISPATCH_SMALL_SIDE);
    if (small || path == PG_PATH_AUTO) {
        if (small || (!sqr && hi > 2 * lo)) {
            // small leaf / mul_unbalanced_into (mul_basic.cpp:207-236) as a batch of one
            pg_job j{};
            j.a = ap; j.an = an; j.b = bp; j.bn = bn; j.r = rp; j.r_cap = an + bn; j.path = PG_PATH_AUTO;
            Batch b;
            b.use_graph = false;
            int st = b.build(&j, 1);
            if (!st) st = b.run();
            if (!st && fft_max_round) {
                std::vector<double> e;
                st = b.read_errors(e);
                if (!st) *fft_max_round = e[0];
            }
            // stream-ordered: the batch's arena returns to the pool, whose
            // blocks are only reused by later work on the same stream
            return st;
        }
        path = PG_PATH_NTT;
    }
    if (path == PG_PATH_NTT || path == PG_PATH_NTT_REF) {
        int st = ntt_mul_device(rp, ap, an, atop, bp, bn, btop, path == PG_PATH_NTT_REF ? MODE_REF : MODE_WIDE);
        if (st != PG_ERANGE) return st;
        path = PG_PATH_FFT;   // beyond the NTT caps the reference multiplies with mul_fft
    }
    const int st = fft_mul_device(rp, ap, an, atop, bp, bn, btop, fft_max_round);
    // a square requested on the FFT past its 2^24-point table: the same exact
    // product from the NTT
    if (st == PG_ERANGE && sqr) return ntt_mul_device(rp, ap, an, atop, bp, bn, btop, MODE_WIDE);
    return st;
}
}  // namespace pg
