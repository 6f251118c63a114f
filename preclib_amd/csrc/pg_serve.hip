// pg_serve.hip -- multiply / Newton-division server with continuous batching.
//
// One worker thread owns the device.  Requests wait in a queue and are
// admitted before every round while their HBM footprint fits the budget and
// a slot is free.  A round collects the next step of every admitted request:
// one linear launch before the products, one batched product launch
// (pg_batch.hip: every product of the round, whatever its request, shape or
// path) and one linear launch after them.  A multiply finishes in one
// round; a division advances through the steps of div_newton.cpp (two
// rounds per precision doubling of the reciprocal, div_newton.cpp:49-65,
// then a*x, b*q and the unit corrections of 114-125), so divisions and
// multiplies that are in flight together share every launch.  Operands,
// reciprocals and partial products stay in HBM until the request finishes;
// finished requests free their memory and slot for the next arrivals.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <new>
#include <thread>

#include "pg_batch.h"
#include "pg_dint.h"
#include "pgpu.h"

namespace pg {

static double now_s() {
    using clk = std::chrono::steady_clock;
    static const clk::time_point t0 = clk::now();
    return std::chrono::duration<double>(clk::now() - t0).count();
}

namespace {

enum Stage { S_MUL, S_IT_A, S_IT_B, S_FIN_AX, S_FIN_P, S_CORR, S_CORR_INC, S_DONE };

size_t host_len(const uint64_t *a, size_t n) {
    while (n && !a[n - 1]) --n;
    return n;
}

size_t host_bits(const uint64_t *a, size_t n) { return n ? (n - 1) * 64 + 64 - (size_t)__builtin_clzll(a[n - 1]) : 0; }

// Bytes of HBM a request may hold at once (operands, results, transform
// scratch of its largest product), used for admission.
size_t footprint(int op, size_t an, size_t bn) {
    const size_t n = an + bn + 1;
    return (op == PG_OP_MUL ? 112 : 256) * n * sizeof(uint64_t) / 8 + (1u << 20);
}

struct Req {
    pg_request *user = nullptr;
    int op = PG_OP_MUL, stage = S_DONE, status = PG_OK;
    size_t an = 0, bn = 0, fp = 0;
    DBuf A, B, R, flag;
    bool read_flag = false;
    uint64_t flag_host = 0;
    // division state (div_newton.cpp:36-129)
    size_t na = 0, nb = 0, scale = 0, target = 0, known = 0;
    size_t xl = 0, nk = 0, next = 0, s = 0, nbl = 0, el = 0, nxl = 0, ql = 0, rm = 0;
    const uint64_t *nbp = nullptr;
    DBuf X, Xn, NB, BX, E, XE, AX, Q, Q2, P, R2;
    uint64_t corr_sign = 0;
    int corr_steps = 0;
    // host results
    bool host_done = false;   // result written without the device (trivial cases)
};

}  // namespace

struct Server {
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<pg_request *> queue;
    std::vector<std::unique_ptr<Req>> active;
    size_t budget = 0, used = 0, max_active = 256;
    size_t outstanding = 0;   // submitted and not finished (guarded by mu)
    bool stop = false;
    pg_server_stats stats{};
    std::thread worker;

    int admit(pg_request *u, std::unique_ptr<Req> &out);
    void plan(Req &q, std::vector<CarryNode> &pre, std::vector<pg_job> &prod, std::vector<CarryNode> &post);
    int advance(Req &q);
    int finish(Req &q);
    void round();
    void run();
};

static void fail(Req &q, int st) {
    q.status = st;
    q.stage = S_DONE;
}

// Uploads the operands and sets up the first step.
int Server::admit(pg_request *u, std::unique_ptr<Req> &out) {
    out.reset(new Req);
    Req &q = *out;
    q.user = u;
    q.op = u->op;
    q.an = host_len(u->a, u->an);
    q.bn = host_len(u->b, u->bn);
    q.fp = footprint(q.op, q.an, q.bn);
    u->t_admit = now_s();
    Context &C = ctx();
    if (q.op == PG_OP_MUL) {
        if (!q.an || !q.bn) {   // mul_basic.cpp:239-242
            q.host_done = true;
            q.stage = S_DONE;
            return PG_OK;
        }
    } else {
        q.na = host_bits(u->a, q.an);
        q.nb = host_bits(u->b, q.bn);
        if (!q.an || !q.bn || q.na < q.nb) {   // div_newton.cpp:90-97
            q.host_done = true;
            q.stage = S_DONE;
            return PG_OK;
        }
    }
    if (!q.A.reset(q.an) || !q.B.reset(q.bn) || !q.flag.reset(2)) return set_error(PG_ENOMEM, "server: operands");
    HIP_TRY(hipMemcpyAsync(q.A.d, u->a, q.an * 8, hipMemcpyHostToDevice, C.stream));
    HIP_TRY(hipMemcpyAsync(q.B.d, u->b, q.bn * 8, hipMemcpyHostToDevice, C.stream));
    if (q.op == PG_OP_MUL) {
        q.stage = S_MUL;
        return PG_OK;
    }
    q.scale = q.na + 128;
    q.target = q.scale - q.nb;
    // reciprocal seed (div_newton.cpp:41-47) from the host copy of b
    q.known = std::min<size_t>(q.target, NEWTON_SEED_BITS);
    const size_t keep = std::min(q.nb, q.known + 128), drop = q.nb - keep, l0 = drop / 64;
    const int r = (int)(drop % 64);
    std::vector<uint64_t> bt(q.bn - l0);
    for (size_t i = 0; i < bt.size(); ++i) {
        const uint64_t w = u->b[l0 + i], w1 = l0 + i + 1 < q.bn ? u->b[l0 + i + 1] : 0;
        bt[i] = r ? (w >> r) | (w1 << (64 - r)) : w;
    }
    if (drop) {
        bt.push_back(0);
        for (auto &v : bt)
            if (++v) break;
    }
    while (!bt.empty() && !bt.back()) bt.pop_back();
    const std::vector<uint64_t> x0 = pow2_div_small(keep + q.known, bt);
    q.xl = limbs_for_bits(q.known + 2) + 1;
    if (!q.X.reset(q.xl)) return set_error(PG_ENOMEM, "server: reciprocal");
    int st = dset(q.X.d, q.xl, x0.data(), x0.size());
    if (st) return st;
    q.stage = q.known < q.target ? S_IT_A : S_FIN_AX;
    return PG_OK;
}

// The next step of a request: linear nodes before / after the products.
void Server::plan(Req &q, std::vector<CarryNode> &pre, std::vector<pg_job> &prod, std::vector<CarryNode> &post) {
    auto job = [&](const uint64_t *a, size_t an, const uint64_t *b, size_t bn, uint64_t *r) {
        pg_job j{};
        j.a = a; j.an = an; j.b = b; j.bn = bn; j.r = r; j.r_cap = an + bn; j.path = PG_PATH_AUTO;
        prod.push_back(j);
    };
    q.read_flag = false;
    switch (q.stage) {
    case S_MUL:
        if (!q.R.reset(q.an + q.bn)) return fail(q, set_error(PG_ENOMEM, "server: product"));
        job(q.A.d, q.an, q.B.d, q.bn, q.R.d);
        return;
    case S_IT_A: {   // next_b, bx = next_b * x, e = 2^(next+1) - (bx << s)
        q.nk = std::min(q.target, q.known * 2);
        const size_t nkeep = std::min(q.nb, q.nk + 128);
        q.next = nkeep + q.nk;
        q.s = q.nk - q.known;
        q.nbp = q.B.d;
        q.nbl = q.bn;
        if (nkeep < q.nb) {
            q.nbl = limbs_for_bits(nkeep + 1);
            if (!q.NB.reset(q.nbl)) return fail(q, set_error(PG_ENOMEM, "server: reciprocal"));
            CarryNode c = lin_node(q.NB.d, q.nbl);
            c.x = q.B.d; c.xn = q.bn; c.xsh = (int64_t)(q.nb - nkeep); c.cin = 1;
            pre.push_back(c);
            q.nbp = q.NB.d;
        }
        q.el = limbs_for_bits(q.next + 2);
        if (!q.BX.reset(q.nbl + q.xl) || !q.E.reset(q.el)) return fail(q, set_error(PG_ENOMEM, "server: reciprocal"));
        job(q.nbp, q.nbl, q.X.d, q.xl, q.BX.d);
        CarryNode e = lin_node(q.E.d, std::max(q.el, q.nbl + q.xl + limbs_for_bits(q.s) + 1));
        e.out_len = q.el;
        e.pow2 = (int64_t)(q.next + 1);
        e.y = q.BX.d; e.yn = q.nbl + q.xl; e.ysh = -(int64_t)q.s; e.yneg = 1; e.cin = 1; e.flag = q.flag.d;
        post.push_back(e);
        q.read_flag = true;
        return;
    }
    case S_IT_B: {   // x = x * e >> (next - s)
        q.nxl = limbs_for_bits(q.nk + 2) + 1;
        if (!q.XE.reset(q.xl + q.el) || !q.Xn.reset(q.nxl)) return fail(q, set_error(PG_ENOMEM, "server: reciprocal"));
        job(q.X.d, q.xl, q.E.d, q.el, q.XE.d);
        CarryNode c = lin_node(q.Xn.d, q.nxl);
        c.x = q.XE.d; c.xn = q.xl + q.el; c.xsh = (int64_t)(q.next - q.s);
        post.push_back(c);
        return;
    }
    case S_FIN_AX: {   // q = a * x >> scale
        q.ql = limbs_for_bits(q.na - q.nb + 2) + 1;
        if (!q.AX.reset(q.an + q.xl) || !q.Q.reset(q.ql)) return fail(q, set_error(PG_ENOMEM, "server: quotient"));
        job(q.A.d, q.an, q.X.d, q.xl, q.AX.d);
        CarryNode c = lin_node(q.Q.d, q.ql);
        c.x = q.AX.d; c.xn = q.an + q.xl; c.xsh = (int64_t)q.scale;
        post.push_back(c);
        return;
    }
    case S_FIN_P: {   // r = a - b q (two's complement, sign word)
        q.rm = std::max(q.an, q.bn + q.ql) + 1;
        if (!q.P.reset(q.bn + q.ql) || !q.R.reset(q.rm) || !q.R2.reset(q.rm) || !q.Q2.reset(q.ql))
            return fail(q, set_error(PG_ENOMEM, "server: remainder"));
        job(q.B.d, q.bn, q.Q.d, q.ql, q.P.d);
        CarryNode c = lin_node(q.R.d, q.rm);
        c.x = q.A.d; c.xn = q.an; c.y = q.P.d; c.yn = q.bn + q.ql; c.yneg = 1; c.cin = 1; c.flag = q.flag.d;
        post.push_back(c);
        q.read_flag = true;
        return;
    }
    case S_CORR:
        if (q.corr_sign == 0) {   // r < 0: q -= 1, r += b
            CarryNode c = lin_node(q.Q2.d, q.ql);
            c.x = q.Q.d; c.xn = q.ql; c.yneg = 1;
            CarryNode d = lin_node(q.R2.d, q.rm);
            d.x = q.R.d; d.xn = q.rm; d.y = q.B.d; d.yn = q.bn; d.flag = q.flag.d;
            post.push_back(c);
            post.push_back(d);
        } else {                  // r >= 0: test r - b
            CarryNode d = lin_node(q.R2.d, q.rm);
            d.x = q.R.d; d.xn = q.rm; d.y = q.B.d; d.yn = q.bn; d.yneg = 1; d.cin = 1; d.flag = q.flag.d;
            post.push_back(d);
        }
        q.read_flag = true;
        return;
    case S_CORR_INC: {
        CarryNode c = lin_node(q.Q2.d, q.ql);
        c.x = q.Q.d; c.xn = q.ql; c.cin = 1;
        post.push_back(c);
        return;
    }
    default:
        return;
    }
}

// After the round: move to the next step (flag_host holds the sign word).
int Server::advance(Req &q) {
    switch (q.stage) {
    case S_MUL:
        q.stage = S_DONE;
        break;
    case S_IT_A:
        if (q.flag_host != 1) return fail(q, set_error(PG_EINTERNAL, "server: reciprocal overshoot")), PG_OK;
        q.stage = S_IT_B;
        break;
    case S_IT_B:
        std::swap(q.X, q.Xn);
        q.xl = q.nxl;
        q.known = q.nk;
        q.BX.release(); q.E.release(); q.XE.release(); q.NB.release();
        q.stage = q.known < q.target ? S_IT_A : S_FIN_AX;
        break;
    case S_FIN_AX:
        q.X.release(); q.Xn.release();
        q.stage = S_FIN_P;
        break;
    case S_FIN_P:
        q.AX.release();
        q.corr_sign = q.flag_host;
        q.stage = S_CORR;
        break;
    case S_CORR:
        if (++q.corr_steps > 64) return fail(q, set_error(PG_EINTERNAL, "server: correction")), PG_OK;
        if (q.corr_sign == 0) {
            std::swap(q.Q, q.Q2);
            std::swap(q.R, q.R2);
            q.corr_sign = q.flag_host;
        } else if (q.flag_host == 0) {
            q.stage = S_DONE;
        } else {
            q.stage = S_CORR_INC;   // r2 = r - b >= 0: q += 1 next
        }
        break;
    case S_CORR_INC:
        std::swap(q.Q, q.Q2);
        std::swap(q.R, q.R2);
        q.corr_sign = 1;
        q.stage = S_CORR;
        break;
    default:
        break;
    }
    return PG_OK;
}

// Copies a finished request's results to its host buffers (after a sync).
int Server::finish(Req &q) {
    pg_request *u = q.user;
    u->r_len = u->r2_len = 0;
    if (q.status) return q.status;
    if (q.host_done) {
        if (q.op == PG_OP_DIV && u->r2 && q.bn && q.an) {   // a < b: r = a
            if (u->r2_cap < q.an) return set_error(PG_EINVAL, "server: remainder capacity");
            std::memcpy(u->r2, u->a, q.an * 8);
            u->r2_len = q.an;
        }
        return PG_OK;
    }
    Context &C = ctx();
    if (q.op == PG_OP_MUL) {
        const size_t n = std::min(q.an + q.bn, u->r_cap);
        HIP_TRY(hipMemcpyAsync(u->r, q.R.d, n * 8, hipMemcpyDeviceToHost, C.stream));
        return PG_OK;
    }
    HIP_TRY(hipMemcpyAsync(u->r, q.Q.d, std::min(q.ql, u->r_cap) * 8, hipMemcpyDeviceToHost, C.stream));
    if (u->r2) HIP_TRY(hipMemcpyAsync(u->r2, q.R.d, std::min(q.bn, u->r2_cap) * 8, hipMemcpyDeviceToHost, C.stream));
    return PG_OK;
}

void Server::round() {
    Context &C = ctx();
    std::lock_guard<std::recursive_mutex> lk(C.mu);
    std::vector<pg_job> prod;
    int st = PG_OK;
    try {
        std::vector<CarryNode> pre, post;
        for (auto &q : active)
            if (q->stage != S_DONE) plan(*q, pre, prod, post);
        st = lin_run(pre.data(), pre.size(), "serve");
        if (!st && !prod.empty()) {
            Batch b;
            b.use_graph = false;
            st = b.build(prod.data(), prod.size());
            if (!st) st = b.run();
        }
        if (!st) st = lin_run(post.data(), post.size(), "serve");
        for (auto &q : active)
            if (!st && q->read_flag && q->stage != S_DONE &&
                hipMemcpyAsync(&q->flag_host, q->flag.d, 8, hipMemcpyDeviceToHost, C.stream) != hipSuccess)
                st = set_error(PG_EDEVICE, "server: flag read");
    } catch (...) {
        // the round's requests fail; work already queued finishes first
        st = exception_status();
    }
    {
        const int sst = C.sync();
        if (!st) st = sst;
    }
    {
        // pg_server_read_stats copies the counters under mu
        std::lock_guard<std::mutex> g(mu);
        stats.rounds++;
        stats.products += prod.size();
    }
    for (auto &q : active) {
        if (st) fail(*q, st);
        else if (q->stage != S_DONE) advance(*q);
    }
    // finished requests: results to the host, memory and slots back
    std::vector<std::unique_ptr<Req>> keep, done;
    for (auto &q : active) (q->stage == S_DONE ? done : keep).push_back(std::move(q));
    active.swap(keep);
    if (done.empty()) return;
    std::vector<int> fst(done.size());
    for (size_t i = 0; i < done.size(); ++i) fst[i] = finish(*done[i]);
    const int sst = C.sync();
    const double t = now_s();
    {
        std::lock_guard<std::mutex> g(mu);
        for (size_t i = 0; i < done.size(); ++i) {
            Req &q = *done[i];
            pg_request *u = q.user;
            int s = q.status ? q.status : (fst[i] ? fst[i] : sst);
            if (!s) {
                const size_t rn = q.host_done ? 0 : (q.op == PG_OP_MUL ? std::min(q.an + q.bn, u->r_cap)
                                                                         : std::min(q.ql, u->r_cap));
                u->r_len = host_len(u->r, rn);
                if (q.op == PG_OP_DIV && u->r2 && !q.host_done) u->r2_len = host_len(u->r2, std::min(q.bn, u->r2_cap));
            }
            u->t_done = t;
            used -= q.fp;
            stats.completed++;
            outstanding--;
            u->status = s;   // last: waiters read it
        }
    }
    cv_done.notify_all();
}

void Server::run() {
    for (;;) {
        {
            std::unique_lock<std::mutex> g(mu);
            cv_work.wait(g, [&] { return stop || !queue.empty() || !active.empty(); });
            if (stop && queue.empty() && active.empty()) return;
            // admission by footprint; one request is always admitted when idle
            while (!queue.empty() && active.size() < max_active) {
                pg_request *u = queue.front();
                const size_t fp = footprint(u->op, host_len(u->a, u->an), host_len(u->b, u->bn));
                if (!active.empty() && used + fp > budget) break;
                queue.pop_front();
                used += fp;
                std::unique_ptr<Req> q;
                int st;
                {
                    std::lock_guard<std::recursive_mutex> lk(ctx().mu);
                    try {
                        st = admit(u, q);
                    } catch (...) {
                        st = exception_status();
                    }
                }
                if (!q) {   // admit failed before its record existed
                    q.reset(new (std::nothrow) Req);
                    if (!q) {   // no record at all: the request completes failed now
                        used -= fp;
                        stats.completed++;
                        outstanding--;
                        u->t_done = now_s();
                        u->status = st ? st : PG_ENOMEM;
                        cv_done.notify_all();
                        continue;
                    }
                    q->user = u;
                    q->op = u->op;
                }
                q->fp = fp;   // what admission added to `used`
                if (st) fail(*q, st);
                active.push_back(std::move(q));
                stats.max_active = std::max<uint64_t>(stats.max_active, active.size());
            }
        }
        round();
    }
}

}  // namespace pg

using namespace pg;

struct pg_server {
    Server s;
};

extern "C" {

double pg_server_clock(void) { return now_s(); }

int pg_server_start(pg_server **out, size_t hbm_budget, size_t max_active) try {
    if (!out) return set_error(PG_EINVAL, "pg_server_start: null");
    if (!ctx().stream) return set_error(PG_ENOINIT, "pg_init not called");
    pg_server *h = new pg_server;
    h->s.budget = hbm_budget ? hbm_budget : ctx().pool.budget() / 2;
    h->s.max_active = max_active ? max_active : 256;
    h->s.worker = std::thread([h] { h->s.run(); });
    *out = h;
    return PG_OK;
} catch (...) {
    return pg::exception_status();
}

int pg_server_submit(pg_server *h, pg_request *req) try {
    if (!h || !req || (req->op != PG_OP_MUL && req->op != PG_OP_DIV) || (req->an && !req->a) || (req->bn && !req->b) ||
        !req->r)
        return set_error(PG_EINVAL, "pg_server_submit: bad request");
    if (req->op == PG_OP_MUL && req->r_cap < host_len(req->a, req->an) + host_len(req->b, req->bn))
        return set_error(PG_EINVAL, "pg_server_submit: product capacity");
    if (req->op == PG_OP_DIV) {
        const size_t an = host_len(req->a, req->an), bn = host_len(req->b, req->bn);
        if (req->r_cap < (an >= bn ? an - bn + 1 : 1)) return set_error(PG_EINVAL, "pg_server_submit: quotient capacity");
        if (req->r2 && req->r2_cap < std::max(bn, an < bn ? an : (size_t)0))
            return set_error(PG_EINVAL, "pg_server_submit: remainder capacity");
    }
    req->status = 1;
    req->r_len = req->r2_len = 0;
    req->t_submit = now_s();
    req->t_admit = req->t_done = 0.0;
    {
        std::lock_guard<std::mutex> g(h->s.mu);
        if (h->s.stop) return set_error(PG_EINVAL, "pg_server_submit: server stopped");
        h->s.queue.push_back(req);
        h->s.stats.submitted++;
        h->s.outstanding++;
    }
    h->s.cv_work.notify_one();
    return PG_OK;
} catch (...) {
    return pg::exception_status();
}

int pg_server_wait(pg_server *h, pg_request *req) try {
    if (!h || !req) return set_error(PG_EINVAL, "pg_server_wait: null");
    std::unique_lock<std::mutex> g(h->s.mu);
    h->s.cv_done.wait(g, [&] { return req->status != 1; });
    return req->status;
} catch (...) {
    return pg::exception_status();
}

int pg_server_drain(pg_server *h) try {
    if (!h) return set_error(PG_EINVAL, "pg_server_drain: null");
    std::unique_lock<std::mutex> g(h->s.mu);
    h->s.cv_done.wait(g, [&] { return h->s.outstanding == 0; });
    return PG_OK;
} catch (...) {
    return pg::exception_status();
}

int pg_server_read_stats(pg_server *h, pg_server_stats *st) try {
    if (!h || !st) return set_error(PG_EINVAL, "pg_server_read_stats: null");
    std::lock_guard<std::mutex> g(h->s.mu);
    *st = h->s.stats;
    return PG_OK;
} catch (...) {
    return pg::exception_status();
}

int pg_server_stop(pg_server *h) try {
    if (!h) return PG_OK;
    {
        std::lock_guard<std::mutex> g(h->s.mu);
        h->s.stop = true;
    }
    h->s.cv_work.notify_all();
    if (h->s.worker.joinable()) h->s.worker.join();
    delete h;
    return PG_OK;
} catch (...) {
    return pg::exception_status();
}

}  // extern "C"
