// pg_dint.hip -- device-resident integer steps and Newton division.
//
// The reciprocal iteration and the quotient correction follow
// src/div_newton.cpp:36-129 step for step (same known/keep/precision
// schedule, same rounded-up divisor tops), with every product on the GPU
// product engine and every add / subtract / shift as one linear carry-scan
// launch.  The host reads back only the divisor's top limbs, the operand
// lengths and one sign word per correction step.
#include "pg_dint.h"
#include <algorithm>
#include <cstring>
#include "pgpu.h"

namespace pg {

bool DBuf::reset(size_t limbs) {
    release();
    d = (uint64_t *)ctx().pool.get(std::max<size_t>(limbs, 1) * sizeof(uint64_t));
    cap = d ? limbs : 0;
    return d != nullptr;
}

void DBuf::release() {
    if (d) ctx().pool.put(d);
    d = nullptr;
    cap = 0;
}

CarryNode lin_node(uint64_t *out, size_t m) {
    CarryNode n{};
    n.kind = NODE_LIN;
    n.parts = 1;
    n.m = m;
    n.out = out;
    n.out_len = m;
    n.xmul = 1;
    n.pow2 = -1;
    return n;
}

static size_t node_chunks(const CarryNode &n) {
    const size_t cover = std::max(n.m + (n.flag ? 1 : 0), n.out_len);
    return std::max<size_t>(1, (cover + CARRY_CHUNK - 1) / CARRY_CHUNK);
}

__global__ void k_fill64(uint64_t *p, size_t n, uint64_t v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// out = max(i + 1) over nonzero a[i] (out zeroed before).
__global__ void __launch_bounds__(256) k_toplen(const uint64_t *a, size_t n, unsigned long long *out) {
    unsigned long long best = 0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        if (a[i]) best = i + 1;
    for (int o = 32; o; o >>= 1) best = std::max(best, (unsigned long long)__shfl_xor(best, o));
    if ((threadIdx.x & 63) == 0 && best) atomicMax(out, best);
}

static unsigned grid_for(size_t n) { return (unsigned)std::min<size_t>(2048, std::max<size_t>(1, (n + 255) / 256)); }

int lin_run(const CarryNode *nodes, size_t n, const char *stage) {
    Context &C = ctx();
    if (!n) return PG_OK;
    CarryLaunch cl{};
    size_t nch = 0;
    double bytes = 0;
    DBuf desc;
    if (n == 1) {
        cl.one = nodes[0];
        cl.one.chunk0 = 0;
        nch = node_chunks(cl.one);
        bytes = 8.0 * ((double)std::min(cl.one.xn, cl.one.m) + (double)std::min(cl.one.yn, cl.one.m) +
                       (double)cl.one.out_len);
    } else {
        std::vector<CarryNode> v(nodes, nodes + n);
        std::vector<uint32_t> map;
        for (size_t i = 0; i < n; ++i) {
            v[i].chunk0 = nch;
            const size_t c = node_chunks(v[i]);
            map.insert(map.end(), c, (uint32_t)i);
            nch += c;
            bytes += 8.0 * ((double)std::min(v[i].xn, v[i].m) + (double)std::min(v[i].yn, v[i].m) +
                            (double)v[i].out_len);
        }
        const size_t nb = n * sizeof(CarryNode), mb = map.size() * sizeof(uint32_t);
        HostImg img = C.take_img(nb + mb);
        std::memcpy(img.data(), v.data(), nb);
        std::memcpy(img.data() + nb, map.data(), mb);
        if (!desc.reset((img.size() + 7) / 8)) return set_error(PG_ENOMEM, "linear descriptors");
        HIP_TRY(hipMemcpyAsync(desc.d, img.data(), img.size(), hipMemcpyHostToDevice, C.stream));
        cl.nodes = (const CarryNode *)desc.d;
        cl.chunk_node = (const uint32_t *)((const char *)desc.d + nb);
        C.hold(std::move(img));
    }
    DBuf flags((nch + 7) / 8);
    if (!flags.d) return set_error(PG_ENOMEM, "carry scratch");
    cl.nchunks = nch;
    cl.nnodes = n;
    cl.agg = (uint8_t *)flags.d;
    bool lin = true;
    for (size_t i = 0; i < n; ++i) lin &= nodes[i].kind == NODE_LIN;
    cl.node_kind = lin ? CARRY_LIN : CARRY_ANY;
    hipEvent_t e0 = nullptr;
    C.stage_begin(e0);
    HIP_TRY(launch_carry_kernels(cl, C.stream));
    C.stage_end(stage, e0, bytes);
    return PG_OK;
}

int dmul(uint64_t *r, const uint64_t *a, size_t an, const uint64_t *b, size_t bn) {
    if (!an || !bn) return dset(r, an + bn, nullptr, 0);
    return mul_device(r, a, an, ~0ull, b, bn, ~0ull, PG_PATH_AUTO, nullptr);
}

int dset(uint64_t *r, size_t n, const uint64_t *value, size_t vn) {
    Context &C = ctx();
    vn = std::min(vn, n);
    if (vn) {
        HostImg img = C.take_img(vn * 8);
        std::memcpy(img.data(), value, vn * 8);
        HIP_TRY(hipMemcpyAsync(r, img.data(), vn * 8, hipMemcpyHostToDevice, C.stream));
        C.hold(std::move(img));
    }
    if (n > vn) hipLaunchKernelGGL(k_fill64, dim3(grid_for(n - vn)), dim3(256), 0, C.stream, r + vn, n - vn, 0ull);
    HIP_TRY(hipGetLastError());
    return PG_OK;
}

int dcopy(uint64_t *r, const uint64_t *a, size_t n) {
    if (n) HIP_TRY(hipMemcpyAsync(r, a, n * 8, hipMemcpyDeviceToDevice, ctx().stream));
    return PG_OK;
}

int read_words(const uint64_t *src, size_t n, uint64_t *host) {
    Context &C = ctx();
    if (n) HIP_TRY(hipMemcpyAsync(host, src, n * 8, hipMemcpyDeviceToHost, C.stream));
    return C.sync();
}

int top_len(const uint64_t *a, size_t n, size_t *len) {
    Context &C = ctx();
    *len = 0;
    if (!n) return PG_OK;
    unsigned long long *w = (unsigned long long *)C.words;
    hipLaunchKernelGGL(k_fill64, dim3(1), dim3(64), 0, C.stream, C.words, (size_t)1, 0ull);
    hipLaunchKernelGGL(k_toplen, dim3(grid_for(n)), dim3(256), 0, C.stream, a, n, w);
    HIP_TRY(hipGetLastError());
    uint64_t v = 0;
    int st = read_words(C.words, 1, &v);
    *len = (size_t)v;
    return st;
}

int bit_length(const uint64_t *a, size_t trimmed, size_t *bits) {
    *bits = 0;
    if (!trimmed) return PG_OK;
    uint64_t top = 0;
    int st = read_words(a + trimmed - 1, 1, &top);
    if (st) return st;
    if (!top) return set_error(PG_EINVAL, "operand not trimmed");
    *bits = (trimmed - 1) * 64 + 64 - (size_t)__builtin_clzll(top);
    return PG_OK;
}

// ---- small host integers (at most a few limbs) ----------------------------------

static void trim(std::vector<uint64_t> &v) {
    while (!v.empty() && !v.back()) v.pop_back();
}

static int cmp_small(const std::vector<uint64_t> &a, const std::vector<uint64_t> &b) {
    const size_t n = std::max(a.size(), b.size());
    for (size_t i = n; i-- > 0;) {
        const uint64_t x = i < a.size() ? a[i] : 0, y = i < b.size() ? b[i] : 0;
        if (x != y) return x < y ? -1 : 1;
    }
    return 0;
}

// floor(2^e / d), d nonzero: restoring long division one bit at a time
// (the role of div_schoolbook(newton64_pow2(precision), bt), div_newton.cpp:47).
std::vector<uint64_t> pow2_div_small(size_t e, const std::vector<uint64_t> &d) {
    std::vector<uint64_t> q(e / 64 + 1, 0), r(d.size() + 1, 0);
    for (size_t i = e + 1; i-- > 0;) {
        uint64_t c = (i == e) ? 1 : 0;
        for (auto &w : r) {
            const uint64_t nc = w >> 63;
            w = (w << 1) | c;
            c = nc;
        }
        if (cmp_small(r, d) >= 0) {
            uint64_t br = 0;
            for (size_t k = 0; k < r.size(); ++k) {
                const uint64_t y = k < d.size() ? d[k] : 0;
                const uint64_t t = r[k] - y - br;
                br = (r[k] < y) || (r[k] - y < br);
                r[k] = t;
            }
            q[i / 64] |= 1ull << (i % 64);
        }
    }
    trim(q);
    return q;
}

// ---- Newton reciprocal and division ---------------------------------------------

// Test hooks (pg_debug_perturb): a one-shot change of the next division's
// first quotient estimate, and a factor on the next reciprocal's seed.
static int64_t g_perturb_q = 0;
static uint64_t g_perturb_seed = 0;

int debug_perturb(int what, int64_t amount) {
    if (what == 0) g_perturb_q = amount;
    else if (what == 1 && amount >= 0) g_perturb_seed = (uint64_t)amount;
    else return set_error(PG_EINVAL, "pg_debug_perturb: bad argument");
    return PG_OK;
}

// x = the reciprocal estimate of div_newton.cpp:36-68.  The fast form reads
// no word from the device inside the loop: a floor estimate never takes
// the overshoot branch (div_newton.cpp:57-62), and each step's sign word is
// read once at the end.  If one shows an overshoot the iteration runs
// again in the checked form, which reads every step's test and halves the
// scaled estimate as the reference does.
static int recip_run(DBuf &x, size_t &xl, const uint64_t *b, size_t bn, size_t nb, size_t bits, bool checked,
                     uint64_t seed_mul, bool &overshoot);

int recip_approx(DBuf &x, size_t &xl, const uint64_t *b, size_t bn, size_t nb, size_t bits) {
    const uint64_t seed_mul = g_perturb_seed;
    g_perturb_seed = 0;
    bool over = false;
    int st = recip_run(x, xl, b, bn, nb, bits, false, seed_mul, over);
    if (st || !over) return st;
    return recip_run(x, xl, b, bn, nb, bits, true, seed_mul, over);
}

static int recip_run(DBuf &x, size_t &xl, const uint64_t *b, size_t bn, size_t nb, size_t bits, bool checked,
                     uint64_t seed_mul, bool &overshoot) {
    overshoot = false;
    xl = 0;
    if (!bn || !nb || bits < nb) return PG_OK;   // div_newton.cpp:37
    const size_t target = bits - nb;
    size_t known = std::min<size_t>(target, NEWTON_SEED_BITS);
    const size_t keep = std::min(nb, known + 128);
    // bt = newton64_top_ceil(b, nb, keep) (div_newton.cpp:26-34) from the top limbs
    std::vector<uint64_t> bt;
    {
        const size_t drop = nb - keep, l0 = drop / 64;
        std::vector<uint64_t> w(bn - l0);
        int st = read_words(b + l0, w.size(), w.data());
        if (st) return st;
        const int r = (int)(drop % 64);
        bt.resize(w.size());
        for (size_t i = 0; i < w.size(); ++i)
            bt[i] = r ? (w[i] >> r) | (i + 1 < w.size() ? w[i + 1] << (64 - r) : 0) : w[i];
        if (drop) {
            bt.push_back(0);
            for (auto &v : bt)
                if (++v) break;
        }
        trim(bt);
    }
    const size_t precision = keep + known;
    std::vector<uint64_t> x0 = pow2_div_small(precision, bt);
    if (seed_mul > 1) {   // test hook: a seed seed_mul times too large
        unsigned __int128 c = 0;
        for (auto &w : x0) {
            c += (unsigned __int128)w * seed_mul;
            w = (uint64_t)c;
            c >>= 64;
        }
        if (c) x0.push_back((uint64_t)c);
    }
    xl = std::max(limbs_for_bits(known + 2) + 1, x0.size());
    if (!x.reset(xl)) return set_error(PG_ENOMEM, "reciprocal");
    int st = dset(x.d, xl, x0.data(), x0.size());
    if (st) return st;

    DBuf flags(64), nbuf, bx, e, xe, xn;
    size_t step = 0;
    while (known < target) {   // div_newton.cpp:49-65
        const size_t nk = std::min(target, known * 2);
        const size_t nkeep = std::min(nb, nk + 128);
        const uint64_t *NB = b;
        size_t nbl = bn;
        if (nkeep < nb) {   // next_b = (b >> (nb - nkeep)) + 1
            nbl = limbs_for_bits(nkeep + 1);
            if (!nbuf.reset(nbl)) return set_error(PG_ENOMEM, "reciprocal");
            CarryNode c = lin_node(nbuf.d, nbl);
            c.x = b;
            c.xn = bn;
            c.xsh = (int64_t)(nb - nkeep);
            c.cin = 1;
            if ((st = lin_one(c, "newton"))) return st;
            NB = nbuf.d;
        }
        const size_t next = nkeep + nk;
        size_t s = nk - known;
        if (!bx.reset(nbl + xl)) return set_error(PG_ENOMEM, "reciprocal");
        if ((st = dmul(bx.d, NB, nbl, x.d, xl))) return st;
        const size_t el = limbs_for_bits(next + 2);
        const size_t em = std::max(el, nbl + xl + limbs_for_bits(s) + 1);
        if (checked) {
            // div_newton.cpp:57-62: if next_b (x << s) >= 2^(next+1), the scaled
            // estimate is halved, x << (s - 1): test 2^(next+1) - 1 - (bx << s) >= 0
            if (!e.reset(em)) return set_error(PG_ENOMEM, "reciprocal");
            CarryNode t = lin_node(e.d, em);
            t.pow2 = (int64_t)(next + 1);
            t.y = bx.d;
            t.yn = nbl + xl;
            t.ysh = -(int64_t)s;
            t.yneg = 1;
            t.flag = flags.d;
            uint64_t ok = 0;
            if ((st = lin_one(t, "newton")) || (st = read_words(flags.d, 1, &ok))) return st;
            if (!ok) {
                overshoot = true;
                --s;   // s >= 1: nk > known
            }
        }
        // e = 2^(next+1) - (bx << s) >= 0
        if (!e.reset(el)) return set_error(PG_ENOMEM, "reciprocal");
        CarryNode c = lin_node(e.d, em);
        c.out_len = el;
        c.pow2 = (int64_t)(next + 1);
        c.y = bx.d;
        c.yn = nbl + xl;
        c.ysh = -(int64_t)s;
        c.yneg = 1;
        c.cin = 1;
        c.flag = flags.d + 1 + std::min<size_t>(step, 62);
        if ((st = lin_one(c, "newton"))) return st;
        if (checked) {   // two_scale - bx of unsigned precn_t saturates at 0 (sub.cpp:14)
            uint64_t pos = 0;
            if ((st = read_words(c.flag, 1, &pos))) return st;
            if (!pos && (st = dset(e.d, el, nullptr, 0))) return st;
        }
        if (!xe.reset(xl + el)) return set_error(PG_ENOMEM, "reciprocal");
        if ((st = dmul(xe.d, x.d, xl, e.d, el))) return st;
        // x = (x << s) * e >> next = x * e >> (next - s)
        const size_t nxl = std::max(limbs_for_bits(nk + 2) + 1, xl);
        if (!xn.reset(nxl)) return set_error(PG_ENOMEM, "reciprocal");
        CarryNode d = lin_node(xn.d, nxl);
        d.x = xe.d;
        d.xn = xl + el;
        d.xsh = (int64_t)(next - s);
        if ((st = lin_one(d, "newton"))) return st;
        std::swap(x, xn);
        xl = nxl;
        known = nk;
        ++step;
    }
    if (step && !checked) {
        std::vector<uint64_t> f(std::min<size_t>(step, 63));
        if ((st = read_words(flags.d + 1, f.size(), f.data()))) return st;
        for (uint64_t v : f) overshoot |= v != 1;
    }
    return PG_OK;
}

int div_mulinv_at(DBuf &q, size_t &ql, DBuf *r, size_t *rl, const uint64_t *a, size_t an, const uint64_t *b,
                  size_t bn, int depth);

// Brings a quotient estimate q (ql limbs) to the floor: r (rm limbs, two's
// complement, sign word at *flag) is the remainder of q.  The two loops of
// div_newton.cpp:114-125: while r < 0, q -= delta and r += b delta with
// delta = max(1, floor(-r / b)); while r >= b, q += delta and r -= b delta
// with delta = floor(r / b).  A good estimate leaves at most a unit or two,
// taken as unit steps; past two steps delta comes from a nested division of
// |r| by b (whose own estimate is within a few units), so any error is
// corrected in a bounded number of rounds.
static int correct(DBuf &q, size_t ql, DBuf &r, size_t rm, uint64_t *flag, const uint64_t *b, size_t bn,
                   int depth = 0) {
    DBuf q2(ql), r2(rm), m, p;
    if (!q2.d || !r2.d) return set_error(PG_ENOMEM, "division correction");
    int st;
    for (int it = 0; it < 2 + 64; ++it) {
        uint64_t f = 0;
        if ((st = read_words(flag, 1, &f))) return st;
        const bool neg = f == 0;
        if (!neg) {   // r >= 0: done unless r >= b
            CarryNode d = lin_node(r2.d, rm);
            d.x = r.d; d.xn = rm; d.y = b; d.yn = bn; d.yneg = 1; d.cin = 1; d.flag = flag;
            if ((st = lin_one(d, "correct"))) return st;
            uint64_t g = 0;
            if ((st = read_words(flag, 1, &g))) return st;
            if (g == 0) return PG_OK;
        }
        if (it < 2) {   // one unit
            if (neg) {   // q -= 1, r += b
                CarryNode c[2] = {lin_node(q2.d, ql), lin_node(r2.d, rm)};
                c[0].x = q.d; c[0].xn = ql; c[0].yneg = 1;          // q + ~0 = q - 1
                c[1].x = r.d; c[1].xn = rm; c[1].y = b; c[1].yn = bn; c[1].flag = flag;
                if ((st = lin_run(c, 2, "correct"))) return st;
            } else {     // q += 1; r2 = r - b and its sign word (>= 0) are already there
                CarryNode c = lin_node(q2.d, ql);
                c.x = q.d; c.xn = ql; c.cin = 1;
                if ((st = lin_one(c, "correct"))) return st;
            }
        } else {   // delta = floor(|r| / b), at least 1
            if (depth >= 3) return set_error(PG_EINTERNAL, "quotient correction did not converge");
            if (!m.reset(rm)) return set_error(PG_ENOMEM, "division correction");
            CarryNode c = lin_node(m.d, rm);   // |r|
            c.x = r.d; c.xn = rm;
            if (neg) {
                c.x = nullptr; c.xn = 0; c.y = r.d; c.yn = rm; c.yneg = 1; c.cin = 1;   // -r
            }
            if ((st = lin_one(c, "correct"))) return st;
            size_t mn = 0;
            if ((st = top_len(m.d, rm, &mn))) return st;
            DBuf dq;
            size_t dl = 0;
            if ((st = div_mulinv_at(dq, dl, nullptr, nullptr, m.d, mn, b, bn, depth + 1))) return st;
            if (!dl) {   // |r| < b: delta = 1
                const uint64_t one = 1;
                if (!dq.reset(1)) return set_error(PG_ENOMEM, "division correction");
                if ((st = dset(dq.d, 1, &one, 1))) return st;
                dl = 1;
            }
            if (!p.reset(bn + dl)) return set_error(PG_ENOMEM, "division correction");
            if ((st = dmul(p.d, b, bn, dq.d, dl))) return st;
            CarryNode u[2] = {lin_node(q2.d, ql), lin_node(r2.d, rm)};
            u[0].x = q.d; u[0].xn = ql; u[0].y = dq.d; u[0].yn = dl;                   // q + delta
            u[1].x = r.d; u[1].xn = rm; u[1].y = p.d; u[1].yn = bn + dl;               // r - b delta
            if (neg) {
                u[0].yneg = 1; u[0].cin = 1;                                          // q - delta
                u[1].flag = flag;                                                     // r + b delta
            } else {
                u[1].yneg = 1; u[1].cin = 1; u[1].flag = flag;
            }
            if ((st = lin_run(u, 2, "correct"))) return st;
        }
        std::swap(q, q2);
        std::swap(r, r2);
    }
    return set_error(PG_EINTERNAL, "quotient correction did not converge");
}

int recip_exact(DBuf &x, size_t &xl, const uint64_t *b, size_t bn, size_t bits) {
    size_t nb = 0;
    int st = bit_length(b, bn, &nb);
    if (st) return st;
    if ((st = recip_approx(x, xl, b, bn, nb, bits))) return st;
    return recip_finish(x, xl, b, bn, bits);
}

int recip_finish(DBuf &x, size_t &xl, const uint64_t *b, size_t bn, size_t bits) {
    int st;
    if (!xl) return PG_OK;
    DBuf p(bn + xl);
    if (!p.d) return set_error(PG_ENOMEM, "reciprocal");
    if ((st = dmul(p.d, b, bn, x.d, xl))) return st;
    const size_t rm = std::max(limbs_for_bits(bits + 1), bn + xl) + 1;
    DBuf r(rm);
    if (!r.d) return set_error(PG_ENOMEM, "reciprocal");
    CarryNode c = lin_node(r.d, rm);
    c.pow2 = (int64_t)bits;
    c.y = p.d; c.yn = bn + xl; c.yneg = 1; c.cin = 1;
    c.flag = ctx().words + 8;
    if ((st = lin_one(c, "correct"))) return st;
    if ((st = correct(x, xl, r, rm, ctx().words + 8, b, bn))) return st;
    return top_len(x.d, xl, &xl);
}

int div_mulinv(DBuf &q, size_t &ql, DBuf *r, size_t *rl, const uint64_t *a, size_t an, const uint64_t *b,
               size_t bn) {
    return div_mulinv_at(q, ql, r, rl, a, an, b, bn, 0);
}

int div_mulinv_at(DBuf &q, size_t &ql, DBuf *r, size_t *rl, const uint64_t *a, size_t an, const uint64_t *b,
                  size_t bn, int depth) {
    ql = 0;
    if (rl) *rl = 0;
    if (!an || !bn) return PG_OK;   // div_newton.cpp:90-93
    size_t na = 0, nb = 0;
    int st = bit_length(a, an, &na);
    if (!st) st = bit_length(b, bn, &nb);
    if (st) return st;
    if (na < nb) {   // a < b
        if (r) {
            if (!r->reset(an)) return set_error(PG_ENOMEM, "division");
            if ((st = dcopy(r->d, a, an))) return st;
            *rl = an;
        }
        return PG_OK;
    }
    const size_t scale = na + 128;   // div_newton.cpp:106
    DBuf x;
    size_t xl = 0;
    if ((st = recip_approx(x, xl, b, bn, nb, scale))) return st;
    DBuf ax(an + xl);
    if (!ax.d) return set_error(PG_ENOMEM, "division");
    if ((st = dmul(ax.d, a, an, x.d, xl))) return st;
    x.release();
    const size_t q0 = limbs_for_bits(na - nb + 2) + 1;
    if (!q.reset(q0)) return set_error(PG_ENOMEM, "division");
    CarryNode c = lin_node(q.d, q0);
    c.x = ax.d; c.xn = an + xl; c.xsh = (int64_t)scale;
    if (depth == 0 && g_perturb_q) {   // test hook: q0 +- |amount| units
        const uint64_t k = (uint64_t)(g_perturb_q < 0 ? -g_perturb_q : g_perturb_q);
        if ((st = dset(ctx().words + 10, 1, &k, 1))) return st;
        c.y = ctx().words + 10; c.yn = 1;
        if (g_perturb_q < 0) { c.yneg = 1; c.cin = 1; }
        g_perturb_q = 0;
    }
    if ((st = lin_one(c, "newton"))) return st;
    ax.release();
    DBuf p(bn + q0);
    if (!p.d) return set_error(PG_ENOMEM, "division");
    if ((st = dmul(p.d, b, bn, q.d, q0))) return st;
    const size_t rm = std::max(an, bn + q0) + 1;
    DBuf rr(rm);
    if (!rr.d) return set_error(PG_ENOMEM, "division");
    CarryNode d = lin_node(rr.d, rm);
    d.x = a; d.xn = an; d.y = p.d; d.yn = bn + q0; d.yneg = 1; d.cin = 1;
    d.flag = ctx().words + 9;
    if ((st = lin_one(d, "correct"))) return st;
    p.release();
    if ((st = correct(q, q0, rr, rm, ctx().words + 9, b, bn, depth))) return st;
    if ((st = top_len(q.d, q0, &ql))) return st;
    if (r) {
        if ((st = top_len(rr.d, std::min(rm, bn), rl))) return st;
        *r = std::move(rr);
    }
    return PG_OK;
}

}  // namespace pg
