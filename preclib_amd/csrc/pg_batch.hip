// pg_batch.hip -- planner and executor for many products in one submission.
//
// A product is decomposed exactly as preclib's mul_into decomposes it
// (src/mul_basic.cpp:238-276): small products (a side <= 24 limbs, or both
// <= 192) are schoolbook leaves; hi > 2 lo is a blocked multiply with blocks
// of 2 lo limbs of the longer operand (mul_unbalanced_into, 207-236), each
// block again dispatched by mul_into; hi > 6144 is an NTT leaf (mul_ntt, with
// its fall back to the FFT beyond 2^24 limbs or 2^26 points); otherwise an
// FFT leaf (mul_fft, with its 16/8-bit digit rule).
//
// Every leaf writes 128-bit limb sums.  Leaves of one kind and transform
// length form a bucket and run as one launch (blockIdx.z = leaf).  Blocked
// multiplies are carry nodes that add the (at most two) block products
// covering each limb; nested blocked multiplies are resolved bottom-up, one
// carry launch per nesting depth for all products.  The whole sequence is
// enqueued on the library stream and can be captured once into a hipGraph
// and replayed.
#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <vector>
#include "pg_batch.h"

namespace pg {

int launch_carry_nodes(size_t nnodes, double bytes, const CarryNode *dnodes, const uint32_t *dchunk,
                       size_t nchunks, uint8_t *agg, int node_kind);

namespace {

constexpr size_t NTT_MAX_LIMBS = (size_t)1 << 24;           // mul_ntt.cpp:27

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace

// Offsets into the device arena are recorded while planning and turned into
// pointers once the arena exists.
struct Batch::Alloc {
    size_t bytes = 0;
    size_t take(size_t n) {
        const size_t off = bytes;
        bytes += align_up(n, 256);
        return off;
    }
};

Batch::~Batch() { release(); }

void Batch::release() {
    if (graph_exec) hipGraphExecDestroy(graph_exec);
    if (graph) hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
    if (arena) ctx().pool.put(arena);
    arena = nullptr;
}

// Sums buffer of `len` limbs (two arrays) in the arena.
Batch::SumRef Batch::new_sums(size_t len) {
    SumRef s;
    s.lo = alloc->take(len * 8);
    s.hi = alloc->take(len * 8);
    s.len = len;
    return s;
}

// Plans x * y into `target`: either limb sums (target.len limbs) or, when
// final is set, normalized limbs at `final`.  Returns 0 or a status.
int Batch::plan(const uint64_t *x, size_t xn, const uint64_t *y, size_t yn, int path, int job, int depth,
                const SumRef *target, uint64_t *final, size_t final_len) {
    const size_t lo = std::min(xn, yn), hi = std::max(xn, yn);
    // Leaves write sums; a final target needs a plain node to normalize them.
    // (pad: sums allocated for len rounded up to a multiple of pad limbs)
    auto leaf_target = [&](size_t len, size_t pad = 1) -> SumRef {
        if (target) return *target;
        SumRef s = new_sums((len + pad - 1) / pad * pad);
        NodePlan np{};
        np.kind = NODE_PLAIN;
        np.parts = 1;
        np.sums = s;
        np.m = len;
        np.out = final;
        np.out_len = final_len;
        np.depth = depth;
        nodes.push_back(np);
        return s;
    };
    const bool sqr = (x == y && xn == yn);
    int kind;   // 0 small, 1 blocked, 2 ntt, 3 fft
    if (path == PG_PATH_NTT || path == PG_PATH_NTT_REF) {
        kind = 2;
    } else if (path == PG_PATH_FFT) {
        kind = 3;
    } else if (lo <= DISPATCH_SMALL_SIDE || hi <= DISPATCH_SMALL_HI) {
        kind = 0;
    } else if (!sqr && hi > 2 * lo) {
        kind = 1;
    } else {
        kind = 2;   // the NTT for every transform-sized product (DISPATCH_*)
    }
    if (kind == 2) {
        const size_t n16 = 4 * (xn + yn);   // digit count with full top limbs
        size_t n = 1;
        while (n < n16) n <<= 1;
        if (hi > NTT_MAX_LIMBS || n > ((size_t)1 << 26)) kind = 3;   // mul_ntt.cpp:497-506
    }
    if (kind == 0 && lo > SMALL_MAX_SHORT) kind = 3;   // forced paths below the threshold never reach here
    const size_t plen = xn + yn;
    switch (kind) {
    case 0: {
        SmallPlan sp{};
        sp.x = x; sp.y = y; sp.xn = xn; sp.yn = yn;
        sp.sums = leaf_target(plen);
        small.push_back(sp);
        return PG_OK;
    }
    case 2: {
        const int mode = path == PG_PATH_NTT_REF ? MODE_REF : MODE_WIDE;
        const int dpl = 64 / mode_digit_bits(mode);
        int logn = 0;
        while (((size_t)1 << logn) < (size_t)dpl * plen) ++logn;
        if (logn < NTT_MIN_LOG) logn = NTT_MIN_LOG;
        if (logn > NTT_MAX_LOG) return set_error(PG_ERANGE, "NTT leaf too long");
        const size_t n = (size_t)1 << logn;
        const int np = mode_primes(mode);
        auto &bucket = ntt[{mode, logn}];
        LeafPlan lp{};
        lp.x = x; lp.y = y; lp.xn = xn; lp.yn = yn;
        // room for whole pieces of packed sums (at most 8 limbs; build())
        if (!target) lp.node = (long)nodes.size();
        lp.sums = leaf_target(plen, 8);
        lp.has_spec = plan_l1(logn) > 0;
        if (lp.has_spec) {
            // An operand already transformed in this bucket (the short operand
            // of every block of a blocked multiply, a factor shared by two
            // products of a level, both operands of a square) is transformed
            // once: the forward column kernel of the bucket's one launch
            // finishes before its row kernel reads the spectra.
            const uint64_t *ptr[2] = {x, y};
            const size_t len[2] = {xn, yn};
            for (int op = 0; op < 2; ++op) {
                const auto key = std::make_tuple(mode, logn, ptr[op], len[op]);
                const auto it = spec_owner.find(key);
                if (it != spec_owner.end()) {
                    lp.ext_leaf[op] = it->second.first;
                    lp.ext_op[op] = it->second.second;
                } else {
                    spec_owner[key] = {(long)bucket.size(), op};
                }
            }
            // spectra this leaf writes: [operand][prime][n] up to its last own operand
            const int nown = lp.ext_leaf[1] < 0 ? 2 : lp.ext_leaf[0] < 0 ? 1 : 0;
            lp.spec = alloc->take((size_t)nown * np * n * 4);
        }
        lp.prod = alloc->take(np * n * 4);
        bucket.push_back(lp);
        return PG_OK;
    }
    case 3: {
        const int bits = hi <= 1024 ? 16 : 8;   // mul_fft.cpp:390-393
        const int dpl = 64 / bits;
        int logn = 0;
        while (((size_t)1 << logn) < (size_t)dpl * plen) ++logn;
        if (logn < FFT_MIN_LOG) logn = FFT_MIN_LOG;
        if (logn > FFT_MAX_LOG) return set_error(PG_ERANGE, "FFT leaf too long");
        const int lt = fft_lt(logn);
        const int seg = std::min(1 << lt, dpl);
        const int parts = dpl / seg;
        LeafPlan lp{};
        lp.x = x; lp.y = y; lp.xn = xn; lp.yn = yn;
        if (parts == 1) {
            lp.sums = leaf_target(plen);
        } else {
            // half-limb parts need their own node; sums have 2 entries per limb
            lp.sums = new_sums(2 * plen);
            NodePlan np{};
            np.kind = NODE_PLAIN;
            np.parts = 2;
            np.sums = lp.sums;
            np.m = plen;
            if (target) { np.out_sums = *target; np.to_sums = true; }
            else { np.out = final; np.out_len = final_len; }
            np.depth = depth;
            nodes.push_back(np);
        }
        lp.parts = parts;
        lp.spec = alloc->take(((size_t)1 << logn) * 16);
        lp.maxerr = alloc->take(8);
        lp.job = job;
        fft[{bits, logn}].push_back(lp);
        return PG_OK;
    }
    default: break;
    }
    // blocked (mul_unbalanced_into): blocks of 2*lo limbs of the longer operand.
    // Block k's product (at most 3 lo = 1.5 B limbs) is summed into layer k & 1
    // of the node at limb k B: the blocks of one layer are 2 B apart and do not
    // overlap, so each layer is one array of the node's length, and the carry
    // reads limb j's sum as the two layers' words at j (where a block covers j)
    // instead of going through per-block descriptors.
    const uint64_t *lx = xn >= yn ? x : y, *sy = xn >= yn ? y : x;
    const size_t ln = std::max(xn, yn), sn = std::min(xn, yn);
    const size_t B = 2 * sn;
    NodePlan np{};
    np.kind = NODE_BLOCKED;
    np.B = B;
    np.blen = B + sn;
    np.nblocks = (int)((ln + B - 1) / B);
    np.m = plen;
    np.depth = depth;
    np.sums = new_sums(plen);
    np.sums1 = new_sums(plen);
    if (target) { np.out_sums = *target; np.to_sums = true; }
    else { np.out = final; np.out_len = final_len; }
    nodes.push_back(np);
    for (size_t off = 0; off < ln; off += B) {
        const size_t pn = std::min(B, ln - off);
        const SumRef &layer = ((off / B) & 1) ? np.sums1 : np.sums;
        SumRef bs;
        bs.lo = layer.lo + off * 8;
        bs.hi = layer.hi + off * 8;
        bs.len = pn + sn;
        int st = plan(lx + off, pn, sy, sn, PG_PATH_AUTO, job, depth + 1, &bs, nullptr, 0);
        if (st) return st;
    }
    return PG_OK;
}

int Batch::build(const pg_job *jobs, size_t n) {
    Context &C = ctx();
    Alloc al;
    alloc = &al;
    njobs = n;
    spec_owner.clear();
    for (size_t i = 0; i < n; ++i) {
        const pg_job &j = jobs[i];
        if (!j.a || !j.b || !j.r || !j.an || !j.bn || j.r_cap < j.an + j.bn)
            return set_error(PG_EINVAL, "pg_batch: bad job " + std::to_string(i));
        int st = plan(j.a, j.an, j.b, j.bn, j.path, (int)i, 0, nullptr, j.r, j.an + j.bn);
        if (st) return st;
    }
    // The plan variant of each NTT bucket, and packed limb sums, as in the
    // single-product path (sum_pack_log), for leaves whose sums only their own
    // plain carry node reads.  Such a node's length is rounded up to whole
    // pieces (at most 8 limbs, room plan() left): the sums beyond the product
    // are exact zeros, and the node writes only out_len limbs.
    for (auto &kv : ntt) {
        const int mode = kv.first.first, logn = kv.first.second;
        const int v = C.ntt_batch_variant(logn, mode, kv.second.size() << logn);
        ntt_variant[kv.first] = v;
        const int pl = sum_pack_log(mode, plan_lt(logn, v));
        if (!pl) continue;
        for (auto &lp : kv.second) {
            if (lp.node < 0) continue;
            NodePlan &nd = nodes[lp.node];
            lp.pack_log = nd.pack_log = pl;
            nd.m = (nd.m + ((size_t)1 << pl) - 1) >> pl << pl;
        }
    }
    // descriptor arrays live at the end of the arena
    size_t nleaf_ntt = 0, nleaf_fft = 0;
    for (auto &kv : ntt) nleaf_ntt += kv.second.size();
    for (auto &kv : fft) nleaf_fft += kv.second.size();
    size_t small_blocks_total = 0;
    for (auto &s : small) small_blocks_total += small_blocks(s.xn, s.yn);
    size_t nchunks = 0;
    int maxdepth = 0;
    for (auto &nd : nodes) maxdepth = std::max(maxdepth, nd.depth);
    const size_t d_ntt = al.take(nleaf_ntt * sizeof(NttLeaf));
    const size_t d_fft = al.take(nleaf_fft * sizeof(FftLeaf));
    const size_t d_small = al.take(small.size() * sizeof(SmallLeaf));
    const size_t d_smallmap = al.take(small_blocks_total * 4);
    const size_t d_nodes = al.take(nodes.size() * sizeof(CarryNode));
    for (auto &nd : nodes) nchunks += (std::max(nd.m, nd.to_sums ? nd.out_sums.len : nd.out_len) + CARRY_CHUNK - 1) / CARRY_CHUNK;
    const size_t d_chunkmap = al.take(nchunks * 4);
    const size_t d_agg = al.take(nchunks);
    arena_bytes = al.bytes;
    arena = (char *)C.pool.get(arena_bytes);
    alloc = nullptr;
    if (!arena) return set_error(PG_ENOMEM, "pg_batch: device budget exhausted (" + std::to_string(arena_bytes) + " bytes)");
    auto P = [&](size_t off) { return arena + off; };
    auto SLO = [&](const SumRef &s) { return (uint64_t *)P(s.lo); };
    auto SHI = [&](const SumRef &s) { return (uint64_t *)P(s.hi); };

    // host images of the descriptor arrays (every descriptor is written
    // below; the chunk aggregates start at zero)
    HostImg img = C.take_img(arena_bytes - d_ntt);
    auto H = [&](size_t off) { return img.data() + (off - d_ntt); };
    std::memset(H(d_agg), 0, nchunks);

    // NTT buckets
    size_t k = 0;
    for (auto &kv : ntt) {
        Bucket b;
        b.kind = 2;
        b.mode = kv.first.first;
        b.logn = kv.first.second;
        b.first = k;
        b.count = kv.second.size();
        const size_t pn = (size_t)mode_primes(b.mode) << b.logn;   // spectra of one operand
        for (auto &lp : kv.second) {
            NttLeaf L{};
            L.a = lp.x; L.b = lp.y; L.an = lp.xn; L.bn = lp.yn;
            L.spec = lp.has_spec ? (uint32_t *)P(lp.spec) : nullptr;
            for (int op = 0; op < 2; ++op)
                if (lp.ext_leaf[op] >= 0)
                    L.ext[op] = (const uint32_t *)P(kv.second[lp.ext_leaf[op]].spec) + lp.ext_op[op] * pn;
            L.prod = (uint32_t *)P(lp.prod);
            L.slo = SLO(lp.sums); L.shi = SHI(lp.sums);
            L.sum_len = lp.pack_log ? nodes[lp.node].m : lp.xn + lp.yn;
            L.pack_log = lp.pack_log;
            std::memcpy(H(d_ntt) + k * sizeof(NttLeaf), &L, sizeof L);
            b.io += (double)(lp.xn + lp.yn) * 16.0;
            ++k;
        }
        b.dleaves = P(d_ntt) + b.first * sizeof(NttLeaf);
        buckets.push_back(b);
    }
    k = 0;
    for (auto &kv : fft) {
        Bucket b;
        b.kind = 3;
        b.bits = kv.first.first;
        b.logn = kv.first.second;
        b.first = k;
        b.count = kv.second.size();
        for (auto &lp : kv.second) {
            FftLeaf L{};
            L.a = lp.x; L.b = lp.y; L.an = lp.xn; L.bn = lp.yn;
            L.spec = (double2 *)P(lp.spec);
            L.slo = SLO(lp.sums); L.shi = SHI(lp.sums);
            L.sum_len = (lp.xn + lp.yn) * lp.parts;
            L.maxerr = (uint64_t *)P(lp.maxerr);
            std::memcpy(H(d_fft) + k * sizeof(FftLeaf), &L, sizeof L);
            fft_err.push_back({lp.job, (uint64_t *)P(lp.maxerr)});
            b.io += (double)(lp.xn + lp.yn) * 16.0;
            ++k;
        }
        b.dleaves = P(d_fft) + b.first * sizeof(FftLeaf);
        buckets.push_back(b);
    }
    // small leaves
    size_t blk = 0;
    uint32_t *smap = (uint32_t *)H(d_smallmap);
    for (size_t i = 0; i < small.size(); ++i) {
        const SmallPlan &sp = small[i];
        SmallLeaf L{};
        L.a = sp.x; L.b = sp.y; L.an = sp.xn; L.bn = sp.yn;
        L.slo = SLO(sp.sums); L.shi = SHI(sp.sums);
        L.sum_len = sp.xn + sp.yn;
        L.blk0 = blk;
        std::memcpy(H(d_small) + i * sizeof(SmallLeaf), &L, sizeof L);
        const size_t nb = small_blocks(sp.xn, sp.yn);
        for (size_t q = 0; q < nb; ++q) smap[blk + q] = (uint32_t)i;
        blk += nb;
        small_io += (double)(sp.xn + sp.yn) * 16.0;
    }
    small_leaves = (SmallLeaf *)P(d_small);
    small_map = (uint32_t *)P(d_smallmap);
    small_nblocks = blk;
    // carry nodes, grouped by depth (deepest first); each depth is two
    // launches: the plain whole-limb nodes (a kernel instance with their loader
    // alone, few registers) and the rest
    size_t chunk = 0, nidx = 0;
    auto plain_node = [](const NodePlan &nd) { return nd.kind == NODE_PLAIN && nd.parts == 1; };
    for (int d = maxdepth; d >= 0; --d)
    for (int pass = 0; pass < 2; ++pass) {
        Level lv;
        lv.first_chunk = chunk;
        lv.first_node = nidx;
        lv.node_kind = pass == 0 ? CARRY_PLAIN : CARRY_ANY;
        size_t count = 0;   // nodes of this level, written straight into the image
        for (auto &nd : nodes) {
            if (nd.depth != d || plain_node(nd) != (pass == 0)) continue;
            CarryNode N{};
            N.kind = nd.kind;
            N.parts = nd.parts;
            N.m = nd.m;
            N.slo = SLO(nd.sums);
            N.shi = SHI(nd.sums);
            if (nd.kind == NODE_PLAIN) {
                N.pack_log = nd.pack_log;
            } else {
                N.slo1 = SLO(nd.sums1);
                N.shi1 = SHI(nd.sums1);
                N.B = nd.B;
                N.blen = nd.blen;
                N.nblocks = nd.nblocks;
            }
            if (nd.to_sums) {
                N.out = SLO(nd.out_sums);
                N.out_hi = SHI(nd.out_sums);
                N.out_len = nd.out_sums.len;
            } else {
                N.out = nd.out;
                N.out_hi = nullptr;
                N.out_len = nd.out_len;
            }
            N.chunk0 = chunk - lv.first_chunk;   // relative to the level's launch
            const size_t nc = (std::max(N.m, N.out_len) + CARRY_CHUNK - 1) / CARRY_CHUNK;
            uint32_t *cmap = (uint32_t *)H(d_chunkmap) + chunk;
            for (size_t q = 0; q < nc; ++q) cmap[q] = (uint32_t)count;
            chunk += nc;
            std::memcpy(H(d_nodes) + (nidx + count) * sizeof(CarryNode), &N, sizeof N);
            lv.bytes += (double)N.m * 16.0 * (N.kind == NODE_BLOCKED ? 2 : 1) + (double)N.out_len * 8.0;
            ++count;
        }
        lv.count = count;
        lv.nchunks = chunk - lv.first_chunk;
        lv.dnodes = (const CarryNode *)(P(d_nodes) + nidx * sizeof(CarryNode));
        lv.dchunk = (const uint32_t *)(P(d_chunkmap) + lv.first_chunk * 4);
        lv.agg = (uint8_t *)P(d_agg) + lv.first_chunk;
        nidx += count;
        if (lv.count) levels.push_back(lv);
    }
    // launch descriptors and twiddle tables now: table creation allocates and
    // copies, which a capturing stream would reject
    for (auto &b : buckets) {
        if (b.kind == 2) {
            NttLaunch &L = b.nl;
            L.leaves = (const NttLeaf *)b.dleaves;
            L.nleaves = (int)b.count;
            L.io_bytes = b.io;
            L.logn = b.logn;
            L.mode = b.mode;
            L.variant = ntt_variant.at({b.mode, b.logn});
            L.order = ctx().ntt_order;
            int st = fill_ntt_tables(L);
            if (st) return st;
        } else {
            FftLaunch &L = b.fl;
            L.leaves = (const FftLeaf *)b.dleaves;
            L.nleaves = (int)b.count;
            L.io_bytes = b.io;
            L.logn = b.logn;
            L.bits = b.bits;
            int st = fill_fft_tables(L);
            if (st) return st;
        }
    }
    // one upload of every descriptor (stream-ordered; the host image stays
    // alive until the context's next sync)
    if (hipMemcpyAsync(P(d_ntt), img.data(), img.size(), hipMemcpyHostToDevice, C.stream) != hipSuccess)
        return set_error(PG_EDEVICE, "pg_batch: descriptor upload failed");
    C.hold(std::move(img));
    return PG_OK;
}

// Enqueues every launch of the batch on the library stream.
int Batch::enqueue() {
    Context &C = ctx();
    const StageHook hk = C.hook();
    for (auto &b : buckets) {
        if (b.kind == 2) {
            if (launch_ntt(b.nl, C.stream, hk) != hipSuccess) return set_error(PG_EDEVICE, "pg_batch: NTT launch failed");
        } else {
            if (launch_fft(b.fl, C.stream, hk) != hipSuccess) return set_error(PG_EDEVICE, "pg_batch: FFT launch failed");
        }
    }
    if (small_nblocks) {
        SmallLaunch S{};
        S.leaves = small_leaves;
        S.blk_leaf = small_map;
        if (launch_small(S, small_nblocks, C.stream, hk, small_io) != hipSuccess)
            return set_error(PG_EDEVICE, "pg_batch: small-product launch failed");
    }
    for (auto &lv : levels) {
        int st = launch_carry_nodes(lv.count, lv.bytes, lv.dnodes, lv.dchunk, lv.nchunks, lv.agg, lv.node_kind);
        if (st) return st;
    }
    return PG_OK;
}

// Replays the captured graph (captured on first use).
int Batch::run() {
    Context &C = ctx();
    if (!use_graph || C.timing) return enqueue();
    if (!graph_exec) {
        if (hipStreamBeginCapture(C.stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return set_error(PG_EDEVICE, "pg_batch: graph capture failed to start");
        int st = enqueue();
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(C.stream, &g);
        if (st) {
            if (g) hipGraphDestroy(g);
            return st;
        }
        if (e != hipSuccess || !g) return set_error(PG_EDEVICE, "pg_batch: graph capture failed");
        graph = g;
        if (hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0) != hipSuccess)
            return set_error(PG_EDEVICE, "pg_batch: graph instantiation failed");
    }
    if (hipGraphLaunch(graph_exec, C.stream) != hipSuccess) return set_error(PG_EDEVICE, "pg_batch: graph launch failed");
    return PG_OK;
}

// Largest FFT rounding distance of each job (waits for the device).
int Batch::read_errors(std::vector<double> &err) {
    Context &C = ctx();
    err.assign(njobs, 0.0);
    if (fft_err.empty()) return PG_OK;
    std::vector<uint64_t> bits(fft_err.size());
    for (size_t i = 0; i < fft_err.size(); ++i)
        if (hipMemcpyAsync(&bits[i], fft_err[i].second, 8, hipMemcpyDeviceToHost, C.stream) != hipSuccess)
            return set_error(PG_EDEVICE, "pg_batch: reading rounding distances failed");
    if (C.wait() != hipSuccess) return set_error(PG_EDEVICE, "pg_batch: device error");
    for (size_t i = 0; i < fft_err.size(); ++i) {
        double e;
        std::memcpy(&e, &bits[i], 8);
        err[fft_err[i].first] = std::max(err[fft_err[i].first], e);
    }
    return PG_OK;
}

}  // namespace pg
