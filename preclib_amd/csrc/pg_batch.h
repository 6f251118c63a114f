// pg_batch.h -- many products planned and launched together (pg_batch.hip).
#pragma once
#include <map>
#include <tuple>
#include <utility>
#include <vector>
#include "pg_runtime.h"
#include "pgpu.h"

namespace pg {

class Batch {
public:
    ~Batch();
    int build(const pg_job *jobs, size_t n);   // job pointers are device pointers
    int run();                                 // enqueue (graph replay when enabled)
    int enqueue();
    int read_errors(std::vector<double> &err);
    void release();
    bool use_graph = true;
    size_t arena_bytes = 0;

private:
    struct Alloc;
    struct SumRef {
        size_t lo = 0, hi = 0, len = 0;   // arena offsets
    };
    struct LeafPlan {
        const uint64_t *x, *y;
        size_t xn, yn;
        SumRef sums;
        size_t spec = 0, prod = 0, maxerr = 0;
        bool has_spec = false;
        int parts = 1, job = -1;
        // operand k's spectra from leaf ext_leaf[k] (its operand ext_op[k]) of
        // the same bucket; -1: transformed by this leaf
        long ext_leaf[2] = {-1, -1};
        int ext_op[2] = {0, 0};
        // the plain carry node that alone reads this leaf's sums (a top-level
        // product), or -1 (a block of a blocked product); pack_log: packed
        // sums (NttLeaf::pack_log)
        long node = -1;
        int pack_log = 0;
    };
    struct SmallPlan {
        const uint64_t *x, *y;
        size_t xn, yn;
        SumRef sums;
    };
    struct NodePlan {
        int kind = NODE_PLAIN, parts = 1, depth = 0;
        SumRef sums, sums1;           // plain: sums; blocked: the two layers (CarryNode)
        size_t B = 0, blen = 0, m = 0;
        int nblocks = 0;
        uint64_t *out = nullptr;
        size_t out_len = 0;
        SumRef out_sums;
        bool to_sums = false;
        int pack_log = 0;
    };
    struct Bucket {
        int kind = 0, mode = 0, bits = 0, logn = 0;
        size_t first = 0, count = 0;
        double io = 0.0;
        const char *dleaves = nullptr;
        NttLaunch nl{};             // prepared at plan time (tables built before any capture)
        FftLaunch fl{};
    };
    struct Level {
        size_t first_chunk = 0, first_node = 0, count = 0, nchunks = 0;
        const CarryNode *dnodes = nullptr;
        const uint32_t *dchunk = nullptr;
        uint8_t *agg = nullptr;
        int node_kind = CARRY_ANY;   // CARRY_PLAIN: plain whole-limb nodes only
        double bytes = 0.0;   // carry traffic, for the stage timing
    };
    SumRef new_sums(size_t len);
    int plan(const uint64_t *x, size_t xn, const uint64_t *y, size_t yn, int path, int job, int depth,
             const SumRef *target, uint64_t *final, size_t final_len);

    Alloc *alloc = nullptr;
    char *arena = nullptr;
    size_t njobs = 0;
    std::map<std::pair<int, int>, std::vector<LeafPlan>> ntt, fft;   // (mode|bits, logn)
    std::map<std::pair<int, int>, int> ntt_variant;                  // plan variant of each NTT bucket
    // first NTT leaf (index in its bucket, operand) transforming each
    // (mode, logn, operand pointer, length)
    std::map<std::tuple<int, int, const uint64_t *, size_t>, std::pair<long, int>> spec_owner;
    std::vector<SmallPlan> small;
    std::vector<NodePlan> nodes;
    std::vector<Bucket> buckets;
    std::vector<Level> levels;
    std::vector<std::pair<int, uint64_t *>> fft_err;
    SmallLeaf *small_leaves = nullptr;
    uint32_t *small_map = nullptr;
    size_t small_nblocks = 0;
    double small_io = 0.0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
};

// Table pointers of a transform length (shared by single and batched launches).
int fill_ntt_tables(NttLaunch &L, bool create = true);
int fill_fft_tables(FftLaunch &L);

}  // namespace pg
