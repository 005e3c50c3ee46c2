// internal.h - declarations shared by the kernel and host translation units.
#pragma once

#include <memory>
#include <mutex>
#include <string>
#include <map>
#include <vector>

#include "common.h"

namespace rinnd {

// Compute units of the current device (the persistent kernels size their grids by it).
inline int device_cus() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) {
        (void)hipGetLastError();
        return 256;
    }
    return cus;
}

// ---- device memory -------------------------------------------------------------------------------
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    bool owned = true;  // false for a view into memory someone else releases
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n), owned(o.owned) { o.p = nullptr; o.n = 0; }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            owned = o.owned;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~DBuf() { release(); }
    void alloc(size_t count) {
        release();
        owned = true;
        if (count == 0) return;
        hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            throw Error{RINND_ERR_OOM, "hipMalloc of " + std::to_string(count * sizeof(T)) +
                                           " bytes failed: " + hipGetErrorString(e)};
        }
        n = count;
    }
    void view(T* ptr, size_t count) {
        release();
        p = ptr;
        n = count;
        owned = false;
    }
    // grow-only reuse
    void ensure(size_t count) {
        if (count > n) alloc(count);
    }
    void release() {
        if (p && owned) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Records of the ordered-apply pipeline.  A generator emits URec (one update, optionally applied to
// both endpoints); the scatter groups them by destination as SegRec; a per-destination pass sorts
// each group by `key` (the reference's generation order) and applies it sequentially.
struct URec {
    uint64_t key;
    int32_t a;        // destination (or first endpoint)
    int32_t b;        // value pushed into a (and destination of the mirrored record)
    float v;          // distance or priority
    int32_t two_sided;
};
struct SegRec {
    uint64_t key;
    int32_t x;
    float v;
};
static_assert(sizeof(SegRec) == 16, "SegRec layout");
static_assert(sizeof(URec) == 24, "URec layout");

// ---- forest --------------------------------------------------------------------------------------
// A tree's arrays either own their memory or are views into `pool`, one allocation shared by all
// trees of a forest (one hipMalloc/hipFree per forest instead of four per tree).
struct TreeDev {
    int64_t n_nodes = 0;
    int64_t cap_nodes = 0;
    std::shared_ptr<DBuf<char>> pool;
    DBuf<float> hyper;       // cap_nodes x dim
    DBuf<float> offsets;     // cap_nodes
    DBuf<int32_t> children;  // cap_nodes x 2
    DBuf<int32_t> indices;   // n
};

}  // namespace rinnd

struct rinnd_forest {
    int64_t dim = 0, n = 0;
    std::vector<rinnd::TreeDev> trees;
};

namespace rinnd {
struct Workspace;
}

struct rinnd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    std::mutex mu;
    rinnd::DBuf<uint64_t> jump_tables;  // 64 x 256 x 4 words
    hipEvent_t ev[12] = {};
    std::vector<hipEvent_t> ev_pool;    // events of deferred launch timings (released by rinnd_ctx_destroy)
    rinnd::Workspace* ws = nullptr;     // build scratch kept between calls (released by rinnd_ctx_destroy)
    std::map<uint64_t, int> join_shapes;  // FP32 join tile shape picked by the tuner, per (metric, row stride, C)
};

namespace rinnd {

// Times one kernel launch with a pair of events on its stream and adds it to a rinnd_kernel_timing
// (a null target makes it a no-op).  end() waits for the launch.
// Launches whose durations are read once, after the build has synchronised anyway: event pairs come from a
// context-owned pool and flush() adds every recorded launch to its rinnd_kernel_timing.
struct LaunchLog {
    std::vector<hipEvent_t>* pool = nullptr;
    size_t used = 0;
    struct Rec {
        hipEvent_t a, b;
        rinnd_kernel_timing* out;
    };
    std::vector<Rec> recs;
    hipEvent_t event() {
        if (used == pool->size()) {
            hipEvent_t e = nullptr;
            // timing only: no system-scope fence (no L2 writeback and invalidate between the timed kernels)
            HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
            pool->push_back(e);
        }
        return (*pool)[used++];
    }
    void flush() {
        for (const Rec& r : recs) {
            HIP_TRY(hipEventSynchronize(r.b));
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, r.a, r.b));
            r.out->ms += ms;
            r.out->launches += 1;
        }
        recs.clear();
        used = 0;
    }
};

struct KTimer {
    hipEvent_t a = nullptr, b = nullptr;
    rinnd_kernel_timing* out = nullptr;
    LaunchLog* log = nullptr;  // set: no wait per launch, durations are read at log->flush()
    mutable hipEvent_t la = nullptr;
    void begin(hipStream_t s) const {
        if (!out) return;
        if (log) la = log->event();
        HIP_TRY(hipEventRecord(log ? la : a, s));
    }
    void stop(hipStream_t s) const {
        if (!out) return;
        if (log) {
            hipEvent_t lb = log->event();
            HIP_TRY(hipEventRecord(lb, s));
            log->recs.push_back(LaunchLog::Rec{la, lb, out});
        } else {
            HIP_TRY(hipEventRecord(b, s));
        }
    }
    // bytes are the launch's algorithmic bytes; without a log this waits for the launch
    void collect(double bytes) const {
        if (!out) return;
        out->bytes += bytes;
        if (log) return;
        HIP_TRY(hipEventSynchronize(b));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, a, b));
        out->ms += ms;
        out->launches += 1;
    }
};

// ---- launchers (kernels.hip) ---------------------------------------------------------------------
void launch_pad_rows(hipStream_t s, const float* src, int64_t n, int dim, float* dst, int64_t ld);
void launch_fill_heap(hipStream_t s, int32_t* idx, float* dist, uint8_t* flags, int64_t count);
void launch_fill_i32(hipStream_t s, int32_t* p, int32_t v, int64_t count);
// zero `bytes` bytes with a kernel (not a memset node, so the call can sit inside a captured graph)
void launch_zero(hipStream_t s, void* p, size_t bytes);
void launch_pair_distances(hipStream_t s, int metric, const float* X, int64_t ld, int dim,
                           const int32_t* a, const int32_t* b, int64_t n_pairs, float* out);

// forest, reference streams (depth-first per tree)
struct ForestDFSArgs {
    const float* X;
    int64_t ld;
    int dim;
    int64_t n;
    int64_t leaf_size;
    int64_t max_depth;
    int angular;
    int n_trees;
    const uint64_t* tree_states;  // 4 words per tree
    float** hyper;
    float** offsets;
    int32_t** children;
    int32_t** indices;
    int64_t* cap_nodes;
    int64_t* n_nodes_out;
    int32_t* status;
    // scratch, n_trees x (n+1) each
    uint8_t* side;
    int32_t* rankR;
    int32_t* PQ;
    int32_t* tmp;
    // Many independent point sets (queue non-null): tree t covers rows tree_base[t] .. tree_base[t] + tree_n[t]
    // of X (its indices are local to that range), and `grid` persistent workgroups pull the trees from a work
    // queue (*queue, zeroed) in the order order[] (largest first).  Workgroup g uses the scratch slot starting at
    // element g * slot_stride of side / rankR / tmp (2 x that of PQ); slot_stride > the largest tree_n.
    const int64_t* tree_base;
    const int64_t* tree_n;
    const int32_t* order;
    int32_t* queue;
    int64_t slot_stride;
    int grid;
    // Cooperative form (coop non-null): coop_groups workgroups per tree; the tree's first walks it and the others
    // join every node of at least coop_min points (margins and partition split into contiguous slices).
    struct CoopCtl* coop;   // one per tree, zeroed
    float* coop_hp;         // dim floats per tree: the node's hyperplane, published by the walking workgroup
    int coop_groups;
    int64_t coop_min;
};
// The control block of one tree's cooperating workgroups.  Every wait on it is bounded: after about four
// seconds without progress a workgroup sets `abort` and the tree reports status 3 (the caller reruns it alone).
constexpr int kCoopMaxGroups = 32;
struct CoopCtl {
    int seq, done, abort, bar;
    int start, end;
    float off;
    int pad;
    int cnt[kCoopMaxGroups];
    int tie[kCoopMaxGroups];
};
void launch_forest_dfs(hipStream_t s, const ForestDFSArgs& a);
// workgroups per tree for the cooperative forest kernel with n_trees trees (0: not all of them fit at once)
int forest_coop_groups(int dim, int n_trees);
// resident workgroups of the work-queue forest kernel on this device (the grid for a.queue launches)
int forest_queue_grid(int dim);

// forest, counter streams (level-synchronous across trees, forest_level.hip).  Trees must be
// allocated with cap_nodes each and zeroed hyperplanes/offsets; returns false when a tree needs
// more than cap_nodes nodes.
struct Workspace;
bool build_forest_level(hipStream_t s, const float* X, int64_t ld, int dim, int64_t n, int64_t leaf_size,
                        int64_t max_depth, int angular, const std::vector<uint64_t>& keys, rinnd_forest& F,
                        const KTimer& margins_timer, Workspace& W);

// leaves of one tree in depth-first order: (start, end) pairs, returns count via device int
void launch_leaf_list(hipStream_t s, const int32_t* children, int64_t n_nodes, int64_t n,
                      int32_t* leaf_mark /* n+1 scratch */, int64_t* leaf_rank /* n+1 */,
                      int32_t* leaf_end_at /* n */, int64_t* leaves /* 2 per leaf */,
                      void* scan_tmp, size_t scan_tmp_bytes);

// The same for a virtual tree of a batch: one entry per index that has tree number t, its positions placed
// at pbase .. pbase + n of a combined position space of n positions (vidx: the combined tree indices, ids
// offset by the index's first row).  max_span bounds every entry's max(n_nodes, n).
struct VTreeEnt {
    const int32_t* children;
    int64_t n_nodes;
    const int32_t* indices;
    int64_t pbase, base, n;
};
void launch_leaf_list_batch(hipStream_t s, const VTreeEnt* ents_dev, int n_ents, int64_t max_span, int64_t n,
                            int32_t* leaf_mark, int64_t* leaf_rank, int32_t* leaf_end_at, int64_t* leaves,
                            int32_t* vidx, void* scan_tmp, size_t scan_tmp_bytes);

// initialize_from_leaves fused per leaf (leaf_init.hip); n_leaves is read on the device,
// max_leaves bounds the grid.  Supported when a 64..256-thread block's heaps fit in LDS.
// ws: the leaf marks and start ranks of launch_leaf_list and a distance scratch of
// leaf_init_scratch_floats() floats (the two-pass kernels are skipped when it is null).
struct Bf16Rows;
struct LeafInitWs {
    const int32_t* mark;
    const int64_t* rank;
    float* D;
    unsigned long long* pairs;  // 2 counters: ordered pairs m(m-1) written whole, survivor pairs (accumulated)
    const Bf16Rows* exact_bf16; // exact-mode bf16 rows: pair distances on MFMA (null: FP32 kernels)
};
// Trees of a build whose leaf pairs are all written and pushed (the heaps start empty, so most pushes of the
// first trees are accepted); later trees write only the pairs below an endpoint's root when the MFMA survivor
// kernel applies.  At 1M x 128, leaf 256, 16 trees the BF16 build's leaf phase takes 10.95 / 10.69 / 10.68 /
// 10.86 / 11.09 ms with 2 / 3 / 4 / 5 / 6 such trees.
#ifndef RINND_LEAF_FULL_TREES
#define RINND_LEAF_FULL_TREES 3
#endif
constexpr int kLeafFullTrees = RINND_LEAF_FULL_TREES;
bool leaf_init_supported(int k, int64_t leaf_size);
int64_t leaf_init_scratch_floats(int64_t n, int64_t ld, int64_t leaf_size);
void launch_leaf_init(hipStream_t s, int metric, const float* X, int64_t ld, int dim, const int32_t* tree_indices,
                      const int64_t* leaves, const int64_t* n_leaves_dev, int64_t max_leaves, int64_t leaf_size,
                      int32_t* idx, float* dist, uint8_t* flags, int k, const LeafInitWs& ws,
                      const KTimer& dist_timer, const KTimer& push_timer,
                      bool early);  // early: one of the first trees of a build (narrower push tiles)

// Survivor trees in groups (exact-mode BF16 rows, squared Euclidean): every tree of a group writes its survivors
// against the roots at the group's start (a superset of the pushes the heaps accept: roots only fall), point-major;
// then one pass per point replays the group's trees in order.  Up to kSurvGroupMax trees per group: a larger group
// saves push passes but its later trees test against staler roots and write more survivors (at 1M x 128, leaf
// 256, 16 trees the 13 survivor trees' leaf phase takes 10.8 ms in groups of 2-4, 11.2 ms one by one, 11.6 ms in
// one group; with the current kernels 10.67 / 10.76 / 10.86 ms with groups of at most 4 / 5 / 7).
#ifndef RINND_SURV_GROUP_MAX
#define RINND_SURV_GROUP_MAX 4
#endif
constexpr int kSurvGroupMax = RINND_SURV_GROUP_MAX;
// Device memory the scratch of one group may take: an eighth of the device's memory, at most 40 GiB.
double surv_group_budget();
struct SurvGroup {
    const int32_t* tidx[kSurvGroupMax];  // tree g's indices
    float* D;                            // tree g: D + g * n * Dl4 (n x Dl4)
    uint32_t* Msk;                       // tree g: Msk + g * n * 8
    int2* info;                          // tree g: info + g * n
    int G;
};
// 0 if the survivor path does not apply, else the D row length (leaf columns) it uses
int64_t leaf_surv_dl4(int metric, const Bf16Rows* exact_bf16, int64_t ld, int64_t leaf_size, int k);
void launch_leaf_surv_tree(hipStream_t s, const Bf16Rows& R, const int32_t* tree_indices, const int64_t* leaves,
                           const int64_t* n_leaves_dev, int64_t n, int64_t leaf_size, const float* hdist, int k,
                           const SurvGroup& grp, int g, unsigned long long* surv, const KTimer& timer);
void launch_leaf_push_group(hipStream_t s, const SurvGroup& grp, const float* X, int64_t ld, int dim, int64_t n,
                            int64_t leaf_size, int32_t* idx, float* dist, uint8_t* flags, int k, const KTimer& timer);

// ordered-apply pipeline
void launch_leaf_pairs(hipStream_t s, int metric, const float* X, int64_t ld, int dim,
                       const int32_t* tree_indices, const int64_t* leaves, int64_t n_leaves,
                       const uint64_t* pair_base, const float* heap_dist, int k, URec* U,
                       unsigned long long* U_count, uint64_t U_cap, unsigned long long* cnt,
                       int32_t* overflow);
void launch_leaf_pair_base(hipStream_t s, const int64_t* leaves, int64_t n_leaves,
                           uint64_t* pair_base_in, uint64_t* pair_base_out, void* scan_tmp,
                           size_t scan_tmp_bytes);
void launch_scatter(hipStream_t s, const URec* U, const unsigned long long* U_count,
                    const uint64_t* off, unsigned long long* cur, SegRec* seg);
// big_list: n_dest ints, big_count: 1 int, bins: 128 counters (scratch of the grouping stages)
void launch_heap_apply(hipStream_t s, const uint64_t* off, SegRec* seg, SegRec* seg2,
                       int64_t n_dest, int32_t* idx, float* dist, uint8_t* flags, int k,
                       unsigned long long* changes, int32_t* big_list, int32_t* big_count, uint32_t* bins);

// candidates
void launch_gather_strided(hipStream_t s, const uint64_t* in, int64_t stride, int64_t n, int count, uint64_t* out);
void launch_row_valid_counts(hipStream_t s, const int32_t* idx, int64_t n, int k, uint64_t* cnt,
                             const uint8_t* row_on = nullptr);
struct RngStream {
    uint64_t s[4];
    int64_t count;
    int64_t out_offset;
};
void launch_rng_streams(hipStream_t s, const uint64_t* jump_tables, const RngStream* streams_dev,
                        int n_streams, int64_t max_count, float* out);
// Candidate sets (CandidateSets::build_from_graph, candidates.rs:51-351) with forward pushes read from
// the heap row and reverse pushes grouped by one stable radix sort; writes new_idx / old_idx
// (pre-filled with -1).  Scratch: fpr, rkey, rval, rkey2, rval2 (n*k each), cnt/off (n+1), order (n),
// bins (128), sort/scan temporaries.  The forward draws need C <= 64 lanes of LDS per destination.
struct CandBuildArgs {
    const int32_t* idx;
    uint8_t* flags;  // the row's new entries that became new candidates are flagged old here
    int64_t n;
    int k, C, mode;
    const uint64_t* rvp;  // exclusive prefix of valid edges per row (n+1)
    const float* prio;
    int64_t block_rows;
    const int64_t* block_prio_base;
    uint64_t counter_key;
    float* fpr;
    uint32_t *rkey, *rkey2;
    uint64_t *rval, *rval2;
    unsigned long long* cnt;
    uint64_t* off;
    int32_t* order;
    uint32_t* bins;
    void* sort_tmp;
    size_t sort_tmp_bytes;
    void* scan_tmp;
    size_t scan_tmp_bytes;
    int32_t *new_idx, *old_idx;
    // per row: pairs the join evaluates, nn(nn-1)/2 + nn*no (rows without candidates keep the caller's
    // zero); counts[0] += valid new+old candidates, counts[1] += pairs
    unsigned long long* row_pairs;
    unsigned long long* counts;
    const uint8_t* row_on;  // optional: rows whose byte is 0 take no part (finished indexes of a batch)
};
size_t cand_sort_tmp_bytes(int64_t items, int end_bit);
void launch_cand_build(hipStream_t s, const CandBuildArgs& a);

// local join.  Row r = i - row_begin owns the region U[rowoff[r], rowoff[r+1]) sized by its valid
// pair count; the join writes its kept pairs there as (slot t = j*2C + kk, d) in any order, the kept
// count to kc[r], and adds each candidate's share to cnt[id] (one atomic per candidate and row).
// A kept pair of a row's join: slot t = j * 2C + kk of the row's candidate pair (j, kk), with flags for the
// endpoints whose heaps it can still enter (d below that endpoint's root when the pair was evaluated; a
// push at or above the root is rejected, and the roots only fall until the records are applied).
struct JRec {
    uint32_t t;
    float d;
};
static_assert(sizeof(JRec) == 8, "JRec layout");
constexpr uint32_t kEmitP = 1u << 31, kEmitQ = 1u << 30, kSlotMask = kEmitQ - 1u;
struct JoinOut {
    const uint64_t* rowoff;
    JRec* U;
    unsigned long long* kc;    // kept pairs per row
    unsigned long long* krec;  // records per row
    unsigned long long* kept;  // kept pairs in all, spread over kKeptSlots counters (summed by the caller)
};
constexpr int kKeptSlots = 256;
// shape: a tile shape (threads per row, rows per workgroup) of join_fp32_shape_info; 0 is the default
void launch_join_fp32(hipStream_t s, int metric, const float* X, int64_t ld, int dim,
                      const int32_t* new_idx, const int32_t* old_idx, int C, int64_t row_begin,
                      int64_t row_end, const float* heap_dist, int k, const JoinOut& out, int shape = 0);
int join_fp32_shape_count();
void join_fp32_shape_info(int shape, int* tpr, int* rpg);
// rows [0, n): times the shapes on slices of the rows, joins the rest with the fastest and returns it
// Many independent builds in one pass over the device (batch.hip): index s owns rows off[s] .. off[s+1] of DX
// and the heap rows of the same numbers; its heap ids come back in its own numbering.
void nn_descent_batch_dev(rinnd_ctx* ctx, const float* DX, int64_t ld, const std::vector<int64_t>& off, int dim,
                          int metric, const rinnd_params& p, const std::vector<int64_t>& n_trees,
                          const std::vector<int64_t>& n_iters, std::vector<Xoshiro>& mains, int32_t* idx,
                          float* dist, uint8_t* flags, bool use_graph, std::vector<int32_t>& iterations,
                          std::vector<int32_t>& converged, rinnd_batch_stats& st);
void join_fp32_cached(rinnd_ctx* ctx, hipStream_t s, int metric, const float* X, int64_t ld, int dim,
                      const int32_t* new_idx, const int32_t* old_idx, int C, int64_t n, const float* heap_dist, int k,
                      const JoinOut& out);
int launch_join_fp32_tuned(hipStream_t s, int metric, const float* X, int64_t ld, int dim, const int32_t* new_idx,
                           const int32_t* old_idx, int C, int64_t n, const float* heap_dist, int k, const JoinOut& out,
                           bool* tuned);
// Kept pairs -> destination-grouped (id, distance) records for the ordered heap update.  Each row's
// kept pairs are laid out in loop order (slot t ascending) as two records each - (p; q, d) and
// (q; p, d) - at 2 * koff[r] (koff: exclusive prefix of kc), then one stable radix sort by destination
// leaves every destination's records in generation order; off = prefix of the per-destination counts
// (cnt).  Then one lane per destination applies them (checked_flagged_push, flag new).
struct JoinApplyArgs {
    const int32_t* new_idx;
    const int32_t* old_idx;
    int C;
    int64_t n;
    JoinOut out;
    uint64_t* koff;          // n + 1: exclusive prefix of the records per row
    uint64_t n_records;      // koff[n]
    uint32_t *rkey, *rkey2;  // n_records each
    uint64_t *rval, *rval2;
    uint64_t* off;           // n + 1: each destination's segment of the sorted records (written here)
    int32_t* order;          // n
    uint32_t* bins;          // 128
    void* sort_tmp;
    size_t sort_tmp_bytes;
    void* scan_tmp;
    size_t scan_tmp_bytes;
    int32_t* idx;
    float* dist;
    uint8_t* flags;
    int k;
    unsigned long long* changes;
    uint32_t* row_changes;  // optional: accepted pushes per destination (rows without records keep their value)
};
size_t join_sort_tmp_bytes(int64_t items, int end_bit);
// records of the kept pairs, sorted by destination (off: each destination's segment, found in the sorted
// keys), applied in generation order
void launch_join_apply(hipStream_t s, const JoinApplyArgs& a);

// local join, BF16 rows on MFMA with FP32 re-evaluation (join_bf16.hip)
struct Bf16Rows {
    DBuf<uint16_t> Xb;  // n x ldb bf16 (centred unless exact)
    DBuf<float> nrm;    // FP32 squared norms of the rounded rows
    DBuf<float> err;    // per row, an upper bound of |bf16(x - mu) - (x - mu)| (non-exact rows)
    DBuf<float> mu;     // column means used for centring
    DBuf<double> aux;   // [flag word | max norm | column sums] of the preparation pass
    int64_t ldb = 0;
    int exact = 0;      // integers with |x| <= 256 and |x|^2 <= 2^22 (2^23 if non-negative): exact screen
};
void prepare_bf16_rows(hipStream_t s, const float* X, int64_t ld, int64_t n, int dim, Bf16Rows& R);


// leaf pair distances on MFMA (exact mode only), same D layout as the FP32 two-pass kernel
size_t leaf_dist_mfma_lds(int64_t ldb, int mcap);
size_t leaf_surv_mfma_lds(int64_t ldb, int mcap);
// survivor pairs only (pairs below an endpoint's heap root when the kernel runs), point-major: per point a
// column mask (8 words), the surviving distances at their leaf columns of its D row, and its leaf's
// (start, size) in info
void launch_leaf_surv_mfma(hipStream_t s, const Bf16Rows& R, const int32_t* tree_indices, const int64_t* leaves,
                           const int64_t* n_leaves_dev, int64_t max_leaves, int mcap, const float* hdist, int k,
                           float* D, int64_t Dl4, uint32_t* Msk, int2* info, unsigned long long* surv);
void launch_leaf_dist_mfma(hipStream_t s, const Bf16Rows& R, const int32_t* tree_indices, const int64_t* leaves,
                           const int64_t* n_leaves_dev, int64_t max_leaves, int mcap, float* D, int64_t Dl4,
                           unsigned long long* pairs);
void launch_join_bf16(hipStream_t s, const float* X, int64_t ld, int dim, const Bf16Rows& R,
                      const int32_t* new_idx, const int32_t* old_idx, int C, int64_t row_begin,
                      int64_t row_end, const float* heap_dist, int k, const JoinOut& out,
                      unsigned long long* rescored);  // kKeptSlots counters

// relabelling: Xr[r] = X[perm[r]], inv[perm[r]] = r; ids[i] = map[ids[i]] (negative ids kept);
// heap row r -> caller row perm[r] with ids mapped through perm
void launch_relabel_points(hipStream_t s, const float* X, int64_t ld, int64_t n, const int32_t* perm, int32_t* inv,
                           float* Xr);
void launch_map_ids_trees(hipStream_t s, int32_t* const* ids_dev, int n_trees, int64_t count, const int32_t* map);
void launch_unpermute_heap(hipStream_t s, const int32_t* idx, const float* dist, const uint8_t* flags,
                           const int32_t* perm, int64_t n, int k, int32_t* oidx, float* odist, uint8_t* oflags);

// rows
void launch_sort_rows(hipStream_t s, int32_t* idx, float* dist, uint8_t* flags, int64_t n, int k);
void launch_correct(hipStream_t s, float* dist, int64_t count, int corr);

// scans (hipcub)
size_t scan_u64_tmp_bytes(int64_t n);
void scan_u64_exclusive(hipStream_t s, const unsigned long long* in, uint64_t* out, int64_t n,
                        void* tmp, size_t tmp_bytes);

// Grouping buffers of the ordered-apply pipeline (grow-only).
struct Pipeline {
    DBuf<URec> U;
    DBuf<unsigned long long> U_count;
    DBuf<unsigned long long> cnt, cur;
    DBuf<uint64_t> off;
    DBuf<SegRec> seg, seg2;
    DBuf<int32_t> big_list, big_count, overflow;
    DBuf<uint32_t> bins;
    DBuf<unsigned long long> changes;
    DBuf<char> scan_tmp;
    size_t scan_bytes = 0;
    int64_t n = 0;
    void init(int64_t n_, uint64_t u_cap) {
        n = n_;
        U.ensure(std::max<uint64_t>(u_cap, 1));
        seg.ensure(2 * std::max<uint64_t>(u_cap, 1));
        seg2.ensure(2 * std::max<uint64_t>(u_cap, 1));
        U_count.ensure(1);
        cnt.ensure(n + 1);
        cur.ensure(n + 1);
        off.ensure(n + 1);
        big_list.ensure(std::max<int64_t>(n, 1));
        big_count.ensure(1);
        bins.ensure(128);
        overflow.ensure(1);
        changes.ensure(1);
        size_t b = scan_u64_tmp_bytes(n + 1);
        if (b > scan_bytes) {
            scan_tmp.alloc(b);
            scan_bytes = b;
        }
    }
    void begin(hipStream_t s) {
        HIP_TRY(hipMemsetAsync(cnt.p, 0, (n + 1) * 8, s));
        HIP_TRY(hipMemsetAsync(U_count.p, 0, 8, s));
        HIP_TRY(hipMemsetAsync(overflow.p, 0, 4, s));
    }
    // counts -> offsets -> grouped records
    void group(hipStream_t s) {
        scan_u64_exclusive(s, cnt.p, off.p, n + 1, scan_tmp.p, scan_bytes);
        HIP_TRY(hipMemsetAsync(cur.p, 0, (n + 1) * 8, s));
        launch_scatter(s, U.p, U_count.p, off.p, cur.p, seg.p);
    }
    void check_overflow(hipStream_t s) {
        int32_t ov = 0;
        HIP_TRY(hipMemcpyAsync(&ov, overflow.p, 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (ov) throw Error{RINND_ERR_INTERNAL, "update buffer overflow"};
    }
};

// Grow-only device buffers a context keeps between builds: allocating and releasing the multi-GB
// scratch of a build would otherwise stall every call.  Each use site owns a fixed slot; contents are
// never assumed to survive from one build to the next.
// Pinned host memory for small device -> host status reads that run while later kernels are queued.
struct PinnedBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void* ensure(size_t b) {
        if (b > bytes) {
            if (p) HIP_TRY(hipHostFree(p));
            p = nullptr;
            HIP_TRY(hipHostMalloc(&p, b, hipHostMallocDefault));
            bytes = b;
        }
        return p;
    }
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
};

// An event that is created on first use and destroyed with its owner.
struct OwnedEvent {
    hipEvent_t e = nullptr;
    hipEvent_t get() {
        if (!e) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        return e;
    }
    OwnedEvent() = default;
    OwnedEvent(const OwnedEvent&) = delete;
    OwnedEvent& operator=(const OwnedEvent&) = delete;
    ~OwnedEvent() {
        if (e) (void)hipEventDestroy(e);
    }
};

struct Workspace {
    std::vector<DBuf<char>> slots;
    Bf16Rows b16;
    std::unique_ptr<Pipeline> pipeline;
    PinnedBuf status;      // device -> host status words
    PinnedBuf stage;       // host -> device parameter staging (candidate streams, block bases)
    OwnedEvent status_ev;  // marks the status copy
    void* get(int slot, size_t bytes) {
        if ((int)slots.size() <= slot) slots.resize(slot + 1);
        slots[slot].ensure(std::max<size_t>(bytes, 16));
        return slots[slot].p;
    }
};
template <class T>
struct WBuf {
    T* p = nullptr;
    size_t n = 0;
    WBuf() = default;
    WBuf(Workspace& w, int slot, size_t count) : p((T*)w.get(slot, count * sizeof(T))), n(count) {}
    size_t bytes() const { return n * sizeof(T); }
};
// slot ranges: NN-Descent build 0..59, caller-facing heap staging 60..62, level forest 100..149
enum : int { kWsHeap = 60, kWsForest = 100 };

}  // namespace rinnd
