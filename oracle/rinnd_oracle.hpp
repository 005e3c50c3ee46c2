// rinnd_oracle.hpp - CPU restatement of lmcinnes/rinnd's NN-Descent build and greedy search.
//
// TEST INFRASTRUCTURE.  This is the checker the GPU library is compared against: only tests/,
// __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it.  Nothing in rinnd_amd/
// links or calls it.
//
// Every function restates the reference algorithm it cites (paths relative to the rinnd repository,
// "core" = crates/rinnd-core/src).  Floating-point expressions keep the reference's operation order:
// the file is compiled with -mavx2 -mfma -ffp-contract=off and the AVX2 kernels use the same
// intrinsics as the reference, so results are bit-identical to a reference run on an AVX2+FMA host.
#pragma once

#include <cstdint>
#include <cstddef>
#include <functional>
#include <vector>
#include <utility>

namespace oracle {

// ---- FastRng: Xoshiro256PlusPlus seeded with SplitMix64 (core/rng.rs:15-80; rand_xoshiro 0.6.0,
//      rand 0.8.5 conversions) --------------------------------------------------------------------
struct FastRng {
    uint64_t s[4];
    FastRng() : FastRng(42) {}
    explicit FastRng(uint64_t seed);
    uint64_t next_u64();
    uint32_t next_u32() { return (uint32_t)(next_u64() >> 32); }
    float next_float();
    bool next_bool() { return (int32_t)next_u32() < 0; }
    uint64_t next_index(uint64_t n);
};

// ---- distances (core/distance/*.rs) -------------------------------------------------------------
enum Metric : int {
    SQEUCLIDEAN = 0, COSINE = 1, INNER_PRODUCT = 2, ALT_DOT = 3, DIRECT_COSINE = 4, ALT_COSINE = 5,
    ALT_INNER_PRODUCT = 6
};
enum Correction : int { CORR_NONE = 0, CORR_SQRT = 1, CORR_ALT_COSINE = 2, CORR_ALT_IP = 3 };

float distance(int metric, const float* a, const float* b, size_t dim);
float correct(int corr, float d);
float tree_dot(const float* a, const float* b, size_t dim);  // builder.rs:29-60

// ---- NeighborHeap (core/heap/neighbor_heap.rs) ------------------------------------------------
struct NeighborHeap {
    std::vector<int32_t> indices;
    std::vector<float> distances;
    std::vector<uint8_t> flags;
    size_t n_points = 0, k = 0;
    NeighborHeap() = default;
    NeighborHeap(size_t n, size_t k_);
    float max_distance(size_t p) const { return distances[p * k]; }
    bool checked_flagged_push(size_t p, int32_t nb, float d, bool is_new);
    bool simple_push(size_t p, int32_t nb, float d);
    void unchecked_flagged_push(size_t p, int32_t nb, float d, bool is_new);
    void sift_down(size_t off, size_t start);
    void sort_all();
};

// ---- FlatTree + forest (core/tree) --------------------------------------------------------------
struct FlatTree {
    std::vector<float> hyperplanes;  // n_nodes x dim
    std::vector<float> offsets;
    std::vector<int32_t> children;   // n_nodes x 2
    std::vector<int32_t> indices;
    size_t dim = 0, n_nodes = 0;
    std::vector<std::pair<size_t, size_t>> get_all_leaves() const;
    size_t search_leaves(const float* point, FastRng& rng, size_t leaf_budget,
                         std::vector<std::pair<size_t, size_t>>& leaves) const;
};

// QuantizedFlatTree (flat_tree.rs:35-183): int8 hyperplanes with one symmetric scale per node; offsets, topology and
// leaves as in the FP32 tree.
struct QuantizedFlatTree {
    std::vector<int8_t> hyperplanes;  // n_nodes x dim
    std::vector<float> scales, offsets;
    std::vector<int32_t> children, indices;
    size_t dim = 0, n_nodes = 0;
    static QuantizedFlatTree from_flat_tree(const FlatTree& t);
    float margin(size_t node, const int8_t* query, float query_scale) const;
    size_t search_leaves(const int8_t* query, float query_scale, FastRng& rng, size_t leaf_budget,
                         std::vector<std::pair<size_t, size_t>>& leaves) const;
};
// quantized_i8_dot (core/distance/quantized.rs:26-40)
int32_t quantized_i8_dot(const int8_t* x, const int8_t* y, size_t n);
// the per-vector symmetric int8 query of a quantized index (crates/rinnd/src/lib.rs:303-347): codes
// round(x * 127 / max|x|) clamped to [-127, 127], returned scale max|x| / 127 (0 for a zero query)
float quantize_query_i8(const float* q, size_t dim, int8_t* codes);

FlatTree build_rp_tree(const float* data, size_t n, size_t dim, size_t leaf_size, FastRng& rng,
                       bool angular, size_t max_depth);
std::vector<FlatTree> build_rp_forest(const float* data, size_t n, size_t dim, size_t n_trees,
                                      size_t leaf_size, FastRng& rng, bool angular,
                                      size_t max_depth, int threads);

// ---- NN-Descent (core/nndescent) ----------------------------------------------------------------
struct Params {
    size_t n_neighbors = 30, n_trees = 8, leaf_size = 64, max_candidates = 60, n_iters = 10;
    float delta = 0.001f;
    bool angular = false;
    size_t max_depth = 200;
};

struct CandidateSets {
    std::vector<int32_t> new_indices, old_indices;
    size_t n_vertices = 0, max_candidates = 0;
};

void initialize_from_leaves(NeighborHeap& g, const std::vector<std::vector<int32_t>>& leaves,
                            const float* data, size_t dim, int metric, int threads);
CandidateSets build_candidates(NeighborHeap& g, size_t max_candidates, FastRng& rng, int threads);
size_t update_iteration(NeighborHeap& g, const CandidateSets& c, const float* data, size_t dim,
                        int metric, int threads);
struct BuildTrace {
    std::vector<size_t> changes;
    double ms_forest = 0, ms_init = 0, ms_candidates = 0, ms_updates = 0;
};
NeighborHeap nn_descent(const float* data, size_t n, size_t dim, int metric, const Params& p,
                        FastRng& rng, int threads, std::vector<FlatTree>* forest_out,
                        BuildTrace* trace);

// ---- search graph (core/graph/search_graph.rs) --------------------------------------------------
struct SearchGraph {
    std::vector<int32_t> indptr, indices;
    size_t n_vertices = 0;
    void reorder(const std::vector<size_t>& order);
};
SearchGraph from_dense_diversified(const int32_t* idx, const float* dist, const float* data,
                                   size_t n, size_t k, size_t dim, int metric, float prob,
                                   float mult, int threads, float* min_distance_out);
SearchGraph weighted_union_pruned(const std::vector<std::vector<std::pair<int32_t, float>>>& fwd,
                                  const std::vector<std::vector<std::pair<int32_t, float>>>& rev,
                                  size_t n, size_t max_degree);

// ---- index (core/index.rs) ----------------------------------------------------------------------
struct IndexParams {
    size_t n_neighbors = 30, n_trees = 8, n_search_trees = 1, search_tree_leaf_budget = 1;
    long leaf_size = -1, max_candidates = -1, n_iters = -1;
    float delta = 0.001f, diversify_prob = 1.0f, pruning_degree_multiplier = 1.5f;
    bool angular = false;
    uint64_t random_seed = 42;
};

// ---- quantized search representations (crates/rinnd/src/lib.rs:133-398) ---------------------------
// int8 (per-vector cosine), sq8u (global symmetric) and sq8p (per-vector symmetric) codes of the rows.
enum I8Encoding : int { I8_PER_VECTOR_COSINE = 0, I8_GLOBAL_SYMMETRIC = 1, I8_PER_VECTOR_SYMMETRIC = 2 };
struct I8Codes {
    int encoding = I8_PER_VECTOR_COSINE;
    size_t dim = 0;
    std::vector<int8_t> codes;     // n x dim
    std::vector<float> inv_norms;  // per row: 1/|codes| (cosine), the global or the row's dequantisation scale
    float global_multiplier = 0, global_dequant_scale = 0;
};
I8Codes encode_i8(const float* data, size_t n, size_t dim, int encoding);
float quantize_query_codes(const I8Codes& c, const float* q, int8_t* codes);
float i8_cosine_distance(const int8_t* q, float q_inv_norm, const int8_t* x, float x_inv_norm, size_t dim);

struct Index {
    std::vector<float> data;  // reordered
    size_t n_points = 0, dim = 0, n_neighbors = 0;
    int metric = SQEUCLIDEAN, corr = CORR_NONE;
    std::vector<int32_t> neighbor_indices;
    std::vector<float> neighbor_distances;
    SearchGraph search_graph;
    std::vector<FlatTree> search_trees;
    std::vector<QuantizedFlatTree> quantized_trees;  // non-empty: the search descends them (tree_quantization int8)
    size_t search_tree_leaf_budget = 1;
    std::vector<size_t> vertex_order;
    float min_distance = 0;
    uint64_t rng_seed = 42;
    BuildTrace trace;
    void query(const float* queries, size_t nq, size_t k, float eps, int32_t* out_idx,
               float* out_dist, int threads) const;
    void query_one(const float* q, size_t query_id, size_t k, float eps, int32_t* out_idx,
                   float* out_dist) const;
    // query_one_quantized_i8_with_workspace (core/index.rs:283-330) over the codes c of this index's rows
    void query_one_quantized(const I8Codes& c, const float* q, size_t query_id, size_t k, size_t candidate_width,
                             size_t rerank_width, float eps, int32_t* out_idx, float* out_dist) const;
    void seed_ids(const float* q, const int8_t* tree_codes, float tree_scale, FastRng& rng,
                  std::vector<int32_t>& ids) const;
};

Index build_index(const float* data, size_t n, size_t dim, int metric, int corr,
                  const IndexParams& p, int threads);

void greedy_from_seeds(const float* q, const float* data, size_t dim, int metric, size_t n_vertices,
                       const int32_t* indptr, const int32_t* indices, const std::vector<int32_t>& seed_ids,
                       size_t k, float eps, float min_distance, FastRng& rng, std::vector<int32_t>& ri,
                       std::vector<float>& rd);
void greedy_core(const std::function<float(int32_t)>& dist_to, size_t n_vertices, const int32_t* indptr,
                 const int32_t* indices, const std::vector<int32_t>& seed_ids, size_t k, float eps,
                 float min_distance, FastRng& rng, std::vector<int32_t>& ri, std::vector<float>& rd);

// std::collections::BinaryHeap<(OrderedFloat<f32>, i32)> as used by BoundedHeap
// (core/heap/candidate_heap.rs:78-163); the layout decides tie order in into_sorted.
struct BoundedHeap {
    std::vector<std::pair<float, int32_t>> data;
    size_t k = 0;
    explicit BoundedHeap(size_t k_) : k(k_) {}
    bool push(float d, int32_t idx);
    float max_distance() const;
    bool full() const { return data.size() >= k; }
    void into_sorted(std::vector<int32_t>& idx, std::vector<float>& dist) const;
    // std BinaryHeap primitives (Rust 1.75 liballoc/collections/binary_heap)
    void std_push(std::pair<float, int32_t> item);
    std::pair<float, int32_t> std_pop();
};

}  // namespace oracle
