/*
 * rinnd_hip.h - C ABI of the MI355X (gfx950) NN-Descent k-NN graph builder and greedy graph search.
 *
 * Every entry point replaces one seam of lmcinnes/rinnd's Rust API.  The reference interface each
 * function stands in for is cited as `file:line` (paths relative to the rinnd repository root,
 * "core" = crates/rinnd-core/src, "py" = crates/rinnd/src/lib.rs).
 *
 * Conventions
 *  - Plain pointers and sizes only.  Array arguments named `data`, `queries` and `out_*` may be host
 *    pointers or HIP device pointers; the library detects which (hipPointerGetAttributes) and stages
 *    host buffers through device memory.  Device-resident inputs avoid every host copy.
 *  - Row-major layouts: data is n x dim float32, neighbour arrays are n x k.
 *  - Every call returns an rinnd_status; on failure rinnd_last_error() holds a message.  The library
 *    never aborts the host process (the reference panics, core/index.rs:65-76, 687-704).
 *  - A context owns one HIP stream on one device.  Calls on one context are serialised by an internal
 *    mutex; use one context per host thread for concurrency.
 */
#ifndef RINND_HIP_H
#define RINND_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rinnd_status {
    RINND_OK = 0,
    RINND_ERR_INVALID_ARG = 1, /* py: PyValueError paths, py:643-709, 787-792 */
    RINND_ERR_OOM = 2,
    RINND_ERR_HIP = 3,
    RINND_ERR_UNSUPPORTED = 4,
    RINND_ERR_INTERNAL = 5
} rinnd_status;

/* Distance functors of the reference (core/distance/traits.rs:7-47 `Distance<f32>`).  The GPU kernels
 * are compiled per functor instead of dispatching through a trait object. */
typedef enum rinnd_distance {
    RINND_DIST_SQEUCLIDEAN = 0,       /* SquaredEuclidean        core/distance/euclidean.rs:14-33, 66-110 */
    RINND_DIST_COSINE = 1,            /* Cosine                  core/distance/cosine.rs:44-118 */
    RINND_DIST_INNER_PRODUCT = 2,     /* InnerProduct            core/distance/inner_product.rs:64-107 */
    RINND_DIST_ALT_DOT = 3,           /* AlternativeDot          core/distance/alternatives.rs:39-124 */
    RINND_DIST_DIRECT_COSINE = 4,     /* DirectNormalizedCosine  core/distance/alternatives.rs:131-154 */
    RINND_DIST_ALT_COSINE = 5,        /* AlternativeCosine       core/distance/alternatives.rs:200-231 */
    RINND_DIST_ALT_INNER_PRODUCT = 6  /* AlternativeInnerProduct core/distance/alternatives.rs:160-193 */
} rinnd_distance;

/* Output distance corrections (core/index.rs:669-673, core/distance/alternatives.rs:276-288). */
typedef enum rinnd_correction {
    RINND_CORR_NONE = 0,
    RINND_CORR_SQRT = 1,             /* euclidean: d.sqrt()                     py:1241-1243 */
    RINND_CORR_ALT_COSINE = 2,       /* 1 - 2^(-d)  correct_alternative_cosine   alternatives.rs:276-278 */
    RINND_CORR_ALT_INNER_PRODUCT = 3 /* -1/d        correct_alternative_inner_product alternatives.rs:282-288 */
} rinnd_correction;

/* Where the random draws of RP-tree splits and candidate priorities come from. */
typedef enum rinnd_rng_mode {
    /* The reference's FastRng streams (core/rng.rs:15-80): split choices consumed in depth-first order
     * per tree (core/tree/builder.rs:172-234) and candidate priorities per Rayon thread block
     * (core/nndescent/candidates.rs:84-131, 170-213).  Results are bit-identical to the reference on
     * the FP32 path; the forest is then built depth-first, one workgroup per tree. */
    RINND_RNG_REFERENCE = 0,
    /* Counter-based draws keyed by (seed, tree, node) and (seed, iteration, edge): the forest is built
     * level-synchronously across all trees and candidate sampling needs no stream jump-ahead.  Same
     * algorithm, different random streams. */
    RINND_RNG_COUNTER = 1
} rinnd_rng_mode;

/* Arithmetic of the local-join distance tiles. */
typedef enum rinnd_join_precision {
    RINND_JOIN_FP32 = 0, /* every candidate pair in FP32, the reference's AVX2 reduction order */
    /* BF16 rows on MFMA (||x||^2 + ||y||^2 - 2 x.y with FP32 accumulation) screen every pair; a pair
     * whose screened distance lies within the proven error bound of its threshold is re-evaluated in
     * FP32 (reference order), so stored distances and indices equal the FP32 path's. */
    RINND_JOIN_BF16 = 1
} rinnd_join_precision;

/* Mirrors NNDescentParams (core/nndescent/mod.rs:21-40) plus the knobs the reference reads implicitly. */
typedef struct rinnd_params {
    int64_t n_neighbors;    /* k */
    int64_t n_trees;        /* RP trees for initialisation */
    int64_t leaf_size;      /* maximum leaf size */
    int64_t max_candidates; /* effective value is min(max_candidates, 60, k)  mod.rs:89 */
    int64_t n_iters;        /* maximum NN-Descent iterations */
    float delta;            /* convergence: stop when changes <= (delta*k*n) as usize  mod.rs:157 */
    int32_t angular;        /* angular (cosine/IP) hyperplanes  builder.rs:281-308 */
    int64_t max_depth;      /* maximum tree depth (200 in the reference builder, index.rs:635) */
    int32_t verbose;        /* print per-phase timings like mod.rs:169-194 */
    int32_t n_threads;      /* rayon::current_num_threads() of the reference run being reproduced;
                               only candidate sampling depends on it (candidates.rs:57-61, 171) */
    int32_t rng_mode;       /* rinnd_rng_mode */
    int32_t join_precision; /* rinnd_join_precision */
} rinnd_params;

/* Launch timing of one kernel family: summed duration of its launches (HIP events recorded on the
 * stream the kernel runs on) and the algorithmic bytes those launches must move. */
typedef struct rinnd_kernel_timing {
    double ms;
    int64_t launches;
    double bytes;
} rinnd_kernel_timing;

typedef enum rinnd_timed_kernel {
    RINND_TIMED_JOIN = 0,      /* local join (FP32 or BF16 tile kernel), one launch per iteration */
    RINND_TIMED_MARGINS = 1,   /* RP-forest hyperplane side tests, one launch per tree level */
    RINND_TIMED_LEAF_DIST = 2, /* leaf pair distances, one launch per tree */
    RINND_TIMED_LEAF_PUSH = 3, /* leaf heap pushes, one launch per tree */
    RINND_TIMED_COUNT = 4
} rinnd_timed_kernel;

/* Per-build counters and phase timings (HIP events on the context stream). */
typedef struct rinnd_build_stats {
    int32_t iterations;          /* NN-Descent iterations run */
    int32_t converged;           /* 1 if stopped by the delta test */
    int64_t changes[64];         /* successful heap pushes per iteration (first 64) */
    int64_t join_pairs[64];      /* candidate pairs evaluated per iteration */
    int64_t join_rescored[64];   /* pairs re-evaluated in FP32 by the BF16 join */
    int64_t join_updates[64];    /* updates that passed the threshold filter */
    double ms_forest;            /* RP forest */
    double ms_init;              /* leaf initialisation */
    double ms_candidates;        /* candidate sampling, all iterations */
    double ms_join;              /* local join (distance tiles + filter), all iterations */
    double ms_apply;             /* ordered heap updates, all iterations */
    double ms_total;             /* whole nn_descent call */
    rinnd_kernel_timing timed[RINND_TIMED_COUNT];  /* per kernel family, see rinnd_timed_kernel */
} rinnd_build_stats;

typedef struct rinnd_ctx rinnd_ctx;
typedef struct rinnd_forest rinnd_forest;
typedef struct rinnd_index rinnd_index;

/* ---- context ----------------------------------------------------------------------------------- */
/* No reference counterpart (the reference runs on the implicit Rayon pool, core/nndescent/mod.rs:17). */
rinnd_status rinnd_ctx_create(int device, rinnd_ctx** out);
void rinnd_ctx_destroy(rinnd_ctx* ctx);
const char* rinnd_last_error(const rinnd_ctx* ctx);
/* Library version string (py: `version()`, crates/rinnd/src/lib.rs:1300-1305). */
const char* rinnd_version(void);

/* ---- reference RNG ----------------------------------------------------------------------------- */
/* FastRng::new(seed) state (core/rng.rs:22-26: Xoshiro256PlusPlus::seed_from_u64). */
void rinnd_rng_seed(uint64_t seed, uint64_t state_out[4]);
/* Advance a FastRng state by `steps` next_u64 draws in O(log steps) (GF(2) jump-ahead). */
void rinnd_rng_jump(uint64_t state[4], uint64_t steps);

/* ---- distances --------------------------------------------------------------------------------- */
/* out[p] = distance(data[a[p]], data[b[p]]) for p < n_pairs, bit-exact to the reference's AVX2
 * kernel of `dist` (core/distance/euclidean.rs:66-110 etc.).  Parity probe for every functor. */
rinnd_status rinnd_pair_distances(rinnd_ctx* ctx, rinnd_distance dist, const float* data, int64_t n,
                                  int64_t dim, const int32_t* a, const int32_t* b, int64_t n_pairs,
                                  float* out);

/* ---- NN-Descent build -------------------------------------------------------------------------- */
/* Replaces `nn_descent(data, n_points, dim, &distance, &params, &mut rng) -> (NeighborHeap,
 * Vec<FlatTree>)` (core/nndescent/mod.rs:79-197).  `rng_state` is the caller's FastRng, consumed and
 * advanced exactly as the reference does (forest seeds first, then each iteration's candidates).
 * Outputs are the unsorted heap arrays (n x k): indices (-1 = empty), distances (+inf = empty),
 * flags (1 = new).  `out_forest` (may be NULL) receives the RP forest; free with rinnd_forest_free. */
rinnd_status rinnd_nn_descent(rinnd_ctx* ctx, const float* data, int64_t n, int64_t dim,
                              rinnd_distance dist, const rinnd_params* params, uint64_t rng_state[4],
                              int32_t* out_idx, float* out_dist, uint8_t* out_flags,
                              rinnd_forest** out_forest, rinnd_build_stats* stats);

/* Counters and phase timings of a batched build (HIP events on the context stream). */
typedef struct rinnd_batch_stats {
    int32_t passes;                 /* passes over the batch: the largest per-index iteration count */
    int32_t graph_replays;          /* passes whose candidate phase and local join replayed the captured graph */
    int32_t graph_memsets_replaced; /* memset nodes of the captured graph replaced by fill kernels */
    int32_t join_threads_per_row;   /* FP32 join tile shape: threads per row ... */
    int32_t join_rows_per_group;    /* ... and rows per workgroup (tuned on the first pass) */
    int64_t changes[64];            /* heap changes per pass, all indexes */
    int64_t join_pairs[64];         /* candidate pairs evaluated per pass, all indexes */
    double ms_forest;               /* every index's RP forest (one work-queue launch) */
    double ms_init;                 /* leaf initialisation */
    double ms_candidates_join;      /* candidate phase + local join, all passes */
    double ms_apply;                /* ordered heap updates + convergence tests, all passes */
    double ms_total;                /* whole call */
} rinnd_batch_stats;

/* rinnd_nn_descent_batch flags */
#define RINND_BATCH_NO_GRAPH 1 /* run every pass eagerly instead of replaying the captured graph */

/* Many independent builds in one call.  The reference builds one index per call (core/index.rs:601-733 calls
 * nn_descent, core/nndescent/mod.rs:79-197, once per index); this runs n_indexes of those builds together.
 * Index s owns rows offsets[s] .. offsets[s+1] of `data` (n_total x dim, n_total = offsets[n_indexes],
 * offsets[0] = 0, non-decreasing) and rows of the same numbers of the outputs (n_total x k, ids in the
 * index's own numbering 0 .. n_s - 1).  Per index: n_trees[s] and n_iters[s] (NULL: params->n_trees /
 * params->n_iters for all), its FastRng state rng_states[4s .. 4s+3] (consumed and advanced), and in
 * out_iterations / out_converged (may be NULL) its iteration count and whether the delta test stopped it.
 * Shared: k, leaf_size, max_candidates, delta, angular, max_depth.  Each index's outputs equal those of
 * rinnd_nn_descent on its rows alone with rng_mode RINND_RNG_REFERENCE and n_threads 1, which the batch
 * requires (and the FP32 join); other settings return RINND_ERR_UNSUPPORTED. */
rinnd_status rinnd_nn_descent_batch(rinnd_ctx* ctx, const float* data, const int64_t* offsets, int64_t n_indexes,
                                    int64_t dim, rinnd_distance dist, const rinnd_params* params,
                                    const int64_t* n_trees, const int64_t* n_iters, uint64_t* rng_states,
                                    int32_t* out_idx, float* out_dist, uint8_t* out_flags, int32_t* out_iterations,
                                    int32_t* out_converged, int32_t flags, rinnd_batch_stats* stats);

/* RP forest built alone: `build_rp_forest` (core/tree/builder.rs:89-118). */
rinnd_status rinnd_build_forest(rinnd_ctx* ctx, const float* data, int64_t n, int64_t dim,
                                int64_t n_trees, int64_t leaf_size, int32_t angular,
                                int64_t max_depth, int32_t rng_mode, uint64_t rng_state[4],
                                rinnd_forest** out_forest);

/* FlatTree (core/tree/flat_tree.rs:14-28) accessors; copies go to host buffers. */
int64_t rinnd_forest_n_trees(const rinnd_forest* f);
rinnd_status rinnd_forest_tree_info(const rinnd_forest* f, int64_t tree, int64_t* n_nodes,
                                    int64_t* dim);
/* hyperplanes n_nodes x dim, offsets n_nodes, children n_nodes x 2, indices n (any may be NULL). */
rinnd_status rinnd_forest_tree_copy(const rinnd_forest* f, int64_t tree, float* hyperplanes,
                                    float* offsets, int32_t* children, int32_t* indices);
void rinnd_forest_free(rinnd_forest* f);

/* `NeighborHeap::sort_all` (core/heap/neighbor_heap.rs:293-320): stable ascending sort of each row by
 * distance, in place. */
rinnd_status rinnd_sort_rows(rinnd_ctx* ctx, int32_t* idx, float* dist, uint8_t* flags, int64_t n,
                             int64_t k);

/* `SearchGraph::from_dense_diversified` (core/graph/search_graph.rs:263-360) on rows sorted by
 * rinnd_sort_rows.  Writes CSR indptr (n+1) and indices (capacity n*round(mult*k)); returns the
 * edge count and min_distance. */
rinnd_status rinnd_search_graph_build(rinnd_ctx* ctx, const int32_t* sorted_idx,
                                      const float* sorted_dist, const float* data, int64_t n,
                                      int64_t k, int64_t dim, rinnd_distance dist,
                                      float diversify_prob, float pruning_degree_multiplier,
                                      int32_t* out_indptr, int32_t* out_indices, int64_t* out_nnz,
                                      float* out_min_distance);

/* ---- index (build_with_distance + query) -------------------------------------------------------- */
/* Mirrors NNDescentBuilder (core/index.rs:449-583) with the optional fields resolved by the caller:
 * leaf_size/max_candidates/n_iters <= 0 select the reference defaults (index.rs:617-625). */
typedef struct rinnd_index_params {
    int64_t n_neighbors;
    int64_t n_trees;
    int64_t n_search_trees;
    int64_t search_tree_leaf_budget;
    int64_t leaf_size;
    int64_t max_candidates;
    int64_t n_iters;
    float delta;
    float diversify_prob;
    float pruning_degree_multiplier;
    int32_t angular;
    uint64_t random_seed;
    int32_t verbose;
    int32_t n_threads;
    int32_t rng_mode;
    int32_t join_precision;
} rinnd_index_params;

/* `NNDescentBuilder::build_with_distance` (core/index.rs:601-733): nn_descent, sort_all, diversified
 * search graph, correction, first-tree vertex reordering; data, CSR graph and trees stay resident in
 * HBM for rinnd_index_query. */
rinnd_status rinnd_index_build(rinnd_ctx* ctx, const float* data, int64_t n, int64_t dim,
                               rinnd_distance dist, rinnd_correction corr,
                               const rinnd_index_params* params, rinnd_index** out,
                               rinnd_build_stats* stats);
void rinnd_index_free(rinnd_index* index);

/* `NNDescentIndex::query` (core/index.rs:406-439): greedy search per query with FastRng(seed+qid),
 * ids mapped to input order, distances corrected, rows padded with (-1, +inf).  out_idx/out_dist are
 * nq x k. */
rinnd_status rinnd_index_query(rinnd_index* index, const float* queries, int64_t nq, int64_t k,
                               float epsilon, int32_t* out_idx, float* out_dist);
/* `query_one_with_workspace` (core/index.rs:253-261): one query with query id 0. */
rinnd_status rinnd_index_query_one(rinnd_index* index, const float* query, int64_t k, float epsilon,
                                   int32_t* out_idx, float* out_dist);
/* FP8 pre-filter, with no counterpart in the reference: the quantized query's structure (query_quantized,
 * rinnd/src/lib.rs:825-870, over query_one_quantized_i8_with_workspace, core/index.rs:283-330) on any index,
 * with candidate distances from FP8 e4m3 codes of the data (per-row symmetric scale, built on the first call) on
 * the matrix cores instead of the reference's int8 codes (rinnd_index_query_i8 has those): the greedy search runs
 * with a candidate width resolved as in resolve_quantized_candidate_width (core/index.rs:55-78), then, when
 * should_rerank_quantized (core/index.rs:81-89) says so, the first rerank_width candidates get their FP32
 * distance and are sorted by (distance, id) (rerank_quantized_candidates, core/index.rs:103-124).  Query ids
 * 0..nq-1 seed the random fill as in rinnd_index_query.  k results per query, ids in input order, the index's
 * correction applied, padded with (-1, +inf); the width rules the reference asserts are returned as
 * RINND_ERR_INVALID_ARG.  dim <= 2048. */
rinnd_status rinnd_index_query_quantized(rinnd_index* index, const float* queries, int64_t nq, int64_t k,
                                         float epsilon, int64_t candidate_width, int64_t rerank_width,
                                         int32_t* out_idx, float* out_dist);

/* tree_quantization = "int8" (crates/rinnd/src/lib.rs:242-248, 681-705): replaces the routing of every later query
 * on this index by int8 copies of its search trees (QuantizedFlatTree::from_flat_tree, core/tree/flat_tree.rs:45-92:
 * per node scale max|h| / 127, codes round(h / scale) clamped to [-127, 127]); the descent takes the query's
 * per-vector symmetric int8 codes (round(x * 127 / max|x|), scale max|x| / 127, lib.rs:303-347) and the margin
 * i8 dot * node scale * query scale + offset (flat_tree.rs:95-164).  The FP32 hyperplanes are released, as the
 * reference takes its FP32 search trees (lib.rs:228-249): rinnd_index_trees returns NULL afterwards.  Idempotent. */
rinnd_status rinnd_index_quantize_trees(rinnd_index* index);

/* cosine_distance_mode "int8" / "sq8u" / "sq8p" (QuantizedI8Index::new, crates/rinnd/src/lib.rs:158-224): int8
 * codes of the index's rows for the quantized search.  encoding 0 (int8): codes round(x * 127 / max|row|) and the
 * inverse norm of each row's codes; 1 (sq8u): one global multiplier 127 / max|data| and its dequantisation
 * scale; 2 (sq8p): per-row multipliers and each row's 1 / multiplier.  Codes are clamped to [-127, 127].  The
 * index must be a DirectNormalizedCosine one (dist RINND_DIST_DIRECT_COSINE over normalized rows). */
rinnd_status rinnd_index_encode_i8(rinnd_index* index, int encoding);
/* query_one_quantized_widths over the int8 codes (lib.rs:366-398 -> query_one_quantized_i8_with_workspace,
 * core/index.rs:283-330; greedy.rs:365-408): the query's codes under the index's encoding, candidate distances
 * 1 - clamp(i8 dot * query inverse norm * row inverse norm, -1, 1), widths as rinnd_index_query_quantized
 * (index.rs:55-90) and the exact FP32 re-ranking of the first rerank_width candidates (index.rs:103-124).
 * Query ids 0..nq-1; results in input ids, padded with -1 / +inf. */
rinnd_status rinnd_index_query_i8(rinnd_index* index, const float* queries, int64_t nq, int64_t k, float epsilon,
                                  int64_t candidate_width, int64_t rerank_width, int32_t* out_idx, float* out_dist);
/* retain_fp32_data = False (lib.rs:250-257): frees the FP32 rows once the int8 codes exist.  Afterwards only
 * rinnd_index_query_i8 without re-ranking runs; the others return RINND_ERR_INVALID_ARG ("exact reranking is
 * unavailable because FP32 vector data was released", lib.rs:287-300). */
rinnd_status rinnd_index_release_data(rinnd_index* index);

/* The query of rinnd_index_query without any device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  queries (nq x dim), out_idx and out_dist (nq x k) must be device pointers; results are
 * identical to rinnd_index_query's.  The search state lives in one fixed 128 MiB device array per GPU
 * (allocated when the library's code is loaded on that GPU, held until the process ends): per concurrent
 * query n + 8*min(n, 65536) + 4 bytes, plus nq * round_up(dim, 8) * 4 bytes for padded query rows when dim is
 * not a multiple of 8 or the rows are not 256-byte aligned.  A call whose smallest layout does not fit returns
 * RINND_ERR_UNSUPPORTED (split the batch); a query whose frontier exceeds 65,536 entries returns
 * RINND_ERR_UNSUPPORTED.  All calls on one GPU share the array: the library serialises them (a per-device lock
 * held until the call's work has finished); kernels of overlapping calls would race on it. */
rinnd_status rinnd_index_query_noalloc(rinnd_index* index, const float* queries, int64_t nq, int64_t k,
                                       float epsilon, int32_t* out_idx, float* out_dist);

/* ---- query service (asynchronous batch submission) ---------------------------------------------------
 * The reference answers each `query` call on the caller's thread (core/index.rs:406-439, a Rayon task per query)
 * and serialises `query_one` on a Mutex<SearchWorkspace> (py:571-586).  A server keeps one index resident and
 * answers requests submitted from any number of host threads: a dispatcher thread gathers the pending requests
 * into one launch (arrival batching; one launch mixes requests of different k and epsilon), a completion thread
 * retires the launches in order.  Each request's results equal those of rinnd_index_query (or, for a quantized
 * server, rinnd_index_query_quantized / rinnd_index_query_i8 with the request's widths) on the same queries: query
 * ids 0..nq-1 within the request seed its random fill (the int8 search seeds every query alike).
 * Admission by memory: a request is admitted into a launch when the launch has room for its queries (at most
 * max_batch per launch; launch buffers are fixed, max_inflight launches at a time, each on its own stream) and its
 * search width fits the beam memory (max_k result-heap entries per query); the concurrent searches share a fixed
 * pool of visited sets (n bytes + an 8-byte entry per frontier vertex, at most 65,536 frontier entries per query)
 * sized by visited_bytes and split evenly between the launch slots (at most 16,384 concurrent searches).  Requests
 * that do not fit wait in submission order.  A search-width floor (min_width) trades throughput for recall at small
 * k: such a request's results are the first k of rinnd_index_query with k = max(k, min_width). */
typedef struct rinnd_server rinnd_server;
typedef struct rinnd_server_params {
    int64_t max_batch;      /* queries per launch (and per request), e.g. 4096 */
    int32_t max_inflight;   /* launches prepared or running at a time (each owns its staging buffers), e.g. 3 */
    int64_t max_k;          /* largest search width a request may use (beam memory per query) */
    int64_t visited_bytes;  /* memory of the visited-set pool; fixes the number of concurrent searches */
    int32_t quantized;      /* 1: FP8 e4m3 candidate distances on MFMA (rinnd_index_query_quantized); 2: the
                               index's int8 codes (rinnd_index_query_i8); then FP32 re-ranking */
    int64_t candidate_width;/* quantized: as rinnd_index_query_quantized's candidate_width */
    int64_t rerank_width;   /* quantized: as rinnd_index_query_quantized's rerank_width */
    int64_t target_batch;   /* a launch closes once it holds this many queries (0: min(1024, max_batch)) */
    int64_t min_width;      /* FP32 server: a request of k results searches with width max(k, min_width) and
                               returns the first k of those results (0: width k, the reference's query) */
} rinnd_server_params;
typedef struct rinnd_server_stats {
    int64_t requests, queries, launches, max_launch_queries;
    double gpu_ms;          /* summed launch durations (HIP events on the server's stream) */
} rinnd_server_stats;
/* The index must outlive the server.  The server owns a HIP stream and two host threads. */
rinnd_status rinnd_server_create(rinnd_index* index, const rinnd_server_params* params, rinnd_server** out);
/* Queues nq queries (nq x dim, host or device memory) with k results each (out_idx / out_dist nq x k, host or
 * device memory); the buffers must stay valid until the request completes.  Returns at once with a ticket. */
rinnd_status rinnd_server_submit(rinnd_server* server, const float* queries, int64_t nq, int64_t k, float epsilon,
                                 int32_t* out_idx, float* out_dist, uint64_t* ticket);
/* The same with the quantized widths of this request, as query_quantized takes them per call (py:825-870); the
 * server must be quantized.  candidate_width / rerank_width resolve as in rinnd_index_query_quantized. */
rinnd_status rinnd_server_submit_quantized(rinnd_server* server, const float* queries, int64_t nq, int64_t k,
                                           float epsilon, int64_t candidate_width, int64_t rerank_width,
                                           int32_t* out_idx, float* out_dist, uint64_t* ticket);
/* *done = 1 once the request has finished (its results are in the caller's buffers); the ticket is then released
 * and the request's status returned. */
rinnd_status rinnd_server_poll(rinnd_server* server, uint64_t ticket, int32_t* done);
/* Blocks until the request has finished; releases the ticket and returns the request's status. */
rinnd_status rinnd_server_wait(rinnd_server* server, uint64_t ticket);
rinnd_status rinnd_server_get_stats(const rinnd_server* server, rinnd_server_stats* out);
const char* rinnd_server_last_error(const rinnd_server* server);
/* Finishes the queued requests, then stops the threads and frees the server. */
void rinnd_server_destroy(rinnd_server* server);

/* `neighbor_graph` getter (py:1087-1103): sorted, corrected n x k arrays in input order. */
rinnd_status rinnd_index_neighbor_graph(const rinnd_index* index, int32_t* out_idx, float* out_dist);
/* Search graph CSR sizes and `search_graph_original_order` (core/index.rs:137-141, py:1013-1025). */
rinnd_status rinnd_index_search_graph_info(const rinnd_index* index, int64_t* n, int64_t* nnz,
                                           float* min_distance);
rinnd_status rinnd_index_search_graph(const rinnd_index* index, int32_t* indptr, int32_t* indices);
/* `vertex_order` (core/index.rs:47, 679-685): internal id -> input id. */
rinnd_status rinnd_index_vertex_order(const rinnd_index* index, int64_t* out);
/* Retained search trees (core/index.rs:679-680, 712-715) in the internal (reordered) id space; NULL once
 * rinnd_index_quantize_trees has replaced them by int8 trees (lib.rs:1117-1120: no retained search tree). */
const rinnd_forest* rinnd_index_trees(const rinnd_index* index);

#ifdef __cplusplus
}
#endif

#endif /* RINND_HIP_H */
