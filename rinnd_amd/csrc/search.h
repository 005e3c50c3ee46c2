// search.h - batched greedy search launcher.
#pragma once

#include "internal.h"

namespace rinnd {

struct SearchArgs {
    int metric;
    const float* X;  // reordered data, row stride ld
    int64_t ld;
    int dim;
    int64_t n;
    const int32_t* indptr;  // reordered CSR
    const int32_t* indices;
    int n_trees;
    const float* const* hyper;  // per tree, n_nodes x hdim (padded)
    int64_t hdim;
    const float* const* offsets;
    // int8 routing trees (tree_quantization = int8; null: the FP32 hyperplanes above): per tree n_nodes x hdim
    // codes and one scale per node; the descent uses the query's per-vector int8 codes
    const int8_t* const* hq;
    const float* const* hqs;
    const int32_t* const* children;
    const int32_t* const* tindices;
    int leaf_budget;
    const float* Q;  // queries, row stride ldq
    int64_t ldq;
    int64_t nq;
    int64_t qid_base;
    int k;
    float eps;
    float min_distance;
    uint64_t seed;
    const int64_t* vertex_order;
    int corr;
    int32_t* out_idx;
    float* out_dist;
    uint8_t* visited;  // slots x n
    int32_t* epochs;   // slots
    float* seed_spill_d;  // slots x seed_cap
    int32_t* seed_spill_i;
    int64_t seed_cap;
    int32_t* status;
    // FP8 e4m3 candidate search (q8 null: FP32 distances throughout).  k above is then the candidate
    // width; k_out results are written per query, after an exact FP32 re-ranking of the first `rerank`
    // candidates when rerank > 0.
    const uint8_t* q8;  // n x ld8 codes in the reordered row order, zero past dim
    int64_t ld8;
    const float* q8s;   // per-row dequantisation scale
    int k_out;
    int rerank;
    const float* q8n;   // squared norms of the dequantized FP8 rows (FP8 search)
    // int8 candidate search over the quantized index's codes (i8 null: none; alone, on a directly normalized
    // cosine index): n x ldi8 codes (zero past dim), per-row inverse norm or dequantisation scale, the encoding
    // (0 int8 per-vector cosine, 1 sq8u global symmetric, 2 sq8p per-vector symmetric) and sq8u's global
    // multiplier and dequantisation scale
    const int8_t* i8;
    int64_t ldi8;
    const float* i8inv;
    int i8enc;
    float i8gm, i8gd;
    int seed_fixed;  // 1: every query's RNG starts from seed itself (the quantized query, core/index.rs:298)
    // Per-query parameters (null: the scalar fields above for every query).  k above is then the largest search
    // width of the launch (it sizes the result heap in LDS).
    const int32_t* q_k;       // search width
    const int32_t* q_kout;    // results written
    const int32_t* q_rerank;  // candidates re-ranked in FP32 (FP8 search)
    const float* q_eps;
    const int64_t* q_qid;     // query id of the RNG seed, FastRng::new(seed + id)
    int32_t* const* q_oi;     // where the query's k_out results go
    float* const* q_od;
    int qcap;  // floats of the query row kept in LDS (0: read from memory); set by launch_search
    int kcap;  // result-heap entries in LDS; set by launch_search
    int c8cap; // bytes of the query's FP8 or int8 codes in LDS; set by launch_search
    int i8cap; // bytes of the query's int8 routing codes in LDS; set by launch_search
};

void launch_search(hipStream_t s, SearchArgs a, int slots);
// per-row symmetric FP8 e4m3 codes: scale = max|x| / 240, code = e4m3(x / scale), rows of ld8 bytes (a multiple
// of 32: whole MFMA k-steps), and the squared norm of each dequantized row
void launch_fp8_encode(hipStream_t s, const float* X, int64_t ld, int dim, int64_t n, uint8_t* codes, int64_t ld8,
                       float* scales, float* norms);
// the quantized index's int8 codes of the rows (encoding as SearchArgs::i8enc) and per-row inverse norms or scales
void launch_i8_encode(hipStream_t s, const float* X, int64_t ld, int dim, int64_t n, int enc, float gm, float gd,
                      int8_t* codes, int64_t ldi8, float* inv);
// max |x| over n rows of dim values (stride ld), as float bits into *out, which must start at zero
void launch_abs_max(hipStream_t s, const float* X, int64_t ld, int dim, int64_t n, unsigned* out);
// int8 copy of a tree's hyperplanes (QuantizedFlatTree::from_flat_tree): per node scale max|h| / 127 and codes
// round(h / scale) clamped to [-127, 127]
void launch_quantize_tree(hipStream_t s, const float* hyper, int64_t n_nodes, int64_t hdim, int dim, int8_t* codes,
                          float* scales);
int search_max_k();
int search_max_degree();
int search_max_budget();

}  // namespace rinnd
