// index.h - the resident index (data, search graph and trees in HBM) shared by the query paths.
#pragma once

#include <vector>

#include "search.h"

struct rinnd_index {
    template <class T>
    using DBuf = rinnd::DBuf<T>;
    rinnd_ctx* ctx = nullptr;
    int64_t n = 0, dim = 0, k = 0, ld = 0;
    int metric = 0, corr = 0;
    DBuf<float> X;  // reordered, padded
    DBuf<int32_t> indptr, indices;
    int64_t nnz = 0;
    float min_distance = 0.f;
    rinnd_forest trees;  // retained search trees, ids in the reordered space
    std::vector<DBuf<float>> hyper_padded;
    DBuf<const float*> d_hyper, d_off;
    DBuf<const int32_t*> d_ch, d_tix;
    // int8 routing trees (rinnd_index_quantize_trees): per tree codes (n_nodes x ld) and scales, and their tables
    std::vector<DBuf<int8_t>> tq_codes;
    std::vector<DBuf<float>> tq_scales;
    DBuf<const int8_t*> d_tq;
    DBuf<const float*> d_tqs;
    bool tree_i8 = false;
    DBuf<int64_t> vertex_order, inv;
    DBuf<int32_t> nb_idx;
    DBuf<float> nb_dist;
    int64_t leaf_budget = 1;
    uint64_t seed = 42;
    // query workspace (grown on demand)
    int slots = 0;
    DBuf<uint8_t> visited;
    DBuf<int32_t> epochs;
    DBuf<float> spill_d;
    DBuf<int32_t> spill_i;
    int64_t seed_cap = 0;
    DBuf<int32_t> status;
    DBuf<float> qbuf;  // padded query rows (grow-only; padding columns stay zero)
    DBuf<int32_t> obuf_i;
    DBuf<float> obuf_d;
    // FP8 e4m3 codes of X for the quantized candidate search (built on the first quantized query)
    DBuf<uint8_t> q8;
    DBuf<float> q8s;
    DBuf<float> q8n;  // squared norms of the dequantized rows
    int64_t ld8 = 0;
    // the quantized index's int8 codes (rinnd_index_encode_i8): n x ldi8 codes of X, per-row inverse norms or
    // dequantisation scales, the encoding and sq8u's global multiplier and dequantisation scale
    DBuf<int8_t> i8c;
    DBuf<float> i8inv;
    int64_t ldi8 = 0;
    int i8enc = -1;
    float i8gm = 0.f, i8gd = 0.f;
    bool data_released = false;  // X freed (rinnd_index_release_data): only the int8 search without reranking
};

namespace rinnd {
// candidate distances of a query: the FP32 rows, their FP8 e4m3 codes or the quantized index's int8 codes
enum QueryRep : int { kQueryF32 = 0, kQueryFP8 = 1, kQueryI8 = 2 };
// the index's part of a launch of the greedy search: data, graph, trees and the query parameters
SearchArgs index_search_args(const rinnd_index& ix, const float* qd, int64_t nq, int64_t k, float eps, int64_t qid_base,
                             int64_t search_k, int64_t rerank, int rep, int32_t* oip, float* odp);
// the FP8 e4m3 codes of the rows (built on first use)
void ensure_fp8_codes(rinnd_index& ix);
bool device_pointer(const void* p);
}  // namespace rinnd
