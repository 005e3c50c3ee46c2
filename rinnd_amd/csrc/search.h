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
};

void launch_search(hipStream_t s, const SearchArgs& a, int slots);
int search_max_k();
int search_max_degree();
int search_max_budget();

}  // namespace rinnd
