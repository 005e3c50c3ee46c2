// graph.h - device search graph (diversified CSR) and vertex reordering helpers.
#pragma once

#include "internal.h"

namespace rinnd {

// Pruned search graph: row v holds deg[v] <= max_degree neighbours (sorted by id) at
// rows[v * max_degree ...] in the original id space.
struct SearchGraphDev {
    int64_t n = 0;
    int max_degree = 0;
    DBuf<int32_t> rows;
    DBuf<int32_t> deg;
    float min_distance = 0.f;
};

void build_search_graph_dev(hipStream_t s, int metric, const float* X, int64_t ld, int dim,
                            const int32_t* sorted_idx, const float* sorted_dist, int64_t n, int k,
                            float prob, float mult, uint64_t seed, SearchGraphDev& G);
void search_graph_csr(hipStream_t s, const SearchGraphDev& G, const int64_t* order, const int64_t* inv,
                      DBuf<int32_t>& indptr, DBuf<int32_t>& indices, int64_t& nnz);
void vertex_order_from_tree(hipStream_t s, const int32_t* tree_indices, int64_t n, DBuf<int64_t>& order,
                            DBuf<int64_t>& inv);
void gather_rows(hipStream_t s, const float* X, int64_t ld, const int64_t* order, int64_t n, float* out);
void remap_ids(hipStream_t s, int32_t* ids, int64_t count, const int64_t* inv);

}  // namespace rinnd
