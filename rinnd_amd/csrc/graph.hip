// graph.hip - search-graph preparation on the GPU:
//   SearchGraph::from_dense_diversified (core/graph/search_graph.rs:263-360) with diversify_row
//   (432-473), transpose_csr (526-567), diversify_sorted_csr (478-523) and weighted_union_pruned
//   (368-426); SearchGraph::reorder (199-229) and the tree/data permutation of build_with_distance
//   (core/index.rs:675-715).
#include <hipcub/hipcub.hpp>

#include "graph.h"

namespace rinnd {

static inline unsigned ggrid(int64_t work, int block, int64_t cap = 1 << 20) {
    int64_t g = ceil_div(work, block);
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

// ---- diversification (one wave per row) ---------------------------------------------------------
// Candidates are visited in `order`; candidate j survives unless a previously kept c has
// d(i,c) > f32::MIN_POSITIVE and d(c,j) < d(i,j).  With diversify_prob < 1 the reference draws from
// rand::thread_rng (not reproducible); here the draw is a counter hash of (seed, row, pos, c).
// Lane c evaluates the kept neighbour c, so up to 64 kept neighbours are tested at once.
__device__ void diversify_row_wave(int metric, const float* __restrict__ X, int64_t ld, int dim,
                                   const int32_t* cand, const float* cdist, const int32_t* order,
                                   int L, bool stop_at_neg, float prob, uint64_t seed, int64_t row,
                                   int max_degree, int32_t* kept_i, float* kept_d, int* n_kept) {
    const int lane = lane_id();
    int nk = 0;
    for (int t = 0; t < L; ++t) {
        const int pos = order ? order[t] : t;
        const int32_t j = cand[pos];
        if (j < 0) {
            if (stop_at_neg) break;
            continue;
        }
        const float dij = cdist[pos];
        bool prune = false;
        for (int c0 = 0; c0 < nk; c0 += RINND_WAVE) {
            const int c = c0 + lane;
            bool hit = false;
            if (c < nk) {
                const float dic = kept_d[c];
                if (dic > 1.17549435e-38f) {
                    const float dcj = distance_exact(metric, X + (int64_t)kept_i[c] * ld,
                                                     X + (int64_t)j * ld, dim);
                    if (dcj < dij) {
                        if (prob >= 1.0f) {
                            hit = true;
                        } else {
                            const uint64_t h = mix64(seed ^ mix64(((uint64_t)row << 24) ^ ((uint64_t)t << 8) ^ c));
                            hit = (float)(uint32_t)(h >> 40) * (1.0f / 16777216.0f) < prob;
                        }
                    }
                }
            }
            if (__ballot(hit)) {
                prune = true;
                break;
            }
        }
        if (!prune) {
            if (lane == 0) {
                kept_i[nk] = j;
                kept_d[nk] = dij;
            }
            wave_sync();
            ++nk;
            if (nk >= max_degree) break;
        }
    }
    if (lane == 0) *n_kept = nk;
}

// forward: rows of the sorted k-NN heap (search_graph.rs:280-296)
__global__ __launch_bounds__(256) void k_diversify_forward(int metric, const float* __restrict__ X,
                                                           int64_t ld, int dim,
                                                           const int32_t* __restrict__ idx,
                                                           const float* __restrict__ dist,
                                                           int64_t n, int k, float prob,
                                                           uint64_t seed, int max_degree,
                                                           int32_t* fwd_i, float* fwd_d,
                                                           int32_t* fwd_n) {
    __shared__ int32_t ki[4][256];
    __shared__ float kd[4][256];
    __shared__ int nk[4];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int64_t i = (int64_t)blockIdx.x * 4 + w; i < n; i += (int64_t)gridDim.x * 4) {
        diversify_row_wave(metric, X, ld, dim, idx + i * k, dist + i * k, nullptr, k, true, prob, seed,
                           i, max_degree, ki[w], kd[w], &nk[w]);
        wave_sync();
        const int m = nk[w];
        for (int c = lane; c < m; c += 64) {
            fwd_i[i * max_degree + c] = ki[w][c];
            fwd_d[i * max_degree + c] = kd[w][c];
        }
        if (lane == 0) fwd_n[i] = m;
        wave_sync();
    }
}

// reverse rows: members ordered by (distance, transposed position) (search_graph.rs:316-346)
constexpr int kRevCap = 1024;
__global__ __launch_bounds__(64) void k_diversify_reverse(int metric, const float* __restrict__ X,
                                                          int64_t ld, int dim,
                                                          const int64_t* __restrict__ rptr,
                                                          const int32_t* __restrict__ ridx,
                                                          const float* __restrict__ rdst, int64_t n,
                                                          float prob, uint64_t seed, int max_degree,
                                                          int32_t* rev_i, float* rev_d, int32_t* rev_n,
                                                          int32_t* order_spill) {
    __shared__ uint64_t keys[kRevCap];
    __shared__ int32_t order[kRevCap];
    __shared__ int32_t ki[256];
    __shared__ float kd[256];
    __shared__ int nk;
    const int lane = threadIdx.x;
    for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
        const int64_t s = rptr[i];
        const int L = (int)(rptr[i + 1] - s);
        if (L == 0) {
            if (lane == 0) rev_n[i] = 0;
            continue;
        }
        const int32_t* ord = order;
        if (L <= kRevCap) {
            int S = 1;
            while (S < L) S <<= 1;
            for (int t = lane; t < S; t += 64)
                keys[t] = t < L ? (((uint64_t)float_order_key(rdst[s + t]) << 32) | (uint32_t)t) : ~0ull;
            wave_sync();
            for (int kk = 2; kk <= S; kk <<= 1)
                for (int j = kk >> 1; j > 0; j >>= 1) {
                    for (int t = lane; t < S; t += 64) {
                        const int ixj = t ^ j;
                        if (ixj > t) {
                            const bool up = (t & kk) == 0;
                            const uint64_t A = keys[t], B = keys[ixj];
                            if ((A > B) == up) {
                                keys[t] = B;
                                keys[ixj] = A;
                            }
                        }
                    }
                    wave_sync();
                }
            for (int t = lane; t < L; t += 64) order[t] = (int32_t)(keys[t] & 0xffffffffu);
            wave_sync();
        } else {
            // long reverse rows: stable selection by rank (O(L^2 / 64)), written to global scratch
            int32_t* og = order_spill + s;
            for (int t = lane; t < L; t += 64) {
                const float dt = rdst[s + t];
                const uint32_t kt = float_order_key(dt);
                int rank = 0;
                for (int u = 0; u < L; ++u) {
                    const uint32_t ku = float_order_key(rdst[s + u]);
                    rank += (ku < kt) || (ku == kt && u < t);
                }
                og[rank] = t;
            }
            __threadfence_block();
            wave_sync();
            ord = og;
        }
        diversify_row_wave(metric, X, ld, dim, ridx + s, rdst + s, ord, L, false, prob, seed ^ 0x5bd1e995ull,
                           i, max_degree, ki, kd, &nk);
        wave_sync();
        const int m = nk;
        for (int c = lane; c < m; c += 64) {
            rev_i[i * max_degree + c] = ki[c];
            rev_d[i * max_degree + c] = kd[c];
        }
        if (lane == 0) rev_n[i] = m;
        wave_sync();
    }
}

// transpose_csr (search_graph.rs:526-567): in-degree, scan, then rows filled in source order.
__global__ void k_indegree(const int32_t* fwd_i, const int32_t* fwd_n, int64_t n, int max_degree,
                           unsigned long long* deg) {
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x)
        for (int c = 0; c < fwd_n[v]; ++c) atomicAdd(&deg[fwd_i[v * max_degree + c]], 1ull);
}
// Slots are claimed with atomics (any order); rows are then restored to source order by the key.
__global__ void k_transpose_fill(const int32_t* fwd_i, const float* fwd_d, const int32_t* fwd_n,
                                 int64_t n, int max_degree, const uint64_t* rptr,
                                 unsigned long long* cur, uint64_t* tkey, float* tdst) {
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x)
        for (int c = 0; c < fwd_n[v]; ++c) {
            const int32_t t = fwd_i[v * max_degree + c];
            const uint64_t pos = rptr[t] + atomicAdd(&cur[t], 1ull);
            tkey[pos] = ((uint64_t)v << 16) | (uint64_t)c;  // source order: (v, position)
            tdst[pos] = fwd_d[v * max_degree + c];
        }
}
// per reverse row: sort (key, dist) by key, emit source vertex ids
__global__ __launch_bounds__(64) void k_transpose_order(const uint64_t* rptr, const uint64_t* tkey,
                                                        float* tdst, int32_t* ridx, int32_t* rank_out,
                                                        int64_t n) {
    __shared__ uint64_t keys[kRevCap];
    __shared__ float ds[kRevCap];
    const int lane = threadIdx.x;
    for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
        const uint64_t s = rptr[i];
        const int L = (int)(rptr[i + 1] - s);
        if (L == 0) continue;
        if (L <= kRevCap) {
            int S = 1;
            while (S < L) S <<= 1;
            for (int t = lane; t < S; t += 64) {
                keys[t] = t < L ? tkey[s + t] : ~0ull;
                ds[t] = t < L ? tdst[s + t] : 0.f;
            }
            wave_sync();
            for (int kk = 2; kk <= S; kk <<= 1)
                for (int j = kk >> 1; j > 0; j >>= 1) {
                    for (int t = lane; t < S; t += 64) {
                        const int ixj = t ^ j;
                        if (ixj > t) {
                            const bool up = (t & kk) == 0;
                            const uint64_t A = keys[t], B = keys[ixj];
                            if ((A > B) == up) {
                                keys[t] = B;
                                keys[ixj] = A;
                                const float x = ds[t];
                                ds[t] = ds[ixj];
                                ds[ixj] = x;
                            }
                        }
                    }
                    wave_sync();
                }
            for (int t = lane; t < L; t += 64) {
                ridx[s + t] = (int32_t)(keys[t] >> 16);
                tdst[s + t] = ds[t];
            }
        } else {  // rank-based placement (keys are unique); ranks kept for the distance permute
            for (int t = lane; t < L; t += 64) {
                const uint64_t kt = tkey[s + t];
                int rank = 0;
                for (int u = 0; u < L; ++u) rank += tkey[s + u] < kt;
                ridx[s + rank] = (int32_t)(kt >> 16);
                rank_out[s + t] = rank;
            }
        }
        wave_sync();
    }
}
__global__ void k_permute_long_rows(const uint64_t* rptr, const int32_t* rank, const float* src,
                                    float* dst, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t s = rptr[i];
        const int64_t L = (int64_t)(rptr[i + 1] - s);
        if (L > kRevCap)
            for (int64_t t = 0; t < L; ++t) dst[s + rank[s + t]] = src[s + t];
    }
}

// weighted_union_pruned (search_graph.rs:368-426), one wave per vertex.
__global__ __launch_bounds__(64) void k_union(const int32_t* fwd_i, const float* fwd_d,
                                              const int32_t* fwd_n, const int32_t* rev_i,
                                              const float* rev_d, const int32_t* rev_n, int64_t n,
                                              int max_degree, int32_t* out_i, int32_t* out_n) {
    __shared__ uint64_t keys[512];
    __shared__ float ds[512];
    __shared__ int cntv;
    const int lane = threadIdx.x;
    for (int64_t v = blockIdx.x; v < n; v += gridDim.x) {
        const int nf = fwd_n[v], nr = rev_n[v];
        const int L = nf + nr;
        int S = 1;
        while (S < L) S <<= 1;
        // filter: idx >= 0, idx != v, finite distance; key = (idx, source order)
        for (int t = lane; t < S; t += 64) {
            uint64_t key = ~0ull;
            float d = 0.f;
            if (t < L) {
                const int32_t x = t < nf ? fwd_i[v * max_degree + t] : rev_i[v * max_degree + (t - nf)];
                d = t < nf ? fwd_d[v * max_degree + t] : rev_d[v * max_degree + (t - nf)];
                if (x >= 0 && x != (int32_t)v && isfinite(d)) key = ((uint64_t)(uint32_t)x << 32) | (uint32_t)t;
            }
            keys[t] = key;
            ds[t] = d;
        }
        wave_sync();
        for (int kk = 2; kk <= S; kk <<= 1)
            for (int j = kk >> 1; j > 0; j >>= 1) {
                for (int t = lane; t < S; t += 64) {
                    const int ixj = t ^ j;
                    if (ixj > t) {
                        const bool up = (t & kk) == 0;
                        const uint64_t A = keys[t], B = keys[ixj];
                        if ((A > B) == up) {
                            keys[t] = B;
                            keys[ixj] = A;
                            const float x = ds[t];
                            ds[t] = ds[ixj];
                            ds[ixj] = x;
                        }
                    }
                }
                wave_sync();
            }
        // dedupe (max distance per id) into (total_cmp(dist), id) keys at the group heads
        if (lane == 0) {
            int m = 0;
            int t = 0;
            while (t < L && keys[t] != ~0ull) {
                const uint32_t x = (uint32_t)(keys[t] >> 32);
                float d = ds[t];
                int u = t + 1;
                while (u < L && keys[u] != ~0ull && (uint32_t)(keys[u] >> 32) == x) {
                    d = fmaxf(d, ds[u]);
                    ++u;
                }
                // f32::total_cmp key (finite distances only reach here)
                uint32_t b = __builtin_bit_cast(uint32_t, d);
                b = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
                keys[m] = ((uint64_t)b << 32) | x;
                ++m;
                t = u;
            }
            cntv = m;
        }
        wave_sync();
        const int m = cntv;
        for (int t = m + lane; t < S; t += 64) keys[t] = ~0ull;
        wave_sync();
        for (int kk = 2; kk <= S; kk <<= 1)
            for (int j = kk >> 1; j > 0; j >>= 1) {
                for (int t = lane; t < S; t += 64) {
                    const int ixj = t ^ j;
                    if (ixj > t) {
                        const bool up = (t & kk) == 0;
                        const uint64_t A = keys[t], B = keys[ixj];
                        if ((A > B) == up) {
                            keys[t] = B;
                            keys[ixj] = A;
                        }
                    }
                }
                wave_sync();
            }
        const int keep = min(m, max_degree);
        // truncate, then order by id: ids are < 2^31, pack into the low bits for a second sort
        for (int t = lane; t < S; t += 64) keys[t] = t < keep ? (keys[t] & 0xffffffffull) : ~0ull;
        wave_sync();
        for (int kk = 2; kk <= S; kk <<= 1)
            for (int j = kk >> 1; j > 0; j >>= 1) {
                for (int t = lane; t < S; t += 64) {
                    const int ixj = t ^ j;
                    if (ixj > t) {
                        const bool up = (t & kk) == 0;
                        const uint64_t A = keys[t], B = keys[ixj];
                        if ((A > B) == up) {
                            keys[t] = B;
                            keys[ixj] = A;
                        }
                    }
                }
                wave_sync();
            }
        for (int t = lane; t < keep; t += 64) out_i[v * max_degree + t] = (int32_t)keys[t];
        if (lane == 0) out_n[v] = keep;
        wave_sync();
    }
}

// min_distance over forward edges (0 -> f32::EPSILON, finite only), search_graph.rs:353-357
__global__ void k_min_distance(const float* fwd_d, const int32_t* fwd_n, int64_t n, int max_degree,
                               unsigned int* out_bits) {
    unsigned int best = 0xffffffffu;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x)
        for (int c = 0; c < fwd_n[v]; ++c) {
            float d = fwd_d[v * max_degree + c];
            if (d == 0.0f) d = 1.1920929e-07f;  // f32::EPSILON
            if (isfinite(d)) best = min(best, float_order_key(d));
        }
    if (best != 0xffffffffu) atomicMin(out_bits, best);
}

__global__ void k_count_to_u64(const int32_t* c, unsigned long long* out, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = i < n ? (unsigned long long)c[i] : 0ull;
}

// reorder (search_graph.rs:199-229): new row nv = old row order[nv] with ids mapped through inv.
__global__ void k_reorder_len(const int32_t* deg, const int64_t* order, int64_t n,
                              unsigned long long* out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = i < n ? (unsigned long long)deg[order[i]] : 0ull;
}
__global__ void k_reorder_rows(const int32_t* rows, int max_degree, const int64_t* order,
                               const int64_t* inv, const uint64_t* nptr, int64_t n, int32_t* out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t ov = order[i];
        const uint64_t s = nptr[i], e = nptr[i + 1];
        for (uint64_t t = 0; t < e - s; ++t) out[s + t] = (int32_t)inv[rows[ov * max_degree + t]];
    }
}
__global__ void k_invert(const int64_t* order, int64_t n, int64_t* inv) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        inv[order[i]] = i;
}
__global__ void k_order_from_tree(const int32_t* tree_idx, int64_t n, int64_t* order) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        order[i] = tree_idx[i];
}
__global__ void k_gather_rows(const float* X, int64_t ld, const int64_t* order, int64_t n,
                              float* out) {
    const int64_t total = n * ld;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = t / ld, c = t - r * ld;
        out[t] = X[order[r] * ld + c];
    }
}
__global__ void k_remap_ids(int32_t* ids, int64_t n, const int64_t* inv) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        ids[i] = (int32_t)inv[ids[i]];
}
__global__ void k_ptr_to_i32(const uint64_t* in, int32_t* out, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (int32_t)in[i];
}

static size_t scan_bytes_for(int64_t n) {
    size_t b = 0;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, b, (unsigned long long*)nullptr,
                                             (unsigned long long*)nullptr, (int)n));
    return std::max<size_t>(b, 16);
}
static void scan(hipStream_t s, unsigned long long* in, uint64_t* out, int64_t n, DBuf<char>& tmp) {
    size_t b = scan_bytes_for(n);
    tmp.ensure(b);
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp.p, b, in, (unsigned long long*)out, (int)n, s));
}

void build_search_graph_dev(hipStream_t s, int metric, const float* X, int64_t ld, int dim,
                            const int32_t* sorted_idx, const float* sorted_dist, int64_t n, int k,
                            float prob, float mult, uint64_t seed, SearchGraphDev& G) {
    const int max_degree = std::max(1, (int)std::lround((double)(mult * (float)k)));
    RINND_CHECK_ARG(max_degree <= 256, "pruning_degree_multiplier * n_neighbors above 256");
    G.n = n;
    G.max_degree = max_degree;
    DBuf<int32_t> fwd_i((size_t)n * max_degree), fwd_n(n), rev_i((size_t)n * max_degree), rev_n(n);
    DBuf<float> fwd_d((size_t)n * max_degree), rev_d((size_t)n * max_degree);
    hipLaunchKernelGGL(k_diversify_forward, dim3(ggrid(n, 4, 1 << 16)), dim3(256), 0, s, metric, X, ld, dim,
                       sorted_idx, sorted_dist, n, k, prob, seed, max_degree, fwd_i.p, fwd_d.p, fwd_n.p);
    // transpose
    DBuf<unsigned long long> deg(n + 1), cur(n + 1);
    DBuf<uint64_t> rptr(n + 1);
    DBuf<char> tmp;
    HIP_TRY(hipMemsetAsync(deg.p, 0, (n + 1) * 8, s));
    hipLaunchKernelGGL(k_indegree, dim3(ggrid(n, 256)), dim3(256), 0, s, fwd_i.p, fwd_n.p, n, max_degree, deg.p);
    scan(s, deg.p, rptr.p, n + 1, tmp);
    uint64_t nnz_t = 0;
    HIP_TRY(hipMemcpyAsync(&nnz_t, rptr.p + n, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    DBuf<uint64_t> tkey(std::max<uint64_t>(nnz_t, 1));
    DBuf<float> tdst(std::max<uint64_t>(nnz_t, 1)), tdst2(std::max<uint64_t>(nnz_t, 1));
    DBuf<int32_t> ridx(std::max<uint64_t>(nnz_t, 1)), spill(std::max<uint64_t>(nnz_t, 1));
    HIP_TRY(hipMemsetAsync(cur.p, 0, (n + 1) * 8, s));
    hipLaunchKernelGGL(k_transpose_fill, dim3(ggrid(n, 256)), dim3(256), 0, s, fwd_i.p, fwd_d.p, fwd_n.p, n,
                       max_degree, rptr.p, cur.p, tkey.p, tdst.p);
    HIP_TRY(hipMemcpyAsync(tdst2.p, tdst.p, std::max<uint64_t>(nnz_t, 1) * 4, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_transpose_order, dim3(ggrid(n, 1, 1 << 16)), dim3(64), 0, s, rptr.p, tkey.p, tdst.p,
                       ridx.p, spill.p, n);
    hipLaunchKernelGGL(k_permute_long_rows, dim3(ggrid(n, 256)), dim3(256), 0, s, rptr.p, spill.p, tdst2.p,
                       tdst.p, n);
    hipLaunchKernelGGL(k_diversify_reverse, dim3(ggrid(n, 1, 1 << 16)), dim3(64), 0, s, metric, X, ld, dim,
                       (const int64_t*)rptr.p, ridx.p, tdst.p, n, prob, seed, max_degree, rev_i.p, rev_d.p,
                       rev_n.p, spill.p);
    // union + prune
    DBuf<int32_t> rows((size_t)n * max_degree), deg_out(n);
    hipLaunchKernelGGL(k_union, dim3(ggrid(n, 1, 1 << 16)), dim3(64), 0, s, fwd_i.p, fwd_d.p, fwd_n.p, rev_i.p,
                       rev_d.p, rev_n.p, n, max_degree, rows.p, deg_out.p);
    DBuf<unsigned int> mdb(1);
    const unsigned int inf_bits = 0xffffffffu;
    HIP_TRY(hipMemcpyAsync(mdb.p, &inf_bits, 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_min_distance, dim3(ggrid(n, 256, 4096)), dim3(256), 0, s, fwd_d.p, fwd_n.p, n,
                       max_degree, mdb.p);
    unsigned int mbits = 0;
    HIP_TRY(hipMemcpyAsync(&mbits, mdb.p, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (mbits == 0xffffffffu) {
        G.min_distance = __builtin_inff();
    } else {
        const uint32_t u = (mbits & 0x80000000u) ? (mbits & 0x7fffffffu) : ~mbits;
        G.min_distance = __builtin_bit_cast(float, u);
    }
    G.rows = std::move(rows);
    G.deg = std::move(deg_out);
}

// CSR in a vertex order (identity when order == nullptr), ids mapped through inv.
void search_graph_csr(hipStream_t s, const SearchGraphDev& G, const int64_t* order, const int64_t* inv,
                      DBuf<int32_t>& indptr, DBuf<int32_t>& indices, int64_t& nnz) {
    const int64_t n = G.n;
    DBuf<int64_t> ident_o, ident_i;
    if (!order) {
        ident_o.alloc(std::max<int64_t>(n, 1));
        std::vector<int64_t> h(n);
        for (int64_t i = 0; i < n; ++i) h[i] = i;
        HIP_TRY(hipMemcpyAsync(ident_o.p, h.data(), n * 8, hipMemcpyHostToDevice, s));
        HIP_TRY(hipStreamSynchronize(s));
        order = ident_o.p;
        inv = ident_o.p;
    }
    DBuf<unsigned long long> len(n + 1);
    DBuf<uint64_t> ptr(n + 1);
    DBuf<char> tmp;
    hipLaunchKernelGGL(k_reorder_len, dim3(ggrid(n + 1, 256)), dim3(256), 0, s, G.deg.p, order, n, len.p);
    scan(s, len.p, ptr.p, n + 1, tmp);
    uint64_t tot = 0;
    HIP_TRY(hipMemcpyAsync(&tot, ptr.p + n, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    nnz = (int64_t)tot;
    indptr.alloc(n + 1);
    indices.alloc(std::max<uint64_t>(tot, 1));
    hipLaunchKernelGGL(k_ptr_to_i32, dim3(ggrid(n + 1, 256)), dim3(256), 0, s, ptr.p, indptr.p, n);
    hipLaunchKernelGGL(k_reorder_rows, dim3(ggrid(n, 256)), dim3(256), 0, s, G.rows.p, G.max_degree, order, inv,
                       ptr.p, n, indices.p);
    HIP_TRY(hipStreamSynchronize(s));
}

void vertex_order_from_tree(hipStream_t s, const int32_t* tree_indices, int64_t n, DBuf<int64_t>& order,
                            DBuf<int64_t>& inv) {
    order.alloc(std::max<int64_t>(n, 1));
    inv.alloc(std::max<int64_t>(n, 1));
    hipLaunchKernelGGL(k_order_from_tree, dim3(ggrid(n, 256)), dim3(256), 0, s, tree_indices, n, order.p);
    hipLaunchKernelGGL(k_invert, dim3(ggrid(n, 256)), dim3(256), 0, s, order.p, n, inv.p);
}

void gather_rows(hipStream_t s, const float* X, int64_t ld, const int64_t* order, int64_t n, float* out) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_gather_rows, dim3(ggrid(n * ld, 256)), dim3(256), 0, s, X, ld, order, n, out);
}

void remap_ids(hipStream_t s, int32_t* ids, int64_t count, const int64_t* inv) {
    if (count == 0) return;
    hipLaunchKernelGGL(k_remap_ids, dim3(ggrid(count, 256)), dim3(256), 0, s, ids, count, inv);
}

}  // namespace rinnd
