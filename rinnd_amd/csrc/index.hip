// index.hip - NNDescentBuilder::build_with_distance (core/index.rs:601-733) and NNDescentIndex::query
// (core/index.rs:157-202, 406-439) with everything resident in HBM.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "graph.h"
#include "index.h"

namespace rinnd {
// defined in capi.hip
void index_nn_descent(rinnd_ctx* ctx, const float* data_dev_padded, int64_t ld, int64_t n, int dim,
                      int metric, const rinnd_params& p, uint64_t state[4], int32_t* idx, float* dist,
                      uint8_t* flags, rinnd_forest& F, rinnd_build_stats* stats);
void make_padded(hipStream_t s, const float* data, int64_t n, int dim, DBuf<float>& X, int64_t& ld);
bool device_pointer(const void* p);
}  // namespace rinnd

using namespace rinnd;


namespace rinnd {

static void index_build_impl(rinnd_ctx* ctx, const float* data, int64_t n, int64_t dim, int metric, int corr,
                             const rinnd_index_params& ip, rinnd_index& ix, rinnd_build_stats* stats) {
    hipStream_t s = ctx->stream;
    const int64_t k = ip.n_neighbors;
    rinnd_params p{};
    p.n_neighbors = k;
    p.n_trees = ip.n_trees;
    p.leaf_size = ip.leaf_size > 0 ? ip.leaf_size : std::max<int64_t>(60, std::min<int64_t>(256, 5 * k));
    p.max_candidates = ip.max_candidates > 0 ? ip.max_candidates : std::min<int64_t>(k, 60);
    p.n_iters = ip.n_iters > 0 ? ip.n_iters
                               : std::max<int64_t>(5, (int64_t)std::ceil(std::log2((double)std::max<int64_t>(n, 1))));
    p.delta = ip.delta;
    p.angular = ip.angular;
    p.max_depth = 200;
    p.verbose = ip.verbose;
    p.n_threads = ip.n_threads;
    p.rng_mode = ip.rng_mode;
    p.join_precision = ip.join_precision;
    ix.ctx = ctx;
    ix.n = n;
    ix.dim = dim;
    ix.k = k;
    ix.metric = metric;
    ix.corr = corr;
    ix.leaf_budget = std::max<int64_t>(1, ip.search_tree_leaf_budget);
    ix.seed = ip.random_seed;
    RINND_CHECK_ARG(ix.leaf_budget <= search_max_budget(), "search_tree_leaf_budget above 64");

    DBuf<float> X0;
    int64_t ld = 0;
    make_padded(s, data, n, (int)dim, X0, ld);
    ix.ld = ld;
    uint64_t st[4];
    rinnd_rng_seed(ip.random_seed, st);  // FastRng::new(random_seed) (index.rs:639)
    DBuf<int32_t> hidx((size_t)n * k);
    DBuf<float> hdist((size_t)n * k);
    DBuf<uint8_t> hflags((size_t)n * k);
    rinnd_forest forest;
    index_nn_descent(ctx, X0.p, ld, n, (int)dim, metric, p, st, hidx.p, hdist.p, hflags.p, forest, stats);
    launch_sort_rows(s, hidx.p, hdist.p, hflags.p, n, (int)k);  // sort_all (index.rs:653)
    SearchGraphDev G;
    build_search_graph_dev(s, metric, X0.p, ld, (int)dim, hidx.p, hdist.p, n, (int)k, ip.diversify_prob,
                           ip.pruning_degree_multiplier, ip.random_seed, G);
    ix.min_distance = G.min_distance;
    launch_correct(s, hdist.p, n * k, corr);  // index.rs:669-673
    ix.nb_idx = std::move(hidx);
    ix.nb_dist = std::move(hdist);
    // retained trees and the first tree's leaf order (index.rs:679-715)
    const int64_t nst = std::min<int64_t>(ip.n_search_trees, (int64_t)forest.trees.size());
    ix.trees.dim = dim;
    ix.trees.n = n;
    for (int64_t t = 0; t < nst; ++t) {  // own, exactly sized copies: the forest's shared pool is released
        const TreeDev& S = forest.trees[t];
        TreeDev T;
        T.n_nodes = T.cap_nodes = S.n_nodes;
        const size_t nn = (size_t)std::max<int64_t>(S.n_nodes, 1);
        T.hyper.alloc(nn * dim);
        T.offsets.alloc(nn);
        T.children.alloc(nn * 2);
        T.indices.alloc((size_t)std::max<int64_t>(n, 1));
        HIP_TRY(hipMemcpyAsync(T.hyper.p, S.hyper.p, (size_t)S.n_nodes * dim * 4, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(T.offsets.p, S.offsets.p, (size_t)S.n_nodes * 4, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(T.children.p, S.children.p, (size_t)S.n_nodes * 8, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(T.indices.p, S.indices.p, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
        ix.trees.trees.push_back(std::move(T));
    }
    HIP_TRY(hipStreamSynchronize(s));
    if (nst > 0) {
        vertex_order_from_tree(s, ix.trees.trees[0].indices.p, n, ix.vertex_order, ix.inv);
    } else {
        std::vector<int64_t> h(n);
        for (int64_t i = 0; i < n; ++i) h[i] = i;
        ix.vertex_order.alloc(std::max<int64_t>(n, 1));
        ix.inv.alloc(std::max<int64_t>(n, 1));
        HIP_TRY(hipMemcpyAsync(ix.vertex_order.p, h.data(), n * 8, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(ix.inv.p, h.data(), n * 8, hipMemcpyHostToDevice, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    ix.X.alloc((size_t)std::max<int64_t>(n, 1) * ld);
    gather_rows(s, X0.p, ld, ix.vertex_order.p, n, ix.X.p);
    search_graph_csr(s, G, ix.vertex_order.p, ix.inv.p, ix.indptr, ix.indices, ix.nnz);
    // trees: ids remapped (remap_indices, flat_tree.rs:329-334); hyperplanes re-laid with stride ld
    std::vector<const float*> hp, of;
    std::vector<const int32_t*> ch, tx;
    for (auto& T : ix.trees.trees) {
        remap_ids(s, T.indices.p, n, ix.inv.p);
        DBuf<float> h((size_t)std::max<int64_t>(T.n_nodes, 1) * ld);
        HIP_TRY(hipMemsetAsync(h.p, 0, h.bytes(), s));
        if (T.n_nodes > 0)
            HIP_TRY(hipMemcpy2DAsync(h.p, ld * 4, T.hyper.p, dim * 4, dim * 4, T.n_nodes, hipMemcpyDeviceToDevice, s));
        hp.push_back(h.p);
        of.push_back(T.offsets.p);
        ch.push_back(T.children.p);
        tx.push_back(T.indices.p);
        ix.hyper_padded.push_back(std::move(h));
    }
    const size_t nt = std::max<size_t>(hp.size(), 1);
    ix.d_hyper.alloc(nt);
    ix.d_off.alloc(nt);
    ix.d_ch.alloc(nt);
    ix.d_tix.alloc(nt);
    if (!hp.empty()) {
        HIP_TRY(hipMemcpyAsync(ix.d_hyper.p, hp.data(), hp.size() * 8, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(ix.d_off.p, of.data(), of.size() * 8, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(ix.d_ch.p, ch.data(), ch.size() * 8, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(ix.d_tix.p, tx.data(), tx.size() * 8, hipMemcpyHostToDevice, s));
    }
    ix.status.alloc(1);
    HIP_TRY(hipMemsetAsync(ix.status.p, 0, 4, s));
    HIP_TRY(hipStreamSynchronize(s));
}

static void ensure_query_workspace(rinnd_index& ix, int slots) {
    hipStream_t s = ix.ctx->stream;
    if (slots <= ix.slots) return;
    ix.visited.alloc((size_t)slots * std::max<int64_t>(ix.n, 1));
    ix.epochs.alloc(slots);
    HIP_TRY(hipMemsetAsync(ix.visited.p, 0, ix.visited.bytes(), s));
    HIP_TRY(hipMemsetAsync(ix.epochs.p, 0, ix.epochs.bytes(), s));
    ix.seed_cap = std::max<int64_t>(ix.n, 1);
    ix.spill_d.alloc((size_t)slots * ix.seed_cap);
    ix.spill_i.alloc((size_t)slots * ix.seed_cap);
    ix.slots = slots;
}

// The query rows as the kernel reads them (stride ld, zero padding): device rows that already have that
// stride are used in place, anything else is copied into the index's grow-only staging buffer.
static const float* stage_queries(rinnd_index& ix, const float* queries, int64_t nq) {
    hipStream_t s = ix.ctx->stream;
    const int64_t dim = ix.dim, ld = ix.ld;
    if (dim == ld && ((uintptr_t)queries & 255) == 0 && device_pointer(queries)) return queries;
    const size_t need = (size_t)nq * ld;
    if (need > ix.qbuf.n) {
        ix.qbuf.alloc(need);
        HIP_TRY(hipMemsetAsync(ix.qbuf.p, 0, need * 4, s));  // padding columns are never written again
    }
    HIP_TRY(hipMemcpy2DAsync(ix.qbuf.p, ld * 4, queries, dim * 4, dim * 4, nq, hipMemcpyDefault, s));
    return ix.qbuf.p;
}

void ensure_fp8_codes(rinnd_index& ix) {
    if (ix.q8.p) return;
    hipStream_t s = ix.ctx->stream;
    ix.ld8 = round_up(ix.dim, (int64_t)32);
    ix.q8.alloc((size_t)std::max<int64_t>(ix.n, 1) * ix.ld8);
    ix.q8s.alloc((size_t)std::max<int64_t>(ix.n, 1));
    ix.q8n.alloc((size_t)std::max<int64_t>(ix.n, 1));
    launch_fp8_encode(s, ix.X.p, ix.ld, (int)ix.dim, ix.n, ix.q8.p, ix.ld8, ix.q8s.p, ix.q8n.p);
    HIP_TRY(hipGetLastError());
}

// search_k: candidate width of the search (k for a plain query); rerank > 0: exact re-ranking of that many
// candidates; fp8: candidate distances from the FP8 codes.  k results are written per query.
// The index's part of a launch of the greedy search: data, graph, trees and the query parameters.
// The searches the index can still run: without its FP32 rows only the int8 search without reranking
// (QuantizedI8Index::ensure_rerank_available, crates/rinnd/src/lib.rs:287-300), and the int8 search only once
// the codes exist.
static void check_query_rep(const rinnd_index& ix, int rep, int64_t rerank) {
    if (rep == kQueryI8 && !ix.i8c.p)
        throw Error{RINND_ERR_INVALID_ARG, "the index has no int8 codes (rinnd_index_encode_i8)"};
    if (ix.data_released && rep == kQueryI8 && rerank > 0)
        throw Error{RINND_ERR_INVALID_ARG, "exact reranking is unavailable because FP32 vector data was released"};
    if (ix.data_released && rep != kQueryI8)
        throw Error{RINND_ERR_INVALID_ARG, "FP32 vector data was released: only the int8 search is available"};
}

SearchArgs index_search_args(const rinnd_index& ix, const float* qd, int64_t nq, int64_t k, float eps,
                                    int64_t qid_base, int64_t search_k, int64_t rerank, int rep, int32_t* oip,
                                    float* odp) {
    check_query_rep(ix, rep, rerank);
    const bool fp8 = rep == kQueryFP8;
    SearchArgs a{};
    a.metric = ix.metric;
    a.X = ix.X.p;
    a.ld = ix.ld;
    a.dim = (int)ix.dim;
    a.n = ix.n;
    a.indptr = ix.indptr.p;
    a.indices = ix.indices.p;
    a.n_trees = (int)ix.trees.trees.size();
    a.hyper = ix.tree_i8 ? nullptr : ix.d_hyper.p;
    a.hdim = ix.ld;
    a.offsets = ix.d_off.p;
    a.hq = ix.tree_i8 ? ix.d_tq.p : nullptr;
    a.hqs = ix.tree_i8 ? ix.d_tqs.p : nullptr;
    a.children = ix.d_ch.p;
    a.tindices = ix.d_tix.p;
    a.leaf_budget = (int)ix.leaf_budget;
    a.Q = qd;
    a.ldq = ix.ld;
    a.nq = nq;
    a.qid_base = qid_base;
    a.k = (int)search_k;
    a.eps = eps;
    a.min_distance = ix.min_distance;
    a.seed = ix.seed;
    a.vertex_order = ix.vertex_order.p;
    a.corr = ix.corr;
    a.out_idx = oip;
    a.out_dist = odp;
    a.q8 = fp8 ? ix.q8.p : nullptr;
    a.ld8 = ix.ld8;
    a.q8s = fp8 ? ix.q8s.p : nullptr;
    a.q8n = fp8 ? ix.q8n.p : nullptr;
    a.i8 = rep == kQueryI8 ? ix.i8c.p : nullptr;
    a.ldi8 = ix.ldi8;
    a.i8inv = rep == kQueryI8 ? ix.i8inv.p : nullptr;
    a.i8enc = ix.i8enc;
    a.i8gm = ix.i8gm;
    a.i8gd = ix.i8gd;
    a.seed_fixed = rep == kQueryI8;
    a.k_out = (int)k;
    a.rerank = (int)rerank;
    return a;
}

static void query_impl(rinnd_index& ix, const float* queries, int64_t nq, int64_t k, float eps, int64_t qid_base,
                       int32_t* out_idx, float* out_dist, int64_t search_k = -1, int64_t rerank = 0,
                       int rep = kQueryF32) {
    hipStream_t s = ix.ctx->stream;
    if (search_k < 0) search_k = k;
    RINND_CHECK_ARG(k >= 0 && k <= search_max_k(), "k must be in [0, 512]");
    RINND_CHECK_ARG(search_k >= k && search_k <= search_max_k(), "the candidate width must be in [k, 512]");
    check_query_rep(ix, rep, rerank);
    if (nq == 0 || k == 0) return;
    if (rep == kQueryFP8) {
        RINND_CHECK_ARG(round_up(ix.dim, (int64_t)32) <= 2048, "the FP8 candidate search supports dim <= 2048");
        ensure_fp8_codes(ix);
    }
    // workspace: one slot per concurrent query (a wave; up to 16 are resident per CU); memory per slot is n
    // bytes of visited epochs + 8n bytes of seed-heap spill (the heap never holds more than n entries), at
    // most 32 GB of the 288 GB in all
    const int64_t per_slot = ix.n * 9;
    int64_t max_slots = std::max<int64_t>(64, std::min<int64_t>(4096, ((int64_t)32 << 30) / std::max<int64_t>(per_slot, 1)));
    const int slots = (int)std::min<int64_t>(nq, max_slots);
    ensure_query_workspace(ix, std::max(slots, ix.slots));
    const float* qd = stage_queries(ix, queries, nq);
    const bool dev_out = device_pointer(out_idx) && device_pointer(out_dist);
    int32_t* oip = out_idx;
    float* odp = out_dist;
    if (!dev_out) {
        ix.obuf_i.ensure((size_t)nq * k);
        ix.obuf_d.ensure((size_t)nq * k);
        oip = ix.obuf_i.p;
        odp = ix.obuf_d.p;
    }
    SearchArgs a = index_search_args(ix, qd, nq, k, eps, qid_base, search_k, rerank, rep, oip, odp);
    a.visited = ix.visited.p;
    a.epochs = ix.epochs.p;
    a.seed_spill_d = ix.spill_d.p;
    a.seed_spill_i = ix.spill_i.p;
    a.seed_cap = ix.seed_cap;
    a.status = ix.status.p;
    launch_search(s, a, slots);
    HIP_TRY(hipGetLastError());
    if (!dev_out) {
        HIP_TRY(hipMemcpyAsync(out_idx, oip, (size_t)nq * k * 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(out_dist, odp, (size_t)nq * k * 4, hipMemcpyDeviceToHost, s));
    }
    int32_t st = 0;
    HIP_TRY(hipMemcpyAsync(&st, ix.status.p, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (st != 0) {
        HIP_TRY(hipMemsetAsync(ix.status.p, 0, 4, s));
        throw Error{RINND_ERR_INTERNAL, "search workspace capacity exceeded"};
    }
}

// ---- the query that allocates nothing ---------------------------------------------------------------
// A fixed device array holds the search's per-slot state (status word, visited epochs, visited bytes, seed-heap
// spill) and, when the query rows need padding, the padded rows.  Its device address is looked up once per
// device (noalloc_scratch_init, called when a context is created) and passed to the kernels; the array's name is
// never used as a device address in host code.  Every call on a device shares that device's array: calls are
// serialised by a per-device lock held until the call's stream work has finished (kernels of overlapping calls
// would race on it), and each call clears what must start at zero.
constexpr size_t kNoAllocScratchBytes = (size_t)128 << 20;  // 128 MiB per GPU, held until the process ends
constexpr int64_t kNoAllocSeedCap = 1 << 16;               // seed-heap entries per query slot
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static std::mutex g_scratch_init_mu;
static void* g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu[kMaxDevices];

hipError_t noalloc_scratch_init(int device) {  // with `device` current
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_init_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = p;
    return hipSuccess;
}

__global__ void k_zero_words(uint32_t* p, int64_t words) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < words; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}
static void zero_bytes(hipStream_t s, void* p, size_t bytes) {  // p and bytes multiples of 4; a kernel, no memset
    const int64_t words = (int64_t)(bytes / 4);
    if (words == 0) return;
    hipLaunchKernelGGL(k_zero_words, dim3((unsigned)std::min<int64_t>(ceil_div(words, (int64_t)256), 8192)),
                       dim3(256), 0, s, (uint32_t*)p, words);
}

static size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

static void query_noalloc_impl(rinnd_index& ix, const float* queries, int64_t nq, int64_t k, float eps,
                               int32_t* out_idx, float* out_dist) {
    RINND_CHECK_ARG(nq >= 0, "negative query count");
    RINND_CHECK_ARG(k >= 0 && k <= search_max_k(), "k must be in [0, 512]");
    if (nq == 0 || k == 0) return;
    RINND_CHECK_ARG(device_pointer(queries) && device_pointer(out_idx) && device_pointer(out_dist),
                    "the no-allocation query takes device pointers for the queries and both outputs");
    const int dev = ix.ctx->device;
    RINND_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "device id out of range");
    if (!g_scratch_ptr[dev]) HIP_TRY(noalloc_scratch_init(dev));
    char* base = (char*)g_scratch_ptr[dev];
    hipStream_t s = ix.ctx->stream;
    // layout: status | epochs | padded query rows (if needed) | visited bytes | seed spill (d, i)
    const bool stage = !(ix.dim == ix.ld && ((uintptr_t)queries & 255) == 0);
    const size_t q_bytes = stage ? al256((size_t)nq * ix.ld * 4) : 0;
    const int64_t seed_cap = std::max<int64_t>(1, std::min<int64_t>(ix.n, kNoAllocSeedCap));
    const size_t per_slot = 4 + (size_t)ix.n + 8 * (size_t)seed_cap;
    const size_t fixed = 256 + q_bytes + 4 * 256;  // the status block, the query rows, rounding of four regions
    if (fixed + per_slot > kNoAllocScratchBytes) {
        throw Error{RINND_ERR_UNSUPPORTED,
                    "the no-allocation query needs " + std::to_string(fixed + per_slot) + " bytes for " +
                        std::to_string(nq) + " queries over " + std::to_string(ix.n) +
                        " points; its device array holds " + std::to_string(kNoAllocScratchBytes) +
                        " bytes (split the batch, or use rinnd_index_query)"};
    }
    const int slots = (int)std::min<int64_t>(std::min<int64_t>(nq, 4096),
                                              (int64_t)((kNoAllocScratchBytes - fixed) / per_slot));
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* p = base + off;
        off += al256(bytes);
        return p;
    };
    int32_t* status = (int32_t*)take(256);
    int32_t* epochs = (int32_t*)(base + 4);  // inside the status block when slots <= 63
    if (slots > 63) epochs = (int32_t*)take((size_t)slots * 4);
    float* qrows = stage ? (float*)take(q_bytes) : nullptr;
    uint8_t* visited = (uint8_t*)take((size_t)slots * ix.n);
    float* spd = (float*)take((size_t)slots * seed_cap * 4);
    int32_t* spi = (int32_t*)take((size_t)slots * seed_cap * 4);
    if (off > kNoAllocScratchBytes) throw Error{RINND_ERR_INTERNAL, "no-allocation scratch layout overflow"};
    std::lock_guard<std::mutex> lk(g_scratch_mu[dev]);
    // what must start at zero: the status word, the epochs and the visited bytes (an earlier call, maybe on
    // another index, left its own epochs there)
    zero_bytes(s, status, 256);
    if (slots > 63) zero_bytes(s, epochs, al256((size_t)slots * 4));
    zero_bytes(s, visited, al256((size_t)slots * ix.n));
    const float* qd = queries;
    if (stage) {
        launch_pad_rows(s, queries, nq, (int)ix.dim, qrows, ix.ld);  // zero padding columns, every call
        qd = qrows;
    }
    SearchArgs a = index_search_args(ix, qd, nq, k, eps, 0, k, 0, kQueryF32, out_idx, out_dist);
    a.visited = visited;
    a.epochs = epochs;
    a.seed_spill_d = spd;
    a.seed_spill_i = spi;
    a.seed_cap = seed_cap;
    a.status = status;
    launch_search(s, a, slots);
    HIP_TRY(hipGetLastError());
    int32_t st = 0;
    HIP_TRY(hipMemcpyAsync(&st, status, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (st == 1)
        throw Error{RINND_ERR_UNSUPPORTED, "a query's frontier exceeded the no-allocation limit of " +
                                               std::to_string(kNoAllocSeedCap) + " entries (use rinnd_index_query)"};
    if (st != 0) throw Error{RINND_ERR_INTERNAL, "search workspace capacity exceeded"};
}

template <class F>
static rinnd_status guard(rinnd_ctx* ctx, F&& f) {
    if (!ctx) return RINND_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        f();
        ctx->last_error.clear();
        return RINND_OK;
    } catch (const Error& e) {
        ctx->last_error = e.msg;
        (void)hipGetLastError();
        return e.code;
    } catch (const std::bad_alloc&) {
        ctx->last_error = "host allocation failed";
        return RINND_ERR_OOM;
    } catch (const std::exception& e) {
        ctx->last_error = e.what();
        return RINND_ERR_INTERNAL;
    }
}

}  // namespace rinnd

extern "C" {

rinnd_status rinnd_search_graph_build(rinnd_ctx* ctx, const int32_t* sorted_idx, const float* sorted_dist,
                                      const float* data, int64_t n, int64_t k, int64_t dim, rinnd_distance dist,
                                      float diversify_prob, float pruning_degree_multiplier, int32_t* out_indptr,
                                      int32_t* out_indices, int64_t* out_nnz, float* out_min_distance) {
    return guard(ctx, [&] {
        RINND_CHECK_ARG(n >= 1 && k >= 1 && dim >= 1, "invalid sizes");
        hipStream_t s = ctx->stream;
        DBuf<float> X;
        int64_t ld = 0;
        make_padded(s, data, n, (int)dim, X, ld);
        DBuf<int32_t> di((size_t)n * k);
        DBuf<float> dd((size_t)n * k);
        HIP_TRY(hipMemcpyAsync(di.p, sorted_idx, (size_t)n * k * 4, hipMemcpyDefault, s));
        HIP_TRY(hipMemcpyAsync(dd.p, sorted_dist, (size_t)n * k * 4, hipMemcpyDefault, s));
        SearchGraphDev G;
        build_search_graph_dev(s, (int)dist, X.p, ld, (int)dim, di.p, dd.p, n, (int)k, diversify_prob,
                               pruning_degree_multiplier, 42, G);
        DBuf<int32_t> ip, ic;
        int64_t nnz = 0;
        search_graph_csr(s, G, nullptr, nullptr, ip, ic, nnz);
        HIP_TRY(hipMemcpyAsync(out_indptr, ip.p, (n + 1) * 4, hipMemcpyDefault, s));
        if (nnz) HIP_TRY(hipMemcpyAsync(out_indices, ic.p, nnz * 4, hipMemcpyDefault, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (out_nnz) *out_nnz = nnz;
        if (out_min_distance) *out_min_distance = G.min_distance;
    });
}

rinnd_status rinnd_index_build(rinnd_ctx* ctx, const float* data, int64_t n, int64_t dim, rinnd_distance dist,
                               rinnd_correction corr, const rinnd_index_params* params, rinnd_index** out,
                               rinnd_build_stats* stats) {
    return guard(ctx, [&] {
        RINND_CHECK_ARG(params && out, "null argument");
        RINND_CHECK_ARG(n >= 1 && dim >= 1, "the index needs at least one point");
        RINND_CHECK_ARG((int)dist >= 0 && (int)dist <= 6, "unknown distance");
        RINND_CHECK_ARG((int)corr >= 0 && (int)corr <= 3, "unknown correction");
        auto ix = std::make_unique<rinnd_index>();
        index_build_impl(ctx, data, n, dim, (int)dist, (int)corr, *params, *ix, stats);
        *out = ix.release();
    });
}

void rinnd_index_free(rinnd_index* index) {
    if (!index) return;
    (void)hipSetDevice(index->ctx->device);
    (void)hipStreamSynchronize(index->ctx->stream);
    delete index;
}

rinnd_status rinnd_index_query(rinnd_index* ix, const float* queries, int64_t nq, int64_t k, float epsilon,
                               int32_t* out_idx, float* out_dist) {
    if (!ix) return RINND_ERR_INVALID_ARG;
    return guard(ix->ctx, [&] {
        RINND_CHECK_ARG(nq >= 0, "negative query count");
        query_impl(*ix, queries, nq, k, epsilon, 0, out_idx, out_dist);
    });
}

rinnd_status rinnd_index_query_noalloc(rinnd_index* ix, const float* queries, int64_t nq, int64_t k, float epsilon,
                                       int32_t* out_idx, float* out_dist) {
    if (!ix) return RINND_ERR_INVALID_ARG;
    return guard(ix->ctx, [&] { query_noalloc_impl(*ix, queries, nq, k, epsilon, out_idx, out_dist); });
}

rinnd_status rinnd_index_query_one(rinnd_index* ix, const float* query, int64_t k, float epsilon,
                                   int32_t* out_idx, float* out_dist) {
    if (!ix) return RINND_ERR_INVALID_ARG;
    return guard(ix->ctx, [&] { query_impl(*ix, query, 1, k, epsilon, 0, out_idx, out_dist); });
}

// resolve_quantized_candidate_width / should_rerank_quantized (core/index.rs:55-90) around a quantized search
static void quantized_query(rinnd_index& ix, const float* queries, int64_t nq, int64_t k, float epsilon,
                            int64_t candidate_width, int64_t rerank_width, int32_t* out_idx, float* out_dist, int rep) {
    RINND_CHECK_ARG(nq >= 0 && k >= 0 && candidate_width >= 0 && rerank_width >= 0, "negative size");
    const int64_t width = candidate_width == 0 ? std::max(k, rerank_width) : candidate_width;
    RINND_CHECK_ARG(width >= k, "quantized candidate width must be at least k");
    RINND_CHECK_ARG(rerank_width <= width, "quantized rerank width must not exceed candidate width");
    RINND_CHECK_ARG(candidate_width == 0 || rerank_width == 0 || rerank_width >= k,
                    "quantized rerank width must be zero or at least k");
    const bool rerank = candidate_width == 0 ? rerank_width > k : rerank_width > 0;
    query_impl(ix, queries, nq, k, epsilon, 0, out_idx, out_dist, width, rerank ? rerank_width : 0, rep);
}

rinnd_status rinnd_index_query_quantized(rinnd_index* ix, const float* queries, int64_t nq, int64_t k,
                                         float epsilon, int64_t candidate_width, int64_t rerank_width,
                                         int32_t* out_idx, float* out_dist) {
    if (!ix) return RINND_ERR_INVALID_ARG;
    return guard(ix->ctx, [&] {
        quantized_query(*ix, queries, nq, k, epsilon, candidate_width, rerank_width, out_idx, out_dist, kQueryFP8);
    });
}

rinnd_status rinnd_index_encode_i8(rinnd_index* ix, int encoding) {
    if (!ix) return RINND_ERR_INVALID_ARG;
    return guard(ix->ctx, [&] {
        RINND_CHECK_ARG(encoding >= 0 && encoding <= 2, "encoding must be 0 (int8), 1 (sq8u) or 2 (sq8p)");
        RINND_CHECK_ARG(ix->metric == RINND_DIST_DIRECT_COSINE,
                        "the quantized codes belong to a directly normalized cosine index");
        RINND_CHECK_ARG(!ix->data_released, "FP32 vector data was released");
        RINND_CHECK_ARG(round_up(ix->dim, (int64_t)32) <= 2048, "the int8 candidate search supports dim <= 2048");
        hipStream_t s = ix->ctx->stream;
        // sq8u's global multiplier 127 / max|data| and dequantisation scale (lib.rs:171-185), in host float
        float gm = 0.f, gd = 0.f;
        if (encoding == 1) {
            DBuf<unsigned> mx(1);
            HIP_TRY(hipMemsetAsync(mx.p, 0, 4, s));
            launch_abs_max(s, ix->X.p, ix->ld, (int)ix->dim, ix->n, mx.p);
            unsigned bits = 0;
            HIP_TRY(hipMemcpyAsync(&bits, mx.p, 4, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            float gmax;
            std::memcpy(&gmax, &bits, 4);
            gm = gmax > 0.f ? 127.0f / gmax : 0.f;
            gd = gm > 0.f ? 1.0f / gm : 0.f;
        }
        ix->ldi8 = round_up(ix->dim, (int64_t)32);
        ix->i8c.alloc((size_t)std::max<int64_t>(ix->n, 1) * ix->ldi8);
        ix->i8inv.alloc((size_t)std::max<int64_t>(ix->n, 1));
        launch_i8_encode(s, ix->X.p, ix->ld, (int)ix->dim, ix->n, encoding, gm, gd, ix->i8c.p, ix->ldi8, ix->i8inv.p);
        HIP_TRY(hipStreamSynchronize(s));
        ix->i8enc = encoding;
        ix->i8gm = gm;
        ix->i8gd = gd;
    });
}

rinnd_status rinnd_index_query_i8(rinnd_index* ix, const float* queries, int64_t nq, int64_t k, float epsilon,
                                  int64_t candidate_width, int64_t rerank_width, int32_t* out_idx, float* out_dist) {
    if (!ix) return RINND_ERR_INVALID_ARG;
    return guard(ix->ctx, [&] {
        quantized_query(*ix, queries, nq, k, epsilon, candidate_width, rerank_width, out_idx, out_dist, kQueryI8);
    });
}

rinnd_status rinnd_index_release_data(rinnd_index* ix) {
    if (!ix) return RINND_ERR_INVALID_ARG;
    return guard(ix->ctx, [&] {
        RINND_CHECK_ARG(ix->i8c.p, "the index has no int8 codes (rinnd_index_encode_i8)");
        HIP_TRY(hipStreamSynchronize(ix->ctx->stream));
        ix->X.release();
        ix->q8.release();
        ix->q8s.release();
        ix->q8n.release();
        ix->data_released = true;
    });
}

rinnd_status rinnd_index_quantize_trees(rinnd_index* ix) {
    if (!ix) return RINND_ERR_INVALID_ARG;
    return guard(ix->ctx, [&] {
        if (ix->tree_i8) return;
        hipStream_t s = ix->ctx->stream;
        const size_t T = ix->trees.trees.size();
        ix->tq_codes.resize(T);
        ix->tq_scales.resize(T);
        std::vector<const int8_t*> tq(T);
        std::vector<const float*> tqs(T);
        for (size_t t = 0; t < T; ++t) {
            const int64_t nn = std::max<int64_t>(ix->trees.trees[t].n_nodes, 1);
            ix->tq_codes[t].alloc((size_t)nn * ix->ld);
            ix->tq_scales[t].alloc((size_t)nn);
            launch_quantize_tree(s, ix->hyper_padded[t].p, ix->trees.trees[t].n_nodes, ix->ld, (int)ix->dim,
                                 ix->tq_codes[t].p, ix->tq_scales[t].p);
            tq[t] = ix->tq_codes[t].p;
            tqs[t] = ix->tq_scales[t].p;
        }
        ix->d_tq.alloc(std::max<size_t>(T, 1));
        ix->d_tqs.alloc(std::max<size_t>(T, 1));
        if (T) {
            HIP_TRY(hipMemcpyAsync(ix->d_tq.p, tq.data(), T * sizeof(void*), hipMemcpyHostToDevice, s));
            HIP_TRY(hipMemcpyAsync(ix->d_tqs.p, tqs.data(), T * sizeof(void*), hipMemcpyHostToDevice, s));
        }
        HIP_TRY(hipStreamSynchronize(s));  // the host tables above go out of scope
        // the FP32 hyperplanes are released as the reference takes its FP32 search trees (lib.rs:228-249);
        // offsets, children and leaf indices stay: they are the int8 trees' own
        for (size_t t = 0; t < T; ++t) {
            ix->hyper_padded[t].release();
            ix->trees.trees[t].hyper.release();
        }
        ix->d_hyper.release();
        ix->tree_i8 = true;
    });
}

rinnd_status rinnd_index_neighbor_graph(const rinnd_index* ix, int32_t* out_idx, float* out_dist) {
    if (!ix) return RINND_ERR_INVALID_ARG;
    return guard(ix->ctx, [&] {
        hipStream_t s = ix->ctx->stream;
        const size_t cnt = (size_t)ix->n * ix->k;
        if (out_idx) HIP_TRY(hipMemcpyAsync(out_idx, ix->nb_idx.p, cnt * 4, hipMemcpyDefault, s));
        if (out_dist) HIP_TRY(hipMemcpyAsync(out_dist, ix->nb_dist.p, cnt * 4, hipMemcpyDefault, s));
        HIP_TRY(hipStreamSynchronize(s));
    });
}

rinnd_status rinnd_index_search_graph_info(const rinnd_index* ix, int64_t* n, int64_t* nnz, float* min_distance) {
    if (!ix) return RINND_ERR_INVALID_ARG;
    if (n) *n = ix->n;
    if (nnz) *nnz = ix->nnz;
    if (min_distance) *min_distance = ix->min_distance;
    return RINND_OK;
}

// search_graph_original_order (core/index.rs:137-141): the internal CSR reordered by inv
rinnd_status rinnd_index_search_graph(const rinnd_index* ix, int32_t* indptr, int32_t* indices) {
    if (!ix) return RINND_ERR_INVALID_ARG;
    return guard(ix->ctx, [&] {
        hipStream_t s = ix->ctx->stream;
        const int64_t n = ix->n;
        std::vector<int32_t> p(n + 1), c(std::max<int64_t>(ix->nnz, 1));
        std::vector<int64_t> inv(n), order(n);
        HIP_TRY(hipMemcpyAsync(p.data(), ix->indptr.p, (n + 1) * 4, hipMemcpyDeviceToHost, s));
        if (ix->nnz) HIP_TRY(hipMemcpyAsync(c.data(), ix->indices.p, ix->nnz * 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(order.data(), ix->vertex_order.p, n * 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        // new (original) row o = internal row inv_orig[o]; ids mapped through vertex_order
        for (int64_t i = 0; i < n; ++i) inv[order[i]] = i;
        std::vector<int32_t> op(n + 1, 0), oc;
        oc.reserve(ix->nnz);
        for (int64_t o = 0; o < n; ++o) {
            const int64_t r = inv[o];
            for (int32_t t = p[r]; t < p[r + 1]; ++t) oc.push_back((int32_t)order[c[t]]);
            op[o + 1] = (int32_t)oc.size();
        }
        std::memcpy(indptr, op.data(), (n + 1) * 4);
        if (!oc.empty()) std::memcpy(indices, oc.data(), oc.size() * 4);
    });
}

rinnd_status rinnd_index_vertex_order(const rinnd_index* ix, int64_t* out) {
    if (!ix) return RINND_ERR_INVALID_ARG;
    return guard(ix->ctx, [&] {
        HIP_TRY(hipMemcpyAsync(out, ix->vertex_order.p, ix->n * 8, hipMemcpyDefault, ix->ctx->stream));
        HIP_TRY(hipStreamSynchronize(ix->ctx->stream));
    });
}

const rinnd_forest* rinnd_index_trees(const rinnd_index* ix) { return ix && !ix->tree_i8 ? &ix->trees : nullptr; }

}  // extern "C"
