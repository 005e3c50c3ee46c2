// batch.hip - many independent NN-Descent builds in one pass over the device (rinnd_nn_descent_batch).
//
// The reference builds one index per call (core/index.rs:601-733 runs nn_descent, core/nndescent/mod.rs:79-197,
// once per index).  Here the indexes' rows are concatenated and every heap holds ids in the concatenated
// numbering; an index's edges never leave it, so candidate building, the local join and the ordered apply run
// once over all rows per pass.  What stays per index is what the reference keeps per build:
//   - its FastRng main stream: the tree seeds (builder.rs:100), then two draws per valid edge per iteration
//     (candidates.rs:84-131), drawn at the index's own offset and jumped on the device;
//   - its trees: one work queue of (index, tree) items, largest first, for the persistent DFS kernel; the
//     leaves of tree t of every index are initialised together (initialize_from_leaves, mod.rs:203-300);
//   - its convergence test (mod.rs:137-167): it stops after an iteration with at most (delta*k*n) heap changes,
//     or after n_iters; from then on its rows count no edges, draw nothing and receive nothing.
// Each index's heap therefore equals the one rinnd_nn_descent returns for it alone (reference draws, one
// candidate thread).  The candidate phase and the local join of every pass after the first replay one captured
// hipGraph; the first pass runs eagerly and times the join tile shapes on slices of its rows.
#include <algorithm>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "internal.h"

namespace rinnd {

// ---- per-index device steps ------------------------------------------------------------------------
// Index s's candidate stream for this pass: 2 draws per valid edge of its rows, written at twice its first
// valid edge's rank; its main state then jumps past them.  A finished index draws nothing.
__global__ void k_seg_streams(const uint64_t* __restrict__ rvp, const int64_t* __restrict__ seg_off, int nseg,
                              const uint8_t* __restrict__ seg_active, uint64_t* main_states,
                              const uint64_t* __restrict__ tables, RngStream* streams) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    uint64_t st[4];
    for (int w = 0; w < 4; ++w) st[w] = main_states[s * 4 + w];
    RngStream r;
    for (int w = 0; w < 4; ++w) r.s[w] = st[w];
    r.count = 0;
    r.out_offset = 0;
    if (seg_active[s]) {
        const uint64_t a = rvp[seg_off[s]], b = rvp[seg_off[s + 1]];
        r.count = (int64_t)(2 * (b - a));
        r.out_offset = (int64_t)(2 * a);
        xoshiro_jump(tables, st, 2 * (b - a));
        for (int w = 0; w < 4; ++w) main_states[s * 4 + w] = st[w];
    }
    streams[s] = r;
}

// Index s's convergence test after a pass (mod.rs:155-160): its heap changes, summed over its rows, against
// its threshold; it stops there, or when it has run its n_iters.  totals[0] counts the indexes still running,
// totals[1] sums the changes.
__global__ __launch_bounds__(256) void k_seg_converge(const uint32_t* __restrict__ row_changes,
                                                      const int64_t* __restrict__ seg_off,
                                                      const int64_t* __restrict__ thr,
                                                      const int32_t* __restrict__ max_iters, int32_t* seg_it,
                                                      int32_t* seg_conv, uint8_t* seg_active, uint8_t* row_on,
                                                      unsigned long long* totals) {
    const int s = blockIdx.x;
    if (!seg_active[s]) return;  // uniform over the workgroup
    const int64_t r0 = seg_off[s], r1 = seg_off[s + 1];
    unsigned long long c = 0;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) c += row_changes[r];
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    __shared__ unsigned long long part[4];
    __shared__ int stop;
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long sum = part[0] + part[1] + part[2] + part[3];
        const int it = seg_it[s] + 1;
        seg_it[s] = it;
        const bool conv = (int64_t)sum <= thr[s];
        seg_conv[s] = conv ? 1 : 0;
        const bool done = conv || it >= max_iters[s];
        stop = done ? 1 : 0;
        if (done) seg_active[s] = 0;
        atomicAdd(&totals[1], sum);
        if (!done) atomicAdd(&totals[0], 1ull);
    }
    __syncthreads();
    if (stop)
        for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) row_on[r] = 0;
}

// rows of index s (blockIdx.y) on or off with it
__global__ void k_seg_rows_on(const int64_t* __restrict__ seg_off, const uint8_t* __restrict__ seg_active,
                              uint8_t* row_on) {
    const int s = blockIdx.y;
    const int64_t r0 = seg_off[s], r1 = seg_off[s + 1];
    const uint8_t v = seg_active[s];
    for (int64_t r = r0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < r1; r += (int64_t)gridDim.x * blockDim.x)
        row_on[r] = v;
}

// concatenated ids -> each index's own ids (blockIdx.y: the index)
__global__ void k_seg_localize(int32_t* idx, const int64_t* __restrict__ seg_off, int k) {
    const int s = blockIdx.y;
    const int64_t b = seg_off[s], e0 = b * k, e1 = seg_off[s + 1] * k;
    for (int64_t t = e0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < e1; t += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = idx[t];
        if (v >= 0) idx[t] = v - (int32_t)b;
    }
}

// ---- memset nodes of a captured graph -> fill kernels ------------------------------------------------
// Library calls (the radix sorts) record memset nodes when captured; those are replaced by an equivalent fill
// kernel node with the same dependencies, so the graph holds kernels and copies only.
__global__ void k_memset_node(char* dst, uint32_t value, int esize, int64_t width, int64_t height, int64_t pitch) {
    const int64_t total = width * height;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t y = t / width, x = t - y * width;
        char* p = dst + y * pitch + x * esize;
        if (esize == 1) *(uint8_t*)p = (uint8_t)value;
        else if (esize == 2) *(uint16_t*)p = (uint16_t)value;
        else *(uint32_t*)p = value;
    }
}

static int graph_memsets_to_kernels(hipGraph_t g) {
    size_t nn = 0;
    HIP_TRY(hipGraphGetNodes(g, nullptr, &nn));
    std::vector<hipGraphNode_t> nodes(nn);
    if (nn) HIP_TRY(hipGraphGetNodes(g, nodes.data(), &nn));
    int replaced = 0;
    for (hipGraphNode_t node : nodes) {
        hipGraphNodeType t;
        HIP_TRY(hipGraphNodeGetType(node, &t));
        if (t != hipGraphNodeTypeMemset) continue;
        hipMemsetParams mp{};
        HIP_TRY(hipGraphMemsetNodeGetParams(node, &mp));
        if (mp.elementSize != 1 && mp.elementSize != 2 && mp.elementSize != 4)
            throw Error{RINND_ERR_INTERNAL, "unexpected memset element size in a captured graph"};
        size_t nd = 0, no = 0;
        HIP_TRY(hipGraphNodeGetDependencies(node, nullptr, &nd));
        std::vector<hipGraphNode_t> deps(nd);
        if (nd) HIP_TRY(hipGraphNodeGetDependencies(node, deps.data(), &nd));
        HIP_TRY(hipGraphNodeGetDependentNodes(node, nullptr, &no));
        std::vector<hipGraphNode_t> outs(no);
        if (no) HIP_TRY(hipGraphNodeGetDependentNodes(node, outs.data(), &no));
        char* dst = (char*)mp.dst;
        uint32_t value = mp.value;
        int esize = (int)mp.elementSize;
        int64_t width = (int64_t)mp.width, height = std::max<int64_t>((int64_t)mp.height, 1);
        int64_t pitch = (int64_t)mp.pitch;
        void* args[] = {&dst, &value, &esize, &width, &height, &pitch};
        hipKernelNodeParams kp{};
        kp.func = reinterpret_cast<void*>(k_memset_node);
        kp.gridDim = dim3(grid_for(std::max<int64_t>(width * height, 1), 256));
        kp.blockDim = dim3(256);
        kp.sharedMemBytes = 0;
        kp.kernelParams = args;
        kp.extra = nullptr;
        hipGraphNode_t kn;
        HIP_TRY(hipGraphAddKernelNode(&kn, g, deps.empty() ? nullptr : deps.data(), deps.size(), &kp));
        for (hipGraphNode_t o : outs) HIP_TRY(hipGraphAddDependencies(g, &kn, &o, 1));
        HIP_TRY(hipGraphDestroyNode(node));
        ++replaced;
    }
    return replaced;
}

struct GraphExec {
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    ~GraphExec() {
        if (x) (void)hipGraphExecDestroy(x);
        if (g) (void)hipGraphDestroy(g);
    }
};

// ---- forest: one work queue of (index, tree) items ---------------------------------------------------
struct BatchTree {
    int seg = 0, tree = 0;
    int64_t n = 0, base = 0, cap = 0, n_nodes = 0;
    int32_t* children = nullptr;
    int32_t* indices = nullptr;
};

static void build_forest_batch(hipStream_t s, const float* DX, int64_t ld, int dim, const std::vector<int64_t>& off,
                               const std::vector<int64_t>& n_trees, int64_t leaf_size, int angular, int64_t max_depth,
                               std::vector<Xoshiro>& mains, std::vector<BatchTree>& trees, DBuf<char>& pool) {
    const int nseg = (int)n_trees.size();
    std::vector<uint64_t> states;
    trees.clear();
    for (int sg = 0; sg < nseg; ++sg)
        for (int64_t t = 0; t < n_trees[sg]; ++t) {  // build_rp_forest seeds, builder.rs:100
            if (off[sg + 1] == off[sg]) break;  // an empty index (the reference returns before the forest)
            const uint64_t seed = mains[sg].next_u64();
            const Xoshiro r = xoshiro_seed(seed);
            for (int w = 0; w < 4; ++w) states.push_back(r.s[w]);
            BatchTree bt;
            bt.seg = sg;
            bt.tree = (int)t;
            bt.base = off[sg];
            bt.n = off[sg + 1] - off[sg];
            trees.push_back(bt);
        }
    const int nt = (int)trees.size();
    if (nt == 0) return;
    RINND_CHECK_ARG(max_depth <= 500, "max_depth above 500 is not supported by the GPU tree builder");
    RINND_CHECK_ARG(dim <= 8192, "dim above 8192 is not supported by the GPU tree builder");
    const int64_t ls = std::max<int64_t>(leaf_size, 1);
    int64_t max_n = 0;
    for (auto& t : trees) max_n = std::max(max_n, t.n);
    // largest trees first: they start early instead of trailing
    std::vector<int32_t> order(nt);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return trees[a].n > trees[b].n; });
    const int grid = std::min(nt, forest_queue_grid(dim));
    const int64_t stride = max_n + 1;
    DBuf<uint8_t> side((size_t)grid * stride);
    DBuf<int32_t> rankR((size_t)grid * stride), PQ((size_t)grid * 2 * stride), tmp((size_t)grid * stride);
    DBuf<float*> d_hp(nt), d_of(nt);
    DBuf<int32_t*> d_ch(nt), d_ix(nt);
    DBuf<int64_t> d_cap(nt), d_nn(nt), d_base(nt), d_n(nt);
    DBuf<int32_t> d_status(nt), d_order(nt), d_queue(1);
    DBuf<uint64_t> d_states((size_t)nt * 4);
    HIP_TRY(hipMemcpyAsync(d_states.p, states.data(), states.size() * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_order.p, order.data(), nt * 4, hipMemcpyHostToDevice, s));
    std::vector<int64_t> hb(nt), hn(nt);
    for (int i = 0; i < nt; ++i) {
        hb[i] = trees[i].base;
        hn[i] = trees[i].n;
    }
    HIP_TRY(hipMemcpyAsync(d_base.p, hb.data(), nt * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_n.p, hn.data(), nt * 8, hipMemcpyHostToDevice, s));
    bool big = false;
    for (int attempt = 0; attempt < 2; ++attempt) {
        auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
        size_t total = 0;
        for (auto& t : trees) {
            t.cap = big ? 2 * t.n + 1 : std::min<int64_t>(2 * t.n + 1, 8 * ceil_div(t.n, ls) + 1024);
            total += al((size_t)t.cap * dim * 4) + al((size_t)t.cap * 4) + al((size_t)t.cap * 8) +
                     al((size_t)std::max<int64_t>(t.n, 1) * 4);
        }
        pool.ensure(total);
        std::vector<float*> hp(nt), of(nt);
        std::vector<int32_t*> chp(nt), ixp(nt);
        std::vector<int64_t> caps(nt);
        char* z = pool.p;
        for (int i = 0; i < nt; ++i) {
            BatchTree& t = trees[i];
            hp[i] = (float*)z;
            z += al((size_t)t.cap * dim * 4);
            of[i] = (float*)z;
            z += al((size_t)t.cap * 4);
            chp[i] = (int32_t*)z;
            z += al((size_t)t.cap * 8);
            ixp[i] = (int32_t*)z;
            z += al((size_t)std::max<int64_t>(t.n, 1) * 4);
            t.children = chp[i];
            t.indices = ixp[i];
            caps[i] = t.cap;
        }
        HIP_TRY(hipMemcpyAsync(d_hp.p, hp.data(), nt * sizeof(float*), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(d_of.p, of.data(), nt * sizeof(float*), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(d_ch.p, chp.data(), nt * sizeof(int32_t*), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(d_ix.p, ixp.data(), nt * sizeof(int32_t*), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(d_cap.p, caps.data(), nt * 8, hipMemcpyHostToDevice, s));
        launch_zero(s, d_queue.p, 4);
        ForestDFSArgs a{};
        a.X = DX;
        a.ld = ld;
        a.dim = dim;
        a.n = max_n;
        a.leaf_size = leaf_size;
        a.max_depth = max_depth;
        a.angular = angular;
        a.n_trees = nt;
        a.tree_states = d_states.p;
        a.hyper = d_hp.p;
        a.offsets = d_of.p;
        a.children = d_ch.p;
        a.indices = d_ix.p;
        a.cap_nodes = d_cap.p;
        a.n_nodes_out = d_nn.p;
        a.status = d_status.p;
        a.side = side.p;
        a.rankR = rankR.p;
        a.PQ = PQ.p;
        a.tmp = tmp.p;
        a.tree_base = d_base.p;
        a.tree_n = d_n.p;
        a.order = d_order.p;
        a.queue = d_queue.p;
        a.slot_stride = stride;
        a.grid = grid;
        launch_forest_dfs(s, a);
        HIP_TRY(hipGetLastError());
        std::vector<int64_t> nn(nt);
        std::vector<int32_t> st(nt);
        HIP_TRY(hipMemcpyAsync(nn.data(), d_nn.p, nt * 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(st.data(), d_status.p, nt * 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        bool overflow = false;
        for (int i = 0; i < nt; ++i) {
            if (st[i] == 2) throw Error{RINND_ERR_INTERNAL, "RP tree stack overflow"};
            if (st[i] == 1) overflow = true;
            trees[i].n_nodes = nn[i];
        }
        if (!overflow) return;
        big = true;  // every split leaves two non-empty children: at most 2n-1 nodes
    }
    throw Error{RINND_ERR_INTERNAL, "RP tree node capacity exceeded"};
}

// ---- the batched build --------------------------------------------------------------------------------
enum : int { kWsBatch = 150 };

static double elapsed_ms_ev(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        (void)hipGetLastError();
        return 0.0;
    }
    return (double)ms;
}

void nn_descent_batch_dev(rinnd_ctx* ctx, const float* DX, int64_t ld, const std::vector<int64_t>& off, int dim,
                          int metric, const rinnd_params& p, const std::vector<int64_t>& n_trees,
                          const std::vector<int64_t>& n_iters, std::vector<Xoshiro>& mains, int32_t* w_idx,
                          float* w_dist, uint8_t* w_flags, bool use_graph, std::vector<int32_t>& iterations,
                          std::vector<int32_t>& converged, rinnd_batch_stats& st) {
    hipStream_t s = ctx->stream;
    Workspace& W = *ctx->ws;
    hipEvent_t* ev = ctx->ev;
    const int nseg = (int)n_trees.size();
    const int64_t N = off[nseg];
    const int k = (int)p.n_neighbors;
    const int C = (int)std::min<int64_t>(std::min<int64_t>(p.max_candidates, 60), p.n_neighbors);
    st = rinnd_batch_stats{};
    iterations.assign(nseg, 0);
    converged.assign(nseg, 0);
    int64_t max_n = 0;
    for (int sg = 0; sg < nseg; ++sg) max_n = std::max(max_n, off[sg + 1] - off[sg]);

    HIP_TRY(hipEventRecord(ev[0], s));
    std::vector<BatchTree> trees;
    DBuf<char> forest_pool;
    build_forest_batch(s, DX, ld, dim, off, n_trees, p.leaf_size, p.angular, p.max_depth, mains, trees, forest_pool);
    HIP_TRY(hipEventRecord(ev[1], s));

    launch_fill_heap(s, w_idx, w_dist, w_flags, N * k);
    // ---- leaves: tree t of every index that has one, together ----
    {
        int64_t max_t = 0;
        for (auto v : n_trees) max_t = std::max(max_t, v);
        WBuf<int32_t> mark(W, kWsBatch + 0, N + 1), end_at(W, kWsBatch + 1, N + 1), vidx(W, kWsBatch + 2, N + 1);
        WBuf<int64_t> rank(W, kWsBatch + 3, 2 * (N + 1)), leaves(W, kWsBatch + 4, 2 * (N + 1));
        const size_t lt_bytes = std::max(scan_u64_tmp_bytes(N + 2), (size_t)16);
        WBuf<char> ltmp(W, kWsBatch + 5, lt_bytes);
        WBuf<float> dscr(W, kWsBatch + 6, std::max<int64_t>(leaf_init_scratch_floats(N, ld, p.leaf_size), 1));
        WBuf<unsigned long long> lpairs(W, kWsBatch + 7, 2);
        WBuf<VTreeEnt> d_ents(W, kWsBatch + 8, std::max(nseg, 1));
        launch_zero(s, lpairs.p, 16);
        const LeafInitWs lws{mark.p, rank.p, dscr.p, lpairs.p, nullptr};
        const KTimer none{nullptr, nullptr, nullptr, nullptr};
        std::vector<std::vector<VTreeEnt>> ents(max_t);
        std::vector<int64_t> Nt(max_t, 0), span(max_t, 0);
        for (const BatchTree& t : trees) {
            VTreeEnt e{t.children, t.n_nodes, t.indices, Nt[t.tree], t.base, t.n};
            ents[t.tree].push_back(e);
            Nt[t.tree] += t.n;
            span[t.tree] = std::max(span[t.tree], std::max(t.n_nodes, t.n));
        }
        for (int64_t t = 0; t < max_t; ++t) {
            if (ents[t].empty()) continue;
            HIP_TRY(hipMemcpyAsync(d_ents.p, ents[t].data(), ents[t].size() * sizeof(VTreeEnt), hipMemcpyHostToDevice, s));
            launch_leaf_list_batch(s, d_ents.p, (int)ents[t].size(), span[t], Nt[t], mark.p, rank.p, end_at.p,
                                   leaves.p, vidx.p, ltmp.p, lt_bytes);
            launch_leaf_init(s, metric, DX, ld, dim, vidx.p, leaves.p, rank.p + Nt[t], Nt[t], p.leaf_size, w_idx, w_dist,
                             w_flags, k, lws, none, none, (int)t < kLeafFullTrees);
            HIP_TRY(hipStreamSynchronize(s));  // the entry table is rewritten for the next tree
        }
    }
    HIP_TRY(hipEventRecord(ev[2], s));

    // ---- per-index state on the device ----
    std::vector<int64_t> thr(nseg);
    std::vector<int32_t> maxit(nseg);
    std::vector<uint8_t> active(nseg);
    std::vector<uint64_t> hstates((size_t)nseg * 4);
    for (int sg = 0; sg < nseg; ++sg) {
        const int64_t n = off[sg + 1] - off[sg];
        thr[sg] = (int64_t)(p.delta * (float)p.n_neighbors * (float)n);  // mod.rs:157
        maxit[sg] = (int32_t)std::min<int64_t>(n_iters[sg], 1 << 30);
        active[sg] = n_iters[sg] > 0 && n > 0 ? 1 : 0;
        for (int w = 0; w < 4; ++w) hstates[sg * 4 + w] = mains[sg].s[w];
    }
    WBuf<int64_t> d_off(W, kWsBatch + 9, nseg + 1), d_thr(W, kWsBatch + 10, nseg);
    WBuf<int32_t> d_maxit(W, kWsBatch + 11, nseg), d_it(W, kWsBatch + 12, nseg), d_conv(W, kWsBatch + 13, nseg);
    WBuf<uint8_t> d_active(W, kWsBatch + 14, nseg), row_on(W, kWsBatch + 15, N);
    WBuf<uint64_t> d_main(W, kWsBatch + 16, (size_t)nseg * 4);
    WBuf<RngStream> d_streams(W, kWsBatch + 17, nseg);
    WBuf<unsigned long long> totals(W, kWsBatch + 18, 2);
    HIP_TRY(hipMemcpyAsync(d_off.p, off.data(), (nseg + 1) * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_thr.p, thr.data(), nseg * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_maxit.p, maxit.data(), nseg * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_active.p, active.data(), nseg, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_main.p, hstates.data(), hstates.size() * 8, hipMemcpyHostToDevice, s));
    launch_zero(s, d_it.p, nseg * 4);
    launch_zero(s, d_conv.p, nseg * 4);
    hipLaunchKernelGGL(k_seg_rows_on, dim3(grid_for(max_n, 256, 64), nseg), dim3(256), 0, s, d_off.p, d_active.p,
                       row_on.p);
    if (!ctx->jump_tables.p) {
        ctx->jump_tables.alloc((size_t)kJumpLevels * 1024);
        HIP_TRY(hipMemcpyAsync(ctx->jump_tables.p, host_jump_tables(), ctx->jump_tables.bytes(), hipMemcpyHostToDevice,
                               s));
    }

    // ---- pass buffers (fixed sizes: the captured graph holds their addresses) ----
    const uint64_t pair_bound = (uint64_t)C * (C - 1) / 2 + (uint64_t)C * C;  // per row: nn(nn-1)/2 + nn*no
    WBuf<int32_t> new_idx(W, kWsBatch + 20, (size_t)N * C), old_idx(W, kWsBatch + 21, (size_t)N * C);
    WBuf<unsigned long long> rvc(W, kWsBatch + 22, N + 1), rpc(W, kWsBatch + 23, N + 1), cstat(W, kWsBatch + 24, 2);
    WBuf<uint64_t> rvp(W, kWsBatch + 25, N + 1), rowoff(W, kWsBatch + 26, N + 1), koff(W, kWsBatch + 27, N + 1);
    WBuf<float> prio(W, kWsBatch + 28, (size_t)2 * N * k + 1), c_fpr(W, kWsBatch + 29, (size_t)N * k);
    WBuf<uint32_t> c_rkey(W, kWsBatch + 30, (size_t)N * k), c_rkey2(W, kWsBatch + 31, (size_t)N * k);
    WBuf<uint64_t> c_rval(W, kWsBatch + 32, (size_t)N * k), c_rval2(W, kWsBatch + 33, (size_t)N * k);
    int c_end_bit = 1;
    while (((int64_t)1 << c_end_bit) <= N) ++c_end_bit;
    WBuf<char> c_sort_tmp(W, kWsBatch + 34, std::max<size_t>(cand_sort_tmp_bytes((int64_t)N * k, c_end_bit), 16));
    WBuf<unsigned long long> ccnt(W, kWsBatch + 35, N + 1), kc(W, kWsBatch + 36, N + 1), krec(W, kWsBatch + 37, N + 1);
    WBuf<unsigned long long> kept(W, kWsBatch + 38, kKeptSlots), changes(W, kWsBatch + 39, 1);
    WBuf<uint64_t> coff(W, kWsBatch + 40, N + 1);
    WBuf<int32_t> order(W, kWsBatch + 41, std::max<int64_t>(N, 1));
    WBuf<uint32_t> bins(W, kWsBatch + 42, 128);
    const size_t scan_bytes = std::max(scan_u64_tmp_bytes(N + 1), (size_t)16);
    WBuf<char> scan_tmp(W, kWsBatch + 43, scan_bytes);
    WBuf<JRec> JU(W, kWsBatch + 44, std::max<uint64_t>((uint64_t)N * pair_bound, 1));
    WBuf<uint32_t> row_changes(W, kWsBatch + 45, std::max<int64_t>(N, 1));
    uint64_t* hst = (uint64_t*)W.status.ensure(64);  // pinned: [0] record count, [1..2] totals, [3..4] cstat
    const hipEvent_t hev = W.status_ev.get();
    HIP_TRY(hipMemsetAsync(krec.p, 0, (N + 1) * 8, s));  // krec[N] stays 0: the scan's total slot
    const JoinOut jo{rowoff.p, JU.p, kc.p, krec.p, kept.p};
    const int64_t maxc = 2 * max_n * k;

    int shape = 0;
    bool shape_known = false;
    {
        const uint64_t key = ((uint64_t)metric << 48) ^ ((uint64_t)ld << 16) ^ (uint64_t)C;
        const auto itj = ctx->join_shapes.find(key);
        if (itj != ctx->join_shapes.end()) {
            shape = itj->second;
            shape_known = true;
        }
    }
    // the candidate phase and the local join: fixed sizes, no host synchronisation
    auto phase_a = [&](bool tune) {
        launch_row_valid_counts(s, w_idx, N, k, (uint64_t*)rvc.p, row_on.p);
        scan_u64_exclusive(s, rvc.p, rvp.p, N + 1, scan_tmp.p, scan_bytes);
        hipLaunchKernelGGL(k_seg_streams, dim3(ceil_div((int64_t)nseg, (int64_t)64)), dim3(64), 0, s, rvp.p, d_off.p,
                           nseg, d_active.p, d_main.p, ctx->jump_tables.p, d_streams.p);
        launch_rng_streams(s, ctx->jump_tables.p, d_streams.p, nseg, maxc, prio.p);
        launch_fill_i32(s, new_idx.p, -1, N * C);
        launch_fill_i32(s, old_idx.p, -1, N * C);
        launch_zero(s, rpc.p, (N + 1) * 8);
        launch_zero(s, cstat.p, 16);
        CandBuildArgs ca{};
        ca.idx = w_idx;
        ca.flags = w_flags;
        ca.n = N;
        ca.k = k;
        ca.C = C;
        ca.mode = 0;
        ca.rvp = rvp.p;
        ca.prio = prio.p;
        ca.block_rows = N;
        ca.block_prio_base = nullptr;
        ca.counter_key = 0;
        ca.fpr = c_fpr.p;
        ca.rkey = c_rkey.p;
        ca.rkey2 = c_rkey2.p;
        ca.rval = c_rval.p;
        ca.rval2 = c_rval2.p;
        ca.cnt = ccnt.p;
        ca.off = coff.p;
        ca.order = order.p;
        ca.bins = bins.p;
        ca.sort_tmp = c_sort_tmp.p;
        ca.sort_tmp_bytes = c_sort_tmp.bytes();
        ca.scan_tmp = scan_tmp.p;
        ca.scan_tmp_bytes = scan_bytes;
        ca.new_idx = new_idx.p;
        ca.old_idx = old_idx.p;
        ca.row_pairs = rpc.p;
        ca.counts = cstat.p;
        ca.row_on = row_on.p;
        launch_cand_build(s, ca);
        scan_u64_exclusive(s, rpc.p, rowoff.p, N + 1, scan_tmp.p, scan_bytes);  // each row's pair region
        launch_zero(s, kept.p, kKeptSlots * 8);
        if (tune) {
            bool tuned = false;
            shape = launch_join_fp32_tuned(s, metric, DX, ld, dim, new_idx.p, old_idx.p, C, N, w_dist, k, jo, &tuned);
            if (tuned) {
                ctx->join_shapes[((uint64_t)metric << 48) ^ ((uint64_t)ld << 16) ^ (uint64_t)C] = shape;
                shape_known = true;
            }
        } else {
            launch_join_fp32(s, metric, DX, ld, dim, new_idx.p, old_idx.p, C, 0, N, w_dist, k, jo, shape);
        }
        scan_u64_exclusive(s, krec.p, koff.p, N + 1, scan_tmp.p, scan_bytes);
        HIP_TRY(hipMemcpyAsync(hst, koff.p + N, 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(hst + 3, cstat.p, 16, hipMemcpyDeviceToHost, s));
    };
    GraphExec G;
    int eb = 1;
    while (((int64_t)1 << eb) < N) ++eb;
    double ms_a = 0, ms_b = 0;
    int pass = 0;
    bool any = false;
    for (int sg = 0; sg < nseg; ++sg) any |= active[sg] != 0;
    while (any) {
        HIP_TRY(hipEventRecord(ev[3], s));
        if (pass == 0 || !use_graph) {
            phase_a(!shape_known);
        } else {
            if (!G.x) {
                HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
                try {
                    phase_a(false);
                } catch (...) {
                    hipGraph_t junk = nullptr;
                    (void)hipStreamEndCapture(s, &junk);
                    if (junk) (void)hipGraphDestroy(junk);
                    throw;
                }
                HIP_TRY(hipStreamEndCapture(s, &G.g));
                st.graph_memsets_replaced = graph_memsets_to_kernels(G.g);
                HIP_TRY(hipGraphInstantiate(&G.x, G.g, nullptr, nullptr, 0));
            }
            HIP_TRY(hipGraphLaunch(G.x, s));
            ++st.graph_replays;
        }
        HIP_TRY(hipEventRecord(ev[4], s));
        HIP_TRY(hipEventRecord(hev, s));
        HIP_TRY(hipEventSynchronize(hev));
        const uint64_t n_rec = hst[0];
        if (pass < 64) st.join_pairs[pass] = (int64_t)hst[4];
        // ordered apply (sizes from the record count), then each index's convergence test
        {
            const uint64_t items = std::max<uint64_t>(n_rec, 1);
            WBuf<uint32_t> jr_key(W, kWsBatch + 46, items), jr_key2(W, kWsBatch + 47, items);
            WBuf<uint64_t> jr_val(W, kWsBatch + 48, items), jr_val2(W, kWsBatch + 49, items);
            WBuf<char> jr_tmp(W, kWsBatch + 50, std::max<size_t>(join_sort_tmp_bytes((int64_t)items, eb), 16));
            launch_zero(s, row_changes.p, (size_t)N * 4);
            launch_zero(s, changes.p, 8);
            JoinApplyArgs ja{};
            ja.new_idx = new_idx.p;
            ja.old_idx = old_idx.p;
            ja.C = C;
            ja.n = N;
            ja.out = jo;
            ja.koff = koff.p;
            ja.n_records = n_rec;
            ja.rkey = jr_key.p;
            ja.rkey2 = jr_key2.p;
            ja.rval = jr_val.p;
            ja.rval2 = jr_val2.p;
            ja.off = coff.p;
            ja.order = order.p;
            ja.bins = bins.p;
            ja.sort_tmp = jr_tmp.p;
            ja.sort_tmp_bytes = jr_tmp.bytes();
            ja.scan_tmp = scan_tmp.p;
            ja.scan_tmp_bytes = scan_bytes;
            ja.idx = w_idx;
            ja.dist = w_dist;
            ja.flags = w_flags;
            ja.k = k;
            ja.changes = changes.p;
            ja.row_changes = row_changes.p;
            launch_join_apply(s, ja);
        }
        launch_zero(s, totals.p, 16);
        hipLaunchKernelGGL(k_seg_converge, dim3(nseg), dim3(256), 0, s, row_changes.p, d_off.p, d_thr.p, d_maxit.p,
                           d_it.p, d_conv.p, d_active.p, row_on.p, totals.p);
        HIP_TRY(hipMemcpyAsync(hst + 1, totals.p, 16, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipEventRecord(ev[5], s));
        HIP_TRY(hipEventRecord(hev, s));
        HIP_TRY(hipEventSynchronize(hev));
        ms_a += elapsed_ms_ev(ev[3], ev[4]);
        ms_b += elapsed_ms_ev(ev[4], ev[5]);
        if (pass < 64) st.changes[pass] = (int64_t)hst[2];
        ++pass;
        any = hst[1] > 0;
    }
    st.passes = pass;
    int tpr = 0, rpg = 0;
    join_fp32_shape_info(shape, &tpr, &rpg);
    st.join_threads_per_row = tpr;
    st.join_rows_per_group = rpg;
    st.ms_forest = elapsed_ms_ev(ev[0], ev[1]);
    st.ms_init = elapsed_ms_ev(ev[1], ev[2]);
    st.ms_candidates_join = ms_a;
    st.ms_apply = ms_b;
    // per-index results: iterations, convergence, the advanced main states; ids back to each index's own
    std::vector<int32_t> hit(nseg), hconv(nseg);
    HIP_TRY(hipMemcpyAsync(hit.data(), d_it.p, nseg * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(hconv.data(), d_conv.p, nseg * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(hstates.data(), d_main.p, hstates.size() * 8, hipMemcpyDeviceToHost, s));
    hipLaunchKernelGGL(k_seg_localize, dim3(grid_for(max_n * k, 256, 256), nseg), dim3(256), 0, s, w_idx, d_off.p, k);
    HIP_TRY(hipStreamSynchronize(s));
    for (int sg = 0; sg < nseg; ++sg) {
        iterations[sg] = hit[sg];
        converged[sg] = hconv[sg];
        for (int w = 0; w < 4; ++w) mains[sg].s[w] = hstates[sg * 4 + w];
    }
}

}  // namespace rinnd
