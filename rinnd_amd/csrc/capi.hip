// capi.hip - host orchestration of the gfx950 NN-Descent build and the C ABI (include/rinnd_hip.h).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "internal.h"


namespace rinnd {

// ---- GF(2) jump tables for FastRng ----------------------------------------------------------------
static void build_jump_tables(std::vector<uint64_t>& tab) {
    tab.assign((size_t)kJumpLevels * 1024, 0);
    // M: column c = transition of basis vector e_c; store as rows.
    std::vector<uint64_t> cols(256 * 4, 0);
    for (int c = 0; c < 256; ++c) {
        Xoshiro x;
        x.s[0] = x.s[1] = x.s[2] = x.s[3] = 0;
        x.s[c >> 6] = 1ull << (c & 63);
        (void)x.next_u64();
        for (int w = 0; w < 4; ++w) cols[c * 4 + w] = x.s[w];
    }
    uint64_t* m0 = tab.data();
    for (int r = 0; r < 256; ++r)
        for (int c = 0; c < 256; ++c)
            if ((cols[c * 4 + (r >> 6)] >> (r & 63)) & 1) m0[r * 4 + (c >> 6)] |= 1ull << (c & 63);
    for (int j = 1; j < kJumpLevels; ++j) {
        const uint64_t* a = tab.data() + (size_t)(j - 1) * 1024;
        uint64_t* out = tab.data() + (size_t)j * 1024;
        for (int r = 0; r < 256; ++r) {
            uint64_t acc[4] = {0, 0, 0, 0};
            for (int t = 0; t < 256; ++t)
                if ((a[r * 4 + (t >> 6)] >> (t & 63)) & 1)
                    for (int w = 0; w < 4; ++w) acc[w] ^= a[t * 4 + w];
            for (int w = 0; w < 4; ++w) out[r * 4 + w] = acc[w];
        }
    }
}

const uint64_t* host_jump_tables() {
    static std::once_flag once;
    static std::vector<uint64_t> tab;
    std::call_once(once, [] { build_jump_tables(tab); });
    return tab.data();
}

// ---- pointers ------------------------------------------------------------------------------------
static bool is_device_ptr(const void* p) {
    if (!p) return false;
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.isManaged;
}

// Device view of a caller array: the pointer itself if it is device memory, else a staged copy.
template <class T>
struct InView {
    const T* d = nullptr;
    DBuf<T> own;
    InView(hipStream_t s, const T* p, size_t count) {
        if (count == 0) return;
        if (is_device_ptr(p)) {
            d = p;
        } else {
            own.alloc(count);
            HIP_TRY(hipMemcpyAsync(own.p, p, count * sizeof(T), hipMemcpyHostToDevice, s));
            d = own.p;
        }
    }
};
template <class T>
static void copy_out(hipStream_t s, T* dst, const T* src, size_t count) {
    if (count == 0 || !dst) return;
    HIP_TRY(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDefault, s));
}

static double elapsed_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        (void)hipGetLastError();
        return 0.0;
    }
    return (double)ms;
}

static double wall_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ---- padded device copy of the data --------------------------------------------------------------
struct DataDev {
    DBuf<float> X;
    int64_t ld = 0;
};
// Device input whose rows already have the padded stride (dim a multiple of 8, 256-byte aligned base)
// is used in place (read-only); anything else is copied into a zero-padded buffer.
static void make_data(hipStream_t s, const float* data, int64_t n, int dim, DataDev& out) {
    out.ld = round_up(std::max(dim, 1), 8);
    if (n > 0 && dim == out.ld && ((uintptr_t)data & 255) == 0 && is_device_ptr(data)) {
        out.X.view(const_cast<float*>(data), (size_t)n * dim);
        return;
    }
    out.X.alloc((size_t)std::max<int64_t>(n, 1) * out.ld);
    InView<float> in(s, data, (size_t)n * dim);
    launch_pad_rows(s, in.d, n, dim, out.X.p, out.ld);
    HIP_TRY(hipStreamSynchronize(s));  // the staged copy dies with `in`
}

// ---- forest ------------------------------------------------------------------------------------------
static void build_forest_dev(rinnd_ctx* ctx, const float* DX, int64_t Dld, int64_t n, int dim, int64_t n_trees,
                             int64_t leaf_size, int angular, int64_t max_depth, int rng_mode,
                             Xoshiro& main, rinnd_forest& F, rinnd_kernel_timing* margins = nullptr) {
    hipStream_t s = ctx->stream;
    F.dim = dim;
    F.n = n;
    F.trees.clear();
    F.trees.resize(n_trees);
    if (n_trees == 0) return;
    RINND_CHECK_ARG(max_depth <= 500, "max_depth above 500 is not supported by the GPU tree builder");
    RINND_CHECK_ARG(dim <= 8192, "dim above 8192 is not supported by the GPU tree builder");
    std::vector<uint64_t> states(n_trees * 4);
    for (int64_t t = 0; t < n_trees; ++t) {  // build_rp_forest seeds (builder.rs:100)
        const uint64_t seed = main.next_u64();
        Xoshiro r = xoshiro_seed(rng_mode == RINND_RNG_COUNTER ? mix64(seed) : seed);
        for (int w = 0; w < 4; ++w) states[t * 4 + w] = r.s[w];
    }
    const int64_t ls = std::max<int64_t>(leaf_size, 1);
    int64_t cap = std::min<int64_t>(2 * n + 1, 8 * ceil_div(n, ls) + 1024);
    for (int attempt = 0; attempt < 2; ++attempt) {
        {  // one pool: [hyper | offsets] of every tree (zeroed together), then children and indices
            auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
            const size_t b_hyper = al((size_t)cap * dim * 4), b_off = al((size_t)cap * 4);
            const size_t b_ch = al((size_t)cap * 8), b_ix = al((size_t)std::max<int64_t>(n, 1) * 4);
            const size_t zeroed = (b_hyper + b_off) * n_trees;
            auto pool = std::make_shared<DBuf<char>>(zeroed + (b_ch + b_ix) * n_trees);
            HIP_TRY(hipMemsetAsync(pool->p, 0, zeroed, s));
            char* z = pool->p;
            char* r = pool->p + zeroed;
            for (auto& T : F.trees) {
                T = TreeDev{};
                T.cap_nodes = cap;
                T.pool = pool;
                T.hyper.view((float*)z, (size_t)cap * dim);
                T.offsets.view((float*)(z + b_hyper), cap);
                T.children.view((int32_t*)r, (size_t)cap * 2);
                T.indices.view((int32_t*)(r + b_ch), std::max<int64_t>(n, 1));
                z += b_hyper + b_off;
                r += b_ch + b_ix;
            }
        }
        if (rng_mode == RINND_RNG_COUNTER) {  // all trees and nodes of a level at once
            std::vector<uint64_t> keys(n_trees);
            for (int64_t t = 0; t < n_trees; ++t) keys[t] = states[t * 4] ^ states[t * 4 + 3];
            if (build_forest_level(s, DX, Dld, dim, n, leaf_size, max_depth, angular, keys, F,
                                   KTimer{ctx->ev[8], ctx->ev[9], margins}, *ctx->ws))
                return;
            cap = 2 * n + 1;
            continue;
        }
        const size_t N1 = (size_t)(n + 1);
        DBuf<uint8_t> side(n_trees * N1);
        DBuf<int32_t> rankR(n_trees * N1), PQ(n_trees * 2 * N1), tmp(n_trees * N1);
        std::vector<float*> hp(n_trees), of(n_trees);
        std::vector<int32_t*> chp(n_trees), ixp(n_trees);
        std::vector<int64_t> caps(n_trees, cap);
        for (int64_t t = 0; t < n_trees; ++t) {
            hp[t] = F.trees[t].hyper.p;
            of[t] = F.trees[t].offsets.p;
            chp[t] = F.trees[t].children.p;
            ixp[t] = F.trees[t].indices.p;
        }
        DBuf<float*> d_hp(n_trees), d_of(n_trees);
        DBuf<int32_t*> d_ch(n_trees), d_ix(n_trees);
        DBuf<int64_t> d_cap(n_trees), d_nn(n_trees);
        DBuf<int32_t> d_status(n_trees);
        DBuf<uint64_t> d_states(n_trees * 4);
        HIP_TRY(hipMemcpyAsync(d_hp.p, hp.data(), n_trees * sizeof(float*), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(d_of.p, of.data(), n_trees * sizeof(float*), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(d_ch.p, chp.data(), n_trees * sizeof(int32_t*), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(d_ix.p, ixp.data(), n_trees * sizeof(int32_t*), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(d_cap.p, caps.data(), n_trees * 8, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(d_states.p, states.data(), n_trees * 32, hipMemcpyHostToDevice, s));
        ForestDFSArgs a{};
        a.X = DX;
        a.ld = Dld;
        a.dim = dim;
        a.n = n;
        a.leaf_size = leaf_size;
        a.max_depth = max_depth;
        a.angular = angular;
        a.n_trees = (int)n_trees;
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
        // Several workgroups per tree when every tree's group fits on the device at once: the walking workgroup
        // hands its large nodes' side tests and partitions to the group.
        const int G = forest_coop_groups(dim, (int)n_trees);
        // a node joins the group from 384 members per workgroup on (below, the group's barriers cost more than the
        // walking workgroup's side tests and partition: measured at 1M x 128, 16 trees)
        const int64_t coop_min = std::max<int64_t>(4096, 384 * (int64_t)G);
        DBuf<CoopCtl> coop_ctl;
        DBuf<float> coop_hp;
        if (G >= 2 && n >= 2 * coop_min) {
            coop_ctl.alloc(n_trees);
            coop_hp.alloc((size_t)n_trees * dim);
            launch_zero(s, coop_ctl.p, n_trees * sizeof(CoopCtl));
            a.coop = coop_ctl.p;
            a.coop_hp = coop_hp.p;
            a.coop_groups = G;
            a.coop_min = coop_min;
        }
        launch_forest_dfs(s, a);
        HIP_TRY(hipGetLastError());
        std::vector<int64_t> nn(n_trees);
        std::vector<int32_t> st(n_trees);
        HIP_TRY(hipMemcpyAsync(nn.data(), d_nn.p, n_trees * 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(st.data(), d_status.p, n_trees * 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (a.coop && std::find(st.begin(), st.end(), 3) != st.end()) {  // a group did not come together: one each
            a.coop = nullptr;
            launch_forest_dfs(s, a);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipMemcpyAsync(nn.data(), d_nn.p, n_trees * 8, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipMemcpyAsync(st.data(), d_status.p, n_trees * 4, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
        }
        bool overflow = false;
        for (int64_t t = 0; t < n_trees; ++t) {
            if (st[t] == 2) throw Error{RINND_ERR_INTERNAL, "RP tree stack overflow"};
            if (st[t] == 1) overflow = true;
            F.trees[t].n_nodes = nn[t];
        }
        if (!overflow) return;
        cap = 2 * n + 1;  // every split leaves two non-empty children: at most 2n-1 nodes
    }
    throw Error{RINND_ERR_INTERNAL, "RP tree node capacity exceeded"};
}

// ---- NN-Descent ---------------------------------------------------------------------------------

// The FP32 join with the context's tile shape for (metric, row stride, C), tuned on the first call.
void join_fp32_cached(rinnd_ctx* ctx, hipStream_t s, int metric, const float* X, int64_t ld, int dim,
                      const int32_t* new_idx, const int32_t* old_idx, int C, int64_t n, const float* heap_dist, int k,
                      const JoinOut& out) {
    const uint64_t key = ((uint64_t)metric << 48) ^ ((uint64_t)ld << 16) ^ (uint64_t)C;
    const auto it = ctx->join_shapes.find(key);
    if (it != ctx->join_shapes.end()) {
        launch_join_fp32(s, metric, X, ld, dim, new_idx, old_idx, C, 0, n, heap_dist, k, out, it->second);
        return;
    }
    bool tuned = false;
    const int shape = launch_join_fp32_tuned(s, metric, X, ld, dim, new_idx, old_idx, C, n, heap_dist, k, out, &tuned);
    if (tuned) ctx->join_shapes[key] = shape;
}

static uint64_t d2h_u64(hipStream_t s, const void* p) {
    uint64_t v = 0;
    HIP_TRY(hipMemcpyAsync(&v, p, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return v;
}

static void nn_descent_dev(rinnd_ctx* ctx, const float* DX, int64_t Dld, int64_t n, int dim, int metric,
                           const rinnd_params& p, Xoshiro& main, int32_t* h_idx, float* h_dist,
                           uint8_t* h_flags, rinnd_forest& F, rinnd_build_stats* stats) {
    hipStream_t s = ctx->stream;
    const int k = (int)p.n_neighbors;
    const int C = (int)std::min<int64_t>(std::min<int64_t>(p.max_candidates, 60), p.n_neighbors);
    const double t_start = wall_ms();
    hipEvent_t* ev = ctx->ev;
    rinnd_build_stats st{};
    LaunchLog tlog;
    tlog.pool = &ctx->ev_pool;

    HIP_TRY(hipEventRecord(ev[0], s));
    build_forest_dev(ctx, DX, Dld, n, dim, p.n_trees, p.leaf_size, p.angular, p.max_depth, p.rng_mode, main, F,
                     &st.timed[RINND_TIMED_MARGINS]);
    HIP_TRY(hipEventRecord(ev[1], s));

    // Counter draws: relabel the points in tree 0's leaf order for the rest of the build, so that
    // points that are close in space are close in memory (leaf rows, candidate rows and heap rows
    // gather from few cache lines).  The results are mapped back to the caller's labels at the end.
    const bool reorder = p.rng_mode == RINND_RNG_COUNTER && !F.trees.empty() && n > 1;
    Workspace& W = *ctx->ws;
    WBuf<int32_t> perm0, inv0, r_idx;
    WBuf<int32_t*> tree_ix;  // device table of the trees' index arrays (relabelling)
    std::vector<int32_t*> tix;  // its host copy: lives until the stream synchronises at the end of the build
    WBuf<float> Xr, r_dist;
    WBuf<uint8_t> r_flags;
    const float* DXw = DX;
    int32_t* w_idx = h_idx;
    float* w_dist = h_dist;
    uint8_t* w_flags = h_flags;
    if (reorder) {
        perm0 = WBuf<int32_t>(W, 0, n);
        inv0 = WBuf<int32_t>(W, 1, n);
        Xr = WBuf<float>(W, 2, (size_t)n * Dld);
        r_idx = WBuf<int32_t>(W, 3, (size_t)n * k);
        r_dist = WBuf<float>(W, 4, (size_t)n * k);
        r_flags = WBuf<uint8_t>(W, 5, (size_t)n * k);
        HIP_TRY(hipMemcpyAsync(perm0.p, F.trees[0].indices.p, n * 4, hipMemcpyDeviceToDevice, s));
        launch_relabel_points(s, DX, Dld, n, perm0.p, inv0.p, Xr.p);
        // a device table of the trees' index arrays: one launch maps them all (here and at the end)
        for (auto& T : F.trees) tix.push_back(T.indices.p);
        tree_ix = WBuf<int32_t*>(W, 42, std::max<size_t>(tix.size(), 1));
        if (!tix.empty())
            HIP_TRY(hipMemcpyAsync(tree_ix.p, tix.data(), tix.size() * sizeof(int32_t*), hipMemcpyHostToDevice, s));
        launch_map_ids_trees(s, tree_ix.p, (int)tix.size(), n, inv0.p);
        DXw = Xr.p;
        w_idx = r_idx.p;
        w_dist = r_dist.p;
        w_flags = r_flags.p;
    }
    launch_fill_heap(s, w_idx, w_dist, w_flags, (int64_t)n * k);
    // BF16 rows (join tiles; leaf tiles too when the data are exact in BF16)
    const bool use_bf16 = p.join_precision == RINND_JOIN_BF16;
    Bf16Rows& B16 = W.b16;
    WBuf<unsigned long long> rescored(W, 6, kKeptSlots);
    if (use_bf16) {
        if (metric != RINND_DIST_SQEUCLIDEAN)
            throw Error{RINND_ERR_UNSUPPORTED, "the BF16 join is implemented for squared Euclidean"};
        prepare_bf16_rows(s, DXw, Dld, n, dim, B16);
    }

    // ---- initialize_from_leaves (mod.rs:203-300), tree by tree ----
    if (!W.pipeline) W.pipeline.reset(new Pipeline());
    Pipeline& PL = *W.pipeline;
    uint64_t u_cap = (uint64_t)2 * n * k;  // candidate records per iteration
    if (leaf_init_supported(k, p.leaf_size)) {  // per-leaf fused pushes, tree by tree
        WBuf<int32_t> mark(W, 7, n + 1), end_at(W, 8, n + 1);
        WBuf<int64_t> rank(W, 9, 2 * (n + 1)), leaves(W, 10, 2 * (n + 1));
        size_t lt_bytes = std::max(scan_u64_tmp_bytes(n + 2), (size_t)16);
        WBuf<char> ltmp(W, 11, lt_bytes);
        WBuf<float> dscr(W, 12, leaf_init_scratch_floats(n, Dld, p.leaf_size));
        WBuf<unsigned long long> lpairs(W, 13, 2);
        HIP_TRY(hipMemsetAsync(lpairs.p, 0, 16, s));
        const bool exact16 = use_bf16 && B16.exact;
        const LeafInitWs lws{mark.p, rank.p, dscr.p, lpairs.p, exact16 ? &B16 : nullptr};
        // deferred timings: no wait per tree
        const KTimer dist_timer{nullptr, nullptr, stats ? &st.timed[RINND_TIMED_LEAF_DIST] : nullptr, &tlog};
        const KTimer push_timer{nullptr, nullptr, stats ? &st.timed[RINND_TIMED_LEAF_PUSH] : nullptr, &tlog};
        const int n_trees = (int)F.trees.size();
        // exact-mode BF16 rows: after the first trees only the pairs below an endpoint's root are written
        // (k_leaf_surv_mfma), in groups of trees sharing one push pass
        const int64_t sdl4 = leaf_surv_dl4(metric, exact16 ? &B16 : nullptr, Dld, p.leaf_size, k);
        const int t_full = sdl4 > 0 ? std::min(n_trees, kLeafFullTrees) : n_trees;
        for (int t = 0; t < t_full; ++t) {
            launch_leaf_list(s, F.trees[t].children.p, F.trees[t].n_nodes, n, mark.p, rank.p, end_at.p,
                             leaves.p, ltmp.p, lt_bytes);
            launch_leaf_init(s, metric, DXw, Dld, dim, F.trees[t].indices.p, leaves.p, rank.p + n, n, p.leaf_size,
                             w_idx, w_dist, w_flags, k, lws, dist_timer, push_timer,
                             t < kLeafFullTrees /* the first trees: most pushes are accepted */);
        }
        if (t_full < n_trees) {
            const double per_tree = (double)n * (sdl4 * 4.0 + 32.0 + 8.0);
            const int G = (int)std::max<double>(1.0, std::min<double>({(double)kSurvGroupMax, (double)(n_trees - t_full),
                                                                       std::floor(surv_group_budget() / per_tree)}));
            WBuf<float> gD(W, 44, (size_t)G * n * sdl4);
            WBuf<uint32_t> gM(W, 39, (size_t)G * n * 8);
            WBuf<int2> gI(W, 45, (size_t)G * n);
            for (int t0 = t_full; t0 < n_trees; t0 += G) {
                SurvGroup grp{};
                grp.D = gD.p;
                grp.Msk = gM.p;
                grp.info = gI.p;
                grp.G = std::min(G, n_trees - t0);
                for (int g = 0; g < grp.G; ++g) {
                    const int t = t0 + g;
                    grp.tidx[g] = F.trees[t].indices.p;
                    launch_leaf_list(s, F.trees[t].children.p, F.trees[t].n_nodes, n, mark.p, rank.p, end_at.p,
                                     leaves.p, ltmp.p, lt_bytes);
                    launch_leaf_surv_tree(s, B16, F.trees[t].indices.p, leaves.p, rank.p + n, n, p.leaf_size, w_dist, k,
                                          grp, g, lpairs.p + 1, dist_timer);
                }
                launch_leaf_push_group(s, grp, DXw, Dld, dim, n, p.leaf_size, w_idx, w_dist, w_flags, k, push_timer);
            }
        }
        if (stats) {  // per-pair bytes: D written (4 B per ordered pair or survivor), D and the partner id read
            unsigned long long lp[2];
            HIP_TRY(hipMemcpyAsync(lp, lpairs.p, 16, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            const double pairs = (double)(lp[0] + lp[1]);
            st.timed[RINND_TIMED_LEAF_DIST].bytes += pairs * 4.0;
            st.timed[RINND_TIMED_LEAF_PUSH].bytes += pairs * 8.0;
        }
        PL.init(n, 1);  // the update buffer is only used by the grouped leaf path below
    } else {
        DBuf<int32_t> mark(n + 1), end_at(n + 1);
        DBuf<int64_t> rank(2 * (n + 1)), leaves(2 * (n + 1));
        DBuf<uint64_t> pc(n + 2), pb(n + 2);
        size_t lt_bytes = std::max(scan_u64_tmp_bytes(n + 2), (size_t)16);
        DBuf<char> ltmp(lt_bytes);
        std::vector<int64_t> n_leaves(F.trees.size()), n_pairs(F.trees.size());
        uint64_t max_pairs = 0;
        for (size_t t = 0; t < F.trees.size(); ++t) {  // sizes first, to size the buffers once
            launch_leaf_list(s, F.trees[t].children.p, F.trees[t].n_nodes, n, mark.p, rank.p, end_at.p,
                             leaves.p, ltmp.p, lt_bytes);
            n_leaves[t] = (int64_t)d2h_u64(s, rank.p + n);
            launch_leaf_pair_base(s, leaves.p, n_leaves[t], pc.p, pb.p, ltmp.p, lt_bytes);
            n_pairs[t] = (int64_t)d2h_u64(s, pb.p + n_leaves[t]);
            max_pairs = std::max<uint64_t>(max_pairs, (uint64_t)n_pairs[t]);
        }
        u_cap = std::max(u_cap, max_pairs);
        PL.init(n, u_cap);
        for (size_t t = 0; t < F.trees.size(); ++t) {
            launch_leaf_list(s, F.trees[t].children.p, F.trees[t].n_nodes, n, mark.p, rank.p, end_at.p,
                             leaves.p, ltmp.p, lt_bytes);
            launch_leaf_pair_base(s, leaves.p, n_leaves[t], pc.p, pb.p, ltmp.p, lt_bytes);
            PL.begin(s);
            launch_leaf_pairs(s, metric, DXw, Dld, dim, F.trees[t].indices.p, leaves.p, n_leaves[t],
                              pb.p, w_dist, k, PL.U.p, PL.U_count.p, u_cap, PL.cnt.p, PL.overflow.p);
            PL.group(s);
            launch_heap_apply(s, PL.off.p, PL.seg.p, PL.seg2.p, n, w_idx, w_dist, w_flags, k,
                              PL.changes.p, PL.big_list.p, PL.big_count.p, PL.bins.p);
        }
        PL.check_overflow(s);
    }
    HIP_TRY(hipEventRecord(ev[2], s));

    // ---- iterations (mod.rs:137-167) ----
    WBuf<int32_t> new_idx(W, 14, (size_t)n * C), old_idx(W, 15, (size_t)n * C);
    WBuf<uint64_t> rvp(W, 16, n + 1);
    WBuf<unsigned long long> rvc(W, 17, n + 1);
    WBuf<float> prio(W, 18, (size_t)2 * n * k + 1);
    WBuf<RngStream> d_streams(W, 19, 64);
    WBuf<unsigned long long> cstat(W, 20, 2);
    WBuf<int64_t> d_bbase(W, 21, 64);
    // candidate build scratch (forward priorities, reverse records before / after the destination sort)
    WBuf<float> c_fpr(W, 22, (size_t)n * k);
    WBuf<uint32_t> c_rkey(W, 23, (size_t)n * k), c_rkey2(W, 24, (size_t)n * k);
    WBuf<uint64_t> c_rval(W, 25, (size_t)n * k), c_rval2(W, 26, (size_t)n * k);
    int c_end_bit = 1;
    while (((int64_t)1 << c_end_bit) <= n) ++c_end_bit;
    WBuf<char> c_sort_tmp(W, 27, std::max<size_t>(cand_sort_tmp_bytes((int64_t)n * k, c_end_bit), 16));
    WBuf<unsigned long long> rpc(W, 28, n + 1);
    WBuf<uint64_t> rowoff(W, 29, n + 1);
    WBuf<unsigned long long> kc(W, 30, n + 1), krec(W, 40, n + 1), kept(W, 41, kKeptSlots);
    HIP_TRY(hipMemsetAsync(krec.p, 0, (n + 1) * 8, s));  // krec[n] stays 0: the scan's total slot
    WBuf<uint64_t> koff(W, 31, n + 1);
    WBuf<uint32_t> jr_key, jr_key2;
    WBuf<uint64_t> jr_val, jr_val2;
    WBuf<char> jr_tmp;
    WBuf<JRec> JU;
    const int T = std::max(1, (int)p.n_threads);
    const int T_eff = (n >= 256 && T > 1) ? T : 1;
    RINND_CHECK_ARG(T_eff <= 64, "n_threads above 64 is not supported");
    const int64_t threshold = (int64_t)(p.delta * (float)p.n_neighbors * (float)n);
    double ms_cand = 0, ms_join = 0, ms_apply = 0;
    int it = 0;
    bool converged = false;
    for (; it < p.n_iters; ++it) {
        HIP_TRY(hipEventRecord(ev[3], s));
        // CandidateSets::build_from_graph
        launch_row_valid_counts(s, w_idx, n, k, (uint64_t*)rvc.p);
        scan_u64_exclusive(s, rvc.p, rvp.p, n + 1, PL.scan_tmp.p, PL.scan_bytes);
        int mode = 2;
        uint64_t counter_key = 0;
        int64_t block_rows = n;
        if (p.rng_mode == RINND_RNG_REFERENCE) {
            std::vector<RngStream> streams;
            if (T_eff == 1) {
                mode = 0;
                const uint64_t V = d2h_u64(s, rvp.p + n);
                RngStream r{};
                for (int w = 0; w < 4; ++w) r.s[w] = main.s[w];
                r.count = (int64_t)(2 * V);
                r.out_offset = 0;
                streams.push_back(r);
                xoshiro_jump(host_jump_tables(), main.s, 2 * V);
            } else {
                mode = 1;
                block_rows = ceil_div(n, T_eff);
                // the valid-edge ranks at the row blocks' boundaries only
                WBuf<uint64_t> d_bnd(W, 43, 65);
                launch_gather_strided(s, rvp.p, block_rows, n, T_eff + 1, d_bnd.p);
                std::vector<uint64_t> hv(T_eff + 1);
                HIP_TRY(hipMemcpyAsync(hv.data(), d_bnd.p, (T_eff + 1) * 8, hipMemcpyDeviceToHost, s));
                HIP_TRY(hipStreamSynchronize(s));
                std::vector<int64_t> bbase(T_eff);
                int64_t acc = 0;
                for (int t = 0; t < T_eff; ++t) {
                    const uint64_t seed = main.next_u64();  // candidates.rs:171
                    RngStream r{};
                    Xoshiro x = xoshiro_seed(seed);
                    for (int w = 0; w < 4; ++w) r.s[w] = x.s[w];
                    r.count = (int64_t)(hv[t + 1] - hv[t]);
                    r.out_offset = acc;
                    bbase[t] = acc;
                    acc += r.count;
                    streams.push_back(r);
                }
                int64_t* pb = (int64_t*)((char*)W.stage.ensure(8192) + 4096);
                std::copy(bbase.begin(), bbase.end(), pb);
                HIP_TRY(hipMemcpyAsync(d_bbase.p, pb, T_eff * 8, hipMemcpyHostToDevice, s));
            }
            if (!ctx->jump_tables.p) {
                ctx->jump_tables.alloc((size_t)kJumpLevels * 1024);
                HIP_TRY(hipMemcpyAsync(ctx->jump_tables.p, host_jump_tables(), ctx->jump_tables.bytes(),
                                       hipMemcpyHostToDevice, s));
                HIP_TRY(hipStreamSynchronize(s));
            }
            int64_t maxc = 0;
            for (auto& r : streams) maxc = std::max(maxc, r.count);
            // staged in pinned memory: no wait here (the next write to it follows this iteration's later waits)
            RngStream* ps = (RngStream*)W.stage.ensure(8192);
            std::copy(streams.begin(), streams.end(), ps);
            HIP_TRY(hipMemcpyAsync(d_streams.p, ps, streams.size() * sizeof(RngStream), hipMemcpyHostToDevice, s));
            launch_rng_streams(s, ctx->jump_tables.p, d_streams.p, (int)streams.size(), maxc, prio.p);
        } else {
            counter_key = main.next_u64();
        }
        launch_fill_i32(s, new_idx.p, -1, (int64_t)n * C);
        launch_fill_i32(s, old_idx.p, -1, (int64_t)n * C);
        HIP_TRY(hipMemsetAsync(rpc.p, 0, (n + 1) * sizeof(unsigned long long), s));  // rows without candidates
        HIP_TRY(hipMemsetAsync(cstat.p, 0, 16, s));
        {
            CandBuildArgs ca{};
            ca.idx = w_idx;
            ca.flags = w_flags;
            ca.n = n;
            ca.k = k;
            ca.C = C;
            ca.mode = mode;
            ca.rvp = rvp.p;
            ca.prio = prio.p;
            ca.block_rows = block_rows;
            ca.block_prio_base = d_bbase.p;
            ca.counter_key = counter_key;
            ca.fpr = c_fpr.p;
            ca.rkey = c_rkey.p;
            ca.rkey2 = c_rkey2.p;
            ca.rval = c_rval.p;
            ca.rval2 = c_rval2.p;
            ca.cnt = PL.cnt.p;
            ca.off = PL.off.p;
            ca.order = PL.big_list.p;
            ca.bins = PL.bins.p;
            ca.sort_tmp = c_sort_tmp.p;
            ca.sort_tmp_bytes = c_sort_tmp.bytes();
            ca.scan_tmp = PL.scan_tmp.p;
            ca.scan_tmp_bytes = PL.scan_bytes;
            ca.new_idx = new_idx.p;
            ca.old_idx = old_idx.p;
            ca.row_pairs = rpc.p;
            ca.counts = cstat.p;
            launch_cand_build(s, ca);
        }
        // update_iteration: generation into per-row regions, grouping by destination, ordered apply.
        // The pair count sizes the join's region buffer.
        scan_u64_exclusive(s, rpc.p, rowoff.p, n + 1, PL.scan_tmp.p, PL.scan_bytes);  // each row's pair region
        uint64_t* hst = (uint64_t*)W.status.ensure(16);
        const hipEvent_t hev = W.status_ev.get();
        HIP_TRY(hipMemcpyAsync(hst, rowoff.p + n, 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipEventRecord(hev, s));
        HIP_TRY(hipMemsetAsync(PL.changes.p, 0, 8, s));
        HIP_TRY(hipMemsetAsync(rescored.p, 0, kKeptSlots * 8, s));
        HIP_TRY(hipMemsetAsync(kept.p, 0, kKeptSlots * 8, s));
        HIP_TRY(hipEventSynchronize(hev));
        JU = WBuf<JRec>(W, 33, std::max<uint64_t>(hst[0], 1));
        const JoinOut jo{rowoff.p, JU.p, kc.p, krec.p, kept.p};
        HIP_TRY(hipEventRecord(ev[4], s));
        if (use_bf16)
            launch_join_bf16(s, DXw, Dld, dim, B16, new_idx.p, old_idx.p, C, 0, n, w_dist, k, jo, rescored.p);
        else
            join_fp32_cached(ctx, s, metric, DXw, Dld, dim, new_idx.p, old_idx.p, C, n, w_dist, k, jo);
        HIP_TRY(hipEventRecord(ev[5], s));
        scan_u64_exclusive(s, krec.p, koff.p, n + 1, PL.scan_tmp.p, PL.scan_bytes);
        // the record count sizes the sort
        HIP_TRY(hipMemcpyAsync(hst + 1, koff.p + n, 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipEventRecord(hev, s));
        HIP_TRY(hipEventSynchronize(hev));
        const uint64_t n_rec = hst[1];
        {
            const uint64_t items = std::max<uint64_t>(n_rec, 1);
            jr_key = WBuf<uint32_t>(W, 34, items);
            jr_key2 = WBuf<uint32_t>(W, 35, items);
            jr_val = WBuf<uint64_t>(W, 36, items);
            jr_val2 = WBuf<uint64_t>(W, 37, items);
            int eb = 1;
            while (((int64_t)1 << eb) < n) ++eb;
            jr_tmp = WBuf<char>(W, 38, std::max<size_t>(join_sort_tmp_bytes((int64_t)items, eb), 16));
            JoinApplyArgs ja{};
            ja.new_idx = new_idx.p;
            ja.old_idx = old_idx.p;
            ja.C = C;
            ja.n = n;
            ja.out = jo;
            ja.koff = koff.p;
            ja.n_records = n_rec;
            ja.rkey = jr_key.p;
            ja.rkey2 = jr_key2.p;
            ja.rval = jr_val.p;
            ja.rval2 = jr_val2.p;
            ja.off = PL.off.p;
            ja.order = PL.big_list.p;
            ja.bins = PL.bins.p;
            ja.sort_tmp = jr_tmp.p;
            ja.sort_tmp_bytes = jr_tmp.bytes();
            ja.scan_tmp = PL.scan_tmp.p;
            ja.scan_tmp_bytes = PL.scan_bytes;
            ja.idx = w_idx;
            ja.dist = w_dist;
            ja.flags = w_flags;
            ja.k = k;
            ja.changes = PL.changes.p;
            launch_join_apply(s, ja);
        }
        HIP_TRY(hipEventRecord(ev[6], s));
        uint64_t ch = 0, cs[2] = {0, 0}, nres = 0, nup = 0, kslots[kKeptSlots], rslots[kKeptSlots];
        HIP_TRY(hipMemcpyAsync(cs, cstat.p, 16, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(rslots, rescored.p, sizeof(rslots), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(kslots, kept.p, sizeof(kslots), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(&ch, PL.changes.p, 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (int j = 0; j < kKeptSlots; ++j) {
            nup += kslots[j];
            nres += rslots[j];
        }
        ms_cand += elapsed_ms(ev[3], ev[4]);  // includes the small counter resets
        const double jm = elapsed_ms(ev[4], ev[5]);
        ms_join += jm;
        ms_apply += elapsed_ms(ev[5], ev[6]);
        if (it < 64) {
            st.changes[it] = (int64_t)ch;
            st.join_updates[it] = (int64_t)nup;
            st.join_pairs[it] = (int64_t)cs[1];
            st.join_rescored[it] = use_bf16 && !B16.exact ? (int64_t)nres : 0;
        }
        // algorithmic bytes of the join launch: gathered candidate rows, both candidate lists, the
        // 2C heap roots read per row, and the 8-byte kept-pair records written
        // (BF16: 2-byte rows + norms, plus the FP32 rows of re-evaluated pairs)
        rinnd_kernel_timing& tj = st.timed[RINND_TIMED_JOIN];
        if (use_bf16)
            tj.bytes += (double)cs[0] * (B16.ldb * 2.0 + 4.0) + (double)n * 2 * C * 4.0 * 2.0 + (double)nup * 8.0 +
                        (B16.exact ? 0.0 : (double)nres * 2.0 * dim * 4.0);
        else
            tj.bytes += (double)cs[0] * dim * 4.0 + (double)n * 2 * C * 4.0 * 2.0 + (double)nup * 8.0;
        tj.ms += jm;
        tj.launches += 1;
        if (p.verbose) std::printf("Iteration %d: %llu updates\n", it + 1, (unsigned long long)ch);
        if ((int64_t)ch <= threshold) {
            converged = true;
            ++it;
            break;
        }
    }
    if (reorder) {
        launch_unpermute_heap(s, w_idx, w_dist, w_flags, perm0.p, n, k, h_idx, h_dist, h_flags);
        launch_map_ids_trees(s, tree_ix.p, (int)F.trees.size(), n, perm0.p);
        HIP_TRY(hipStreamSynchronize(s));  // the relabelled buffers die here
    }
    tlog.flush();
    st.iterations = it;
    st.converged = converged ? 1 : 0;
    st.ms_forest = elapsed_ms(ev[0], ev[1]);
    st.ms_init = elapsed_ms(ev[1], ev[2]);
    st.ms_candidates = ms_cand;
    st.ms_join = ms_join;
    st.ms_apply = ms_apply;
    st.ms_total = wall_ms() - t_start;
    if (p.verbose)
        std::printf("forest %.3f ms, init %.3f ms, candidates %.3f ms, join %.3f ms, apply %.3f ms\n",
                    st.ms_forest, st.ms_init, ms_cand, ms_join, ms_apply);
    if (stats) *stats = st;
}

// index.hip: looks up (once per device, with it current) the device address of the no-allocation query's array
hipError_t noalloc_scratch_init(int device);

void make_padded(hipStream_t s, const float* data, int64_t n, int dim, DBuf<float>& X, int64_t& ld) {
    DataDev D;
    make_data(s, data, n, dim, D);
    X = std::move(D.X);
    ld = D.ld;
}
bool device_pointer(const void* p) { return is_device_ptr(p); }
static void validate_params(const rinnd_params* p, int64_t n, int64_t dim);
void index_nn_descent(rinnd_ctx* ctx, const float* X, int64_t ld, int64_t n, int dim, int metric,
                      const rinnd_params& p, uint64_t state[4], int32_t* idx, float* dist, uint8_t* flags,
                      rinnd_forest& F, rinnd_build_stats* stats) {
    validate_params(&p, n, dim);
    Xoshiro main;
    for (int w = 0; w < 4; ++w) main.s[w] = state[w];
    nn_descent_dev(ctx, X, ld, n, dim, metric, p, main, idx, dist, flags, F, stats);
    for (int w = 0; w < 4; ++w) state[w] = main.s[w];
}

static void validate_params(const rinnd_params* p, int64_t n, int64_t dim) {
    RINND_CHECK_ARG(p, "params is NULL");
    RINND_CHECK_ARG(n >= 0 && dim > 0, "invalid data shape");
    RINND_CHECK_ARG(n < ((int64_t)1 << 31), "more than 2^31-1 points");
    RINND_CHECK_ARG(p->n_neighbors >= 1 && p->n_neighbors <= 256, "n_neighbors must be in [1, 256]");
    RINND_CHECK_ARG(p->n_trees >= 0 && p->n_trees <= 256, "n_trees must be in [0, 256]");
    RINND_CHECK_ARG(p->leaf_size >= 1, "leaf_size must be positive");
    RINND_CHECK_ARG(p->max_candidates >= 1, "max_candidates must be positive");
    RINND_CHECK_ARG(p->n_iters >= 0, "n_iters must be non-negative");
    RINND_CHECK_ARG(p->rng_mode == RINND_RNG_REFERENCE || p->rng_mode == RINND_RNG_COUNTER,
                    "unknown rng_mode");
    RINND_CHECK_ARG(p->join_precision == RINND_JOIN_FP32 || p->join_precision == RINND_JOIN_BF16,
                    "unknown join_precision");
}

}  // namespace rinnd

using namespace rinnd;

// =================================================================================================
// C ABI
// =================================================================================================
template <class F>
static rinnd_status guarded(rinnd_ctx* ctx, F&& f) {
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

extern "C" {

const char* rinnd_version(void) { return "rinnd-mi355x 0.1.0"; }

rinnd_status rinnd_ctx_create(int device, rinnd_ctx** out) {
    if (!out) return RINND_ERR_INVALID_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        (void)hipGetLastError();
        return RINND_ERR_HIP;
    }
    if (device < 0 || device >= count) return RINND_ERR_INVALID_ARG;
    if (hipSetDevice(device) != hipSuccess) return RINND_ERR_HIP;
    if (rinnd::noalloc_scratch_init(device) != hipSuccess) {
        (void)hipGetLastError();
        return RINND_ERR_HIP;
    }
    auto* c = new rinnd_ctx;
    c->ws = new rinnd::Workspace();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return RINND_ERR_HIP;
    }
    for (auto& e : c->ev)  // phase and launch timings only: no system-scope fence at each record
        if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return RINND_ERR_HIP;
    // the FastRng jump tables (built once per process) on the device, so that no build pays for them
    try {
        const uint64_t* jt = rinnd::host_jump_tables();
        c->jump_tables.alloc((size_t)rinnd::kJumpLevels * 1024);
        if (hipMemcpy(c->jump_tables.p, jt, c->jump_tables.bytes(), hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipGetLastError();
            c->jump_tables.release();
        }
    } catch (...) {
        c->jump_tables.release();
    }
    *out = c;
    return RINND_OK;
}

void rinnd_ctx_destroy(rinnd_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : ctx->ev_pool) (void)hipEventDestroy(e);
    ctx->jump_tables.release();
    (void)hipStreamDestroy(ctx->stream);
    delete ctx->ws;  // frees the build scratch with the device current
    ctx->ws = nullptr;
    delete ctx;
}

const char* rinnd_last_error(const rinnd_ctx* ctx) {
    return ctx ? ctx->last_error.c_str() : "null context";
}

void rinnd_rng_seed(uint64_t seed, uint64_t state_out[4]) {
    Xoshiro x = xoshiro_seed(seed);
    for (int w = 0; w < 4; ++w) state_out[w] = x.s[w];
}

void rinnd_rng_jump(uint64_t state[4], uint64_t steps) { xoshiro_jump(host_jump_tables(), state, steps); }

rinnd_status rinnd_pair_distances(rinnd_ctx* ctx, rinnd_distance dist, const float* data, int64_t n,
                                  int64_t dim, const int32_t* a, const int32_t* b, int64_t n_pairs,
                                  float* out) {
    return guarded(ctx, [&] {
        RINND_CHECK_ARG(n > 0 && dim > 0 && n_pairs >= 0, "invalid sizes");
        RINND_CHECK_ARG((int)dist >= 0 && (int)dist <= 6, "unknown distance");
        hipStream_t s = ctx->stream;
        DataDev D;
        make_data(s, data, n, (int)dim, D);
        InView<int32_t> da(s, a, n_pairs), db(s, b, n_pairs);
        DBuf<float> o(std::max<int64_t>(n_pairs, 1));
        launch_pair_distances(s, (int)dist, D.X.p, D.ld, (int)dim, da.d, db.d, n_pairs, o.p);
        HIP_TRY(hipGetLastError());
        copy_out(s, out, o.p, n_pairs);
        HIP_TRY(hipStreamSynchronize(s));
    });
}

rinnd_status rinnd_build_forest(rinnd_ctx* ctx, const float* data, int64_t n, int64_t dim,
                                int64_t n_trees, int64_t leaf_size, int32_t angular,
                                int64_t max_depth, int32_t rng_mode, uint64_t rng_state[4],
                                rinnd_forest** out_forest) {
    return guarded(ctx, [&] {
        RINND_CHECK_ARG(out_forest && rng_state, "null output");
        RINND_CHECK_ARG(n >= 1 && dim >= 1 && leaf_size >= 1, "invalid sizes");
        hipStream_t s = ctx->stream;
        DataDev D;
        make_data(s, data, n, (int)dim, D);
        Xoshiro main;
        for (int w = 0; w < 4; ++w) main.s[w] = rng_state[w];
        auto F = std::make_unique<rinnd_forest>();
        build_forest_dev(ctx, D.X.p, D.ld, n, (int)dim, n_trees, leaf_size, angular, max_depth, rng_mode, main, *F);
        for (int w = 0; w < 4; ++w) rng_state[w] = main.s[w];
        *out_forest = F.release();
    });
}

int64_t rinnd_forest_n_trees(const rinnd_forest* f) { return f ? (int64_t)f->trees.size() : 0; }

rinnd_status rinnd_forest_tree_info(const rinnd_forest* f, int64_t tree, int64_t* n_nodes,
                                    int64_t* dim) {
    if (!f || tree < 0 || tree >= (int64_t)f->trees.size()) return RINND_ERR_INVALID_ARG;
    if (n_nodes) *n_nodes = f->trees[tree].n_nodes;
    if (dim) *dim = f->dim;
    return RINND_OK;
}

rinnd_status rinnd_forest_tree_copy(const rinnd_forest* f, int64_t tree, float* hyperplanes,
                                    float* offsets, int32_t* children, int32_t* indices) {
    if (!f || tree < 0 || tree >= (int64_t)f->trees.size()) return RINND_ERR_INVALID_ARG;
    const TreeDev& T = f->trees[tree];
    const size_t nn = (size_t)T.n_nodes;
    if (hyperplanes && nn) (void)hipMemcpy(hyperplanes, T.hyper.p, nn * f->dim * 4, hipMemcpyDefault);
    if (offsets && nn) (void)hipMemcpy(offsets, T.offsets.p, nn * 4, hipMemcpyDefault);
    if (children && nn) (void)hipMemcpy(children, T.children.p, nn * 8, hipMemcpyDefault);
    if (indices && f->n) (void)hipMemcpy(indices, T.indices.p, f->n * 4, hipMemcpyDefault);
    return hipGetLastError() == hipSuccess ? RINND_OK : RINND_ERR_HIP;
}

void rinnd_forest_free(rinnd_forest* f) { delete f; }

rinnd_status rinnd_nn_descent(rinnd_ctx* ctx, const float* data, int64_t n, int64_t dim,
                              rinnd_distance dist, const rinnd_params* params, uint64_t rng_state[4],
                              int32_t* out_idx, float* out_dist, uint8_t* out_flags,
                              rinnd_forest** out_forest, rinnd_build_stats* stats) {
    return guarded(ctx, [&] {
        validate_params(params, n, dim);
        RINND_CHECK_ARG(rng_state, "rng_state is NULL");
        RINND_CHECK_ARG((int)dist >= 0 && (int)dist <= 6, "unknown distance");
        hipStream_t s = ctx->stream;
        auto F = std::make_unique<rinnd_forest>();
        Xoshiro main;
        for (int w = 0; w < 4; ++w) main.s[w] = rng_state[w];
        const int64_t k = params->n_neighbors;
        if (n == 0) {
            F->dim = dim;
            if (out_forest) *out_forest = F.release();
            if (stats) *stats = rinnd_build_stats{};
            return;
        }
        DataDev D;
        make_data(s, data, n, (int)dim, D);
        // device outputs are written in place; otherwise the heap lives in the context's workspace
        const size_t cnt = (size_t)n * k;
        const bool direct = is_device_ptr(out_idx) && is_device_ptr(out_dist) && is_device_ptr(out_flags);
        Workspace& W = *ctx->ws;
        WBuf<int32_t> hi;
        WBuf<float> hd;
        WBuf<uint8_t> hf;
        if (!direct) {
            hi = WBuf<int32_t>(W, kWsHeap + 0, cnt);
            hd = WBuf<float>(W, kWsHeap + 1, cnt);
            hf = WBuf<uint8_t>(W, kWsHeap + 2, cnt);
        }
        int32_t* h_idx = direct ? out_idx : hi.p;
        float* h_dist = direct ? out_dist : hd.p;
        uint8_t* h_flags = direct ? out_flags : hf.p;
        nn_descent_dev(ctx, D.X.p, D.ld, n, (int)dim, (int)dist, *params, main, h_idx, h_dist, h_flags, *F, stats);
        if (!direct) {
            copy_out(s, out_idx, h_idx, cnt);
            copy_out(s, out_dist, h_dist, cnt);
            copy_out(s, out_flags, h_flags, cnt);
        }
        HIP_TRY(hipStreamSynchronize(s));
        for (int w = 0; w < 4; ++w) rng_state[w] = main.s[w];
        if (out_forest) *out_forest = F.release();
    });
}

rinnd_status rinnd_nn_descent_batch(rinnd_ctx* ctx, const float* data, const int64_t* offsets, int64_t n_indexes,
                                    int64_t dim, rinnd_distance dist, const rinnd_params* params,
                                    const int64_t* n_trees, const int64_t* n_iters, uint64_t* rng_states,
                                    int32_t* out_idx, float* out_dist, uint8_t* out_flags, int32_t* out_iterations,
                                    int32_t* out_converged, int32_t flags, rinnd_batch_stats* stats) {
    return guarded(ctx, [&] {
        const double t0 = wall_ms();
        RINND_CHECK_ARG(n_indexes >= 0 && n_indexes <= 65535, "n_indexes must be in [0, 65535]");
        RINND_CHECK_ARG(n_indexes == 0 || (offsets && rng_states), "offsets or rng_states is NULL");
        RINND_CHECK_ARG((int)dist >= 0 && (int)dist <= 6, "unknown distance");
        std::vector<int64_t> off(n_indexes + 1, 0), nt(n_indexes), ni(n_indexes);
        for (int64_t i = 0; i <= n_indexes; ++i) off[i] = n_indexes ? offsets[i] : 0;
        RINND_CHECK_ARG(off[0] == 0, "offsets[0] must be 0");
        for (int64_t i = 0; i < n_indexes; ++i) {
            RINND_CHECK_ARG(off[i + 1] >= off[i], "offsets must be non-decreasing");
            nt[i] = n_trees ? n_trees[i] : params ? params->n_trees : 0;
            ni[i] = n_iters ? n_iters[i] : params ? params->n_iters : 0;
            RINND_CHECK_ARG(nt[i] >= 0 && nt[i] <= 256, "n_trees must be in [0, 256]");
            RINND_CHECK_ARG(ni[i] >= 0, "n_iters must be non-negative");
        }
        const int64_t N = off[n_indexes];
        validate_params(params, N, dim);
        if (params->rng_mode != RINND_RNG_REFERENCE || params->n_threads > 1)
            throw Error{RINND_ERR_UNSUPPORTED, "the batched build reproduces reference draws with one candidate thread"};
        if (params->join_precision != RINND_JOIN_FP32)
            throw Error{RINND_ERR_UNSUPPORTED, "the batched build runs the FP32 join"};
        if (!leaf_init_supported((int)params->n_neighbors, params->leaf_size))
            throw Error{RINND_ERR_UNSUPPORTED, "the batched build needs the fused leaf initialisation (k, leaf_size)"};
        if (stats) *stats = rinnd_batch_stats{};
        if (N == 0) {
            for (int64_t i = 0; i < n_indexes; ++i) {
                if (out_iterations) out_iterations[i] = 0;
                if (out_converged) out_converged[i] = 0;
            }
            return;
        }
        hipStream_t s = ctx->stream;
        std::vector<Xoshiro> mains(n_indexes);
        for (int64_t i = 0; i < n_indexes; ++i)
            for (int w = 0; w < 4; ++w) mains[i].s[w] = rng_states[4 * i + w];
        DataDev D;
        make_data(s, data, N, (int)dim, D);
        const int64_t k = params->n_neighbors;
        const size_t cnt = (size_t)N * k;
        const bool direct = is_device_ptr(out_idx) && is_device_ptr(out_dist) && is_device_ptr(out_flags);
        Workspace& W = *ctx->ws;
        WBuf<int32_t> hi;
        WBuf<float> hd;
        WBuf<uint8_t> hf;
        if (!direct) {
            hi = WBuf<int32_t>(W, kWsHeap + 0, cnt);
            hd = WBuf<float>(W, kWsHeap + 1, cnt);
            hf = WBuf<uint8_t>(W, kWsHeap + 2, cnt);
        }
        int32_t* h_idx = direct ? out_idx : hi.p;
        float* h_dist = direct ? out_dist : hd.p;
        uint8_t* h_flags = direct ? out_flags : hf.p;
        std::vector<int32_t> its, conv;
        rinnd_batch_stats st{};
        nn_descent_batch_dev(ctx, D.X.p, D.ld, off, (int)dim, (int)dist, *params, nt, ni, mains, h_idx, h_dist,
                             h_flags, (flags & RINND_BATCH_NO_GRAPH) == 0, its, conv, st);
        if (!direct) {
            copy_out(s, out_idx, h_idx, cnt);
            copy_out(s, out_dist, h_dist, cnt);
            copy_out(s, out_flags, h_flags, cnt);
        }
        HIP_TRY(hipStreamSynchronize(s));
        for (int64_t i = 0; i < n_indexes; ++i) {
            for (int w = 0; w < 4; ++w) rng_states[4 * i + w] = mains[i].s[w];
            if (out_iterations) out_iterations[i] = its[i];
            if (out_converged) out_converged[i] = conv[i];
        }
        st.ms_total = wall_ms() - t0;
        if (stats) *stats = st;
    });
}

rinnd_status rinnd_sort_rows(rinnd_ctx* ctx, int32_t* idx, float* dist, uint8_t* flags, int64_t n,
                             int64_t k) {
    return guarded(ctx, [&] {
        RINND_CHECK_ARG(n >= 0 && k >= 1 && k <= 256, "invalid sizes");
        hipStream_t s = ctx->stream;
        const size_t cnt = (size_t)n * k;
        if (cnt == 0) return;
        // device rows are sorted in place; host rows are staged through the context's workspace
        Workspace& W = *ctx->ws;
        const bool dev_i = is_device_ptr(idx), dev_d = is_device_ptr(dist), dev_f = flags && is_device_ptr(flags);
        WBuf<int32_t> di;
        WBuf<float> dd;
        WBuf<uint8_t> df;
        int32_t* pi = idx;
        float* pd = dist;
        uint8_t* pf = flags;
        if (!dev_i) {
            di = WBuf<int32_t>(W, kWsHeap + 0, cnt);
            HIP_TRY(hipMemcpyAsync(di.p, idx, cnt * 4, hipMemcpyDefault, s));
            pi = di.p;
        }
        if (!dev_d) {
            dd = WBuf<float>(W, kWsHeap + 1, cnt);
            HIP_TRY(hipMemcpyAsync(dd.p, dist, cnt * 4, hipMemcpyDefault, s));
            pd = dd.p;
        }
        if (!dev_f) {
            df = WBuf<uint8_t>(W, kWsHeap + 2, cnt);
            if (flags) HIP_TRY(hipMemcpyAsync(df.p, flags, cnt, hipMemcpyDefault, s));
            else HIP_TRY(hipMemsetAsync(df.p, 0, cnt, s));
            pf = df.p;
        }
        launch_sort_rows(s, pi, pd, pf, n, (int)k);
        HIP_TRY(hipGetLastError());
        if (!dev_i) copy_out(s, idx, pi, cnt);
        if (!dev_d) copy_out(s, dist, pd, cnt);
        if (flags && !dev_f) copy_out(s, flags, pf, cnt);
        HIP_TRY(hipStreamSynchronize(s));
    });
}

}  // extern "C"
