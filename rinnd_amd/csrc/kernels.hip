// kernels.hip - gfx950 kernels of the NN-Descent build path.
//
// Phases and the reference code they reproduce:
//   k_forest_dfs          RP trees with the reference's FastRng streams (core/tree/builder.rs:89-348)
//   k_leaf_pairs          leaf pairs of initialize_from_leaves when the heaps do not fit in LDS
//                         (core/nndescent/mod.rs:203-300; the LDS form is leaf_init.hip)
//   k_scatter/k_seg_*     grouping + ordered apply of those leaf pairs (mod.rs:450-487)
//   k_cand_gen_rev/_push_rev  CandidateSets::build_from_graph (core/nndescent/candidates.rs:51-484)
//   k_join_fp32           update_iteration's candidate-pair distances + filter (mod.rs:321-448)
//   k_join_records/k_heap_push_rec  per-vertex heap pushes in the reference's generation order
//                         (mod.rs:450-487)
//   k_sort_rows           NeighborHeap::sort_all (core/heap/neighbor_heap.rs:293-320)
//
// Ordering argument used by the apply pipeline: the reference applies, for every vertex, its pushes
// in global generation order (Rayon buckets are concatenated in source-thread order), and its
// threshold filter only drops pushes the heap would reject anyway (roots only decrease).  Each
// update therefore carries its generation key; a destination sorts its updates by key and replays
// checked_flagged_push sequentially, which gives the reference's heaps and change counts exactly.
#include <hipcub/hipcub.hpp>
#include <rocprim/device/device_radix_sort.hpp>

#include "internal.h"

namespace rinnd {



// ---- fills / copies ------------------------------------------------------------------------------
__global__ void k_pad_rows(const float* __restrict__ src, int64_t n, int dim,
                           float* __restrict__ dst, int64_t ld) {
    const int64_t total = n * ld;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = t / ld, c = t - r * ld;
        dst[t] = c < dim ? src[r * dim + c] : 0.f;
    }
}
void launch_pad_rows(hipStream_t s, const float* src, int64_t n, int dim, float* dst, int64_t ld) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_pad_rows, dim3(grid_for(n * ld, 256)), dim3(256), 0, s, src, n, dim, dst, ld);
}

// NeighborHeap::new (neighbor_heap.rs:35-44): (-1, +inf, 0)
__global__ void k_fill_heap(int32_t* idx, float* dist, uint8_t* flags, int64_t count) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < count;
         t += (int64_t)gridDim.x * blockDim.x) {
        idx[t] = -1;
        dist[t] = __builtin_inff();
        flags[t] = 0;
    }
}
void launch_fill_heap(hipStream_t s, int32_t* idx, float* dist, uint8_t* flags, int64_t count) {
    if (count == 0) return;
    hipLaunchKernelGGL(k_fill_heap, dim3(grid_for(count, 256)), dim3(256), 0, s, idx, dist, flags, count);
}

__global__ void k_fill_i32(int32_t* p, int32_t v, int64_t count) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < count;
         t += (int64_t)gridDim.x * blockDim.x)
        p[t] = v;
}
void launch_fill_i32(hipStream_t s, int32_t* p, int32_t v, int64_t count) {
    if (count == 0) return;
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(count, 256)), dim3(256), 0, s, p, v, count);
}
__global__ void k_zero_u8(uint8_t* p, int64_t count) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < count; t += (int64_t)gridDim.x * blockDim.x)
        p[t] = 0;
}
void launch_zero(hipStream_t s, void* p, size_t bytes) {
    if (bytes == 0) return;
    if (((uintptr_t)p & 3) == 0 && (bytes & 3) == 0)
        launch_fill_i32(s, (int32_t*)p, 0, (int64_t)(bytes >> 2));
    else
        hipLaunchKernelGGL(k_zero_u8, dim3(grid_for((int64_t)bytes, 256)), dim3(256), 0, s, (uint8_t*)p,
                           (int64_t)bytes);
}

__global__ void k_pair_distances(int metric, const float* __restrict__ X, int64_t ld, int dim,
                                 const int32_t* __restrict__ a, const int32_t* __restrict__ b,
                                 int64_t n_pairs, float* __restrict__ out) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n_pairs;
         t += (int64_t)gridDim.x * blockDim.x)
        out[t] = distance_exact(metric, X + (int64_t)a[t] * ld, X + (int64_t)b[t] * ld, dim);
}
void launch_pair_distances(hipStream_t s, int metric, const float* X, int64_t ld, int dim,
                           const int32_t* a, const int32_t* b, int64_t n_pairs, float* out) {
    if (n_pairs == 0) return;
    hipLaunchKernelGGL(k_pair_distances, dim3(grid_for(n_pairs, 256)), dim3(256), 0, s, metric, X, ld,
                       dim, a, b, n_pairs, out);
}

// ---- RP forest with the reference streams ---------------------------------------------------------
// One workgroup per tree walks the tree depth-first exactly like TreeBuilder::build_recursive
// (builder.rs:172-226): node ids in pre-order, split points from the tree's FastRng, hyperplane
// sums in the reference's sequential order, margins in the AVX2 dot order (8 lanes per point).
// The two-pointer partition (builder.rs:312-348) is computed in parallel from its closed form:
// with s = #L, a left element at p < s stays; the b-th left element counted from the back fills
// the b-th right element counted from the front; the r-th right element from the front (p <= s)
// moves to (r == 0 ? m : Q[r-1]) - 1 where Q[b] is the position of the b-th left element from the
// back; a right element at p > s moves to p - 1.  A node containing a margin tie (|m| < 1e-8)
// consumes next_bool() in examination order, so it is partitioned sequentially instead.
constexpr int kForestThreads = 1024;
constexpr int kForestStack = 512;
struct StackEnt {
    int32_t start, end, depth, pslot;
};

// In the work-queue form (many trees in flight), subtrees of at most kWaveSub points are walked by one wave
// with their positions in LDS (no workgroup barriers): the nodes, their order and the draws are those of the
// workgroup walk.
constexpr int kWaveSub = 1024;

// ---- partition_inplace with margin ties (builder.rs:319-347) on one wave ------------------------------------
// The reference's two pointers examine the node's elements in an order fixed by the decisions: the first
// element from the front; after an element goes left the next front element, after one goes right the next
// element from the back.  A left-goer lands at (lefts so far), a right-goer at m - 1 - (rights so far), and a
// tie takes the next draw (go left iff !next_bool()) in that order.  Wave 0 replays the order with 64-element
// windows of both ends in registers, writes every element's final position to dst[] and advances rs; returns
// the number of left-goers.  Every lane runs the same (uniform) steps; lane 0 writes.
__device__ int tie_partition_wave(const uint8_t* __restrict__ side, int m, uint64_t* rs, int32_t* dst) {
    const int lane = threadIdx.x & 63;
    Xoshiro r;
    for (int w = 0; w < 4; ++w) r.s[w] = rs[w];
    int L = 0, R = 0, fi = 0, bj = 0;
    bool fromF = true;
    int fv = lane < m ? side[lane] : 0;            // front window: side[fi + lane]
    int bv = m - 1 - lane >= 0 ? side[m - 1 - lane] : 0;  // back window: side[m - 1 - (bj + lane)]
    for (int k = 0; k < m; ++k) {
        int sd, p;
        if (fromF) {
            if (L - fi >= 64) {
                fi += 64;
                fv = fi + lane < m ? side[fi + lane] : 0;
            }
            sd = __builtin_amdgcn_readlane(fv, __builtin_amdgcn_readfirstlane(L - fi));
            p = L;
        } else {
            const int j = R - 1;
            if (j - bj >= 64) {
                bj += 64;
                bv = m - 1 - (bj + lane) >= 0 ? side[m - 1 - (bj + lane)] : 0;
            }
            sd = __builtin_amdgcn_readlane(bv, __builtin_amdgcn_readfirstlane(j - bj));
            p = m - 1 - j;
        }
        const bool left = sd == 2 ? !r.next_bool() : sd == 0;
        if (lane == 0) dst[p] = left ? L : m - 1 - R;
        if (left) {
            ++L;
            fromF = true;
        } else {
            ++R;
            fromF = false;
        }
    }
    if (lane == 0)
        for (int w = 0; w < 4; ++w) rs[w] = r.s[w];
    return L;
}

// The same with few ties (thread 0): only the ties need the examination order.  With every decision before a
// tie known, the step at which the two pointers reach it has a closed form: the front element at position x
// (r right-goers before it in front order) is examined after the first r back runs, i.e. after the back
// elements up to the r-th left-goer from the end; the back element at position t (l left-goers behind it) after
// the front elements up to the (l+1)-th right-goer.  The undecided ties always form a contiguous range of the
// sorted tie positions (the next one examined is the first or the last of them): both candidates' steps are
// computed (a candidate is blocked when an undecided tie would be examined on the way), the earlier takes the
// next draw.  cntR[p] = right-goers (side 1) before p, p in [0, m]; ties[] ascending; dec[i] = 0 left / 1 right.
constexpr int kMaxTies = 64;
__device__ void tie_resolve(const int32_t* __restrict__ cntR, int m, const int* ties, int T, int8_t* dec,
                            uint64_t* rs) {
    for (int i = 0; i < T; ++i) dec[i] = -1;
    auto cntT = [&](int p) {  // ties before p
        int c = 0;
        for (int i = 0; i < T; ++i) c += ties[i] < p;
        return c;
    };
    auto cntL = [&](int p) { return p - cntR[p] - cntT(p); };  // side-0 elements before p
    const int totL = cntL(m);
    auto g = [&](int q) {  // known left-goers in [q, m)
        int c = totL - cntL(q);
        for (int i = 0; i < T; ++i) c += dec[i] == 0 && ties[i] >= q;
        return c;
    };
    auto h = [&](int p) {  // known right-goers in [0, p)
        int c = cntR[p];
        for (int i = 0; i < T; ++i) c += dec[i] == 1 && ties[i] < p;
        return c;
    };
    Xoshiro r;
    for (int w = 0; w < 4; ++w) r.s[w] = rs[w];
    int lo_i = 0, hi_i = T - 1;
    while (lo_i <= hi_i) {
        const int tlo = ties[lo_i], thi = ties[hi_i];
        int fpos = -1, bpos = -1;
        {  // tlo from the front
            const int rr = h(tlo);
            if (rr == 0) {
                fpos = tlo;
            } else if (g(0) >= rr) {
                int lo = 0, hi = m;  // the rr-th left-goer from the end: max { q : g(q) >= rr }
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (g(mid) >= rr) lo = mid;
                    else hi = mid - 1;
                }
                if (thi <= lo) fpos = tlo + (m - lo);
            }
        }
        {  // thi from the back
            const int need = g(thi + 1) + 1;
            if (h(m) >= need) {
                int lo = 1, hi = m;  // the need-th right-goer from the front sits at min { p : h(p) >= need } - 1
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (h(mid) >= need) hi = mid;
                    else lo = mid + 1;
                }
                const int ps = lo - 1;
                if (tlo >= ps) bpos = (m - 1 - thi) + ps + 1;
            }
        }
        const int d = r.next_bool() ? 1 : 0;  // goes left iff !next_bool()
        if (fpos >= 0 && (bpos < 0 || fpos < bpos)) dec[lo_i++] = (int8_t)d;
        else dec[hi_i--] = (int8_t)d;
    }
    for (int w = 0; w < 4; ++w) rs[w] = r.s[w];
}

// ---- cooperating workgroups of one tree ----------------------------------------------------------------
// Waits are bounded by the constant-rate wall clock (wall_clock64, 100 MHz on this hardware): about four
// seconds without progress sets the tree's abort flag, which every waiting workgroup sees and leaves on.
constexpr uint64_t kCoopTimeoutTicks = 400000000ull;

// thread 0 only: wait until *p >= target; false on abort or timeout.  The polls are relaxed (an acquire per
// poll would invalidate the caches the other workgroups are using); one acquire fence follows success.
__device__ bool coop_wait(int* p, int target, int* abortp) {
    const uint64_t t0 = wall_clock64();
    for (;;) {
        if (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            return true;
        }
        if (__hip_atomic_load(abortp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
        if (wall_clock64() - t0 > kCoopTimeoutTicks) {
            __hip_atomic_store(abortp, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(4);
    }
}

// All G workgroups of the tree: every wave waits for its own stores, then thread 0 releases the workgroup's
// writes at device scope (one L2 write-back per workgroup) and counts in; false (in every thread) on abort.
__device__ bool coop_barrier(CoopCtl* c, int G, int* sh_nb, int* sh_ok) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        atomicAdd(&c->bar, 1);
        const int nb = *sh_nb + 1;
        *sh_nb = nb;
        *sh_ok = coop_wait(&c->bar, G * nb, &c->abort) ? 1 : 0;
    }
    __syncthreads();
    return *sh_ok != 0;
}

// One node's side tests and partition by workgroup g of G (contiguous slices of the node's m positions), the
// same closed form as the single-workgroup walk: returns the split, -1 if some margin is a tie (side[] then
// holds every position's side and the walking workgroup partitions alone), -2 on abort.
template <int THREADS>
__device__ int coop_node(const float* X, int64_t ld, int dim, const float* hp, float off, int start, int m, int g,
                         int G, CoopCtl* c, int32_t* idx, uint8_t* side, int32_t* rankR, int32_t* P, int32_t* Q,
                         int32_t* tmp, int* sh_nb) {
    __shared__ int cw_tot[THREADS / RINND_WAVE];
    __shared__ int cw_R, cw_tie, cw_ok, cw_base, cw_totR;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const unsigned long long lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    const int slice = (int)round_up(ceil_div(m, G), 64);
    const int s0 = min(m, g * slice), s1 = min(m, s0 + slice);
    if (tid == 0) cw_R = cw_tie = 0;
    __syncthreads();
    {  // margins: two lanes per point (as the single-workgroup walk)
        const int pg = tid >> 1, h = tid & 1;
        const int chunks = dim >> 3;
        int myR = 0, mytie = 0;
        for (int base = s0; base < s1; base += THREADS / 2) {
            const int pos = base + pg;
            const bool act = pos < s1;
            const float* x = X + (int64_t)(act ? idx[start + pos] : 0) * ld;
            float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
            if (act)
                for (int cc = 0; cc < chunks; ++cc) {
                    const float4 u = *reinterpret_cast<const float4*>(x + cc * 8 + 4 * h);
                    const float4 w = *reinterpret_cast<const float4*>(hp + cc * 8 + 4 * h);
                    a0 = __builtin_fmaf(u.x, w.x, a0);
                    a1 = __builtin_fmaf(u.y, w.y, a1);
                    a2 = __builtin_fmaf(u.z, w.z, a2);
                    a3 = __builtin_fmaf(u.w, w.w, a3);
                }
            const float b0 = __shfl_xor(a0, 1), b1 = __shfl_xor(a1, 1);
            const float b2 = __shfl_xor(a2, 1), b3 = __shfl_xor(a3, 1);
            if (act && h == 0) {
                float r = ((a0 + b0) + (a1 + b1)) + ((a2 + b2) + (a3 + b3));
                for (int i = chunks * 8; i < dim; ++i) r += x[i] * hp[i];
                const float margin = r + off;
                uint8_t sd;
                if (fabsf(margin) < 1e-8f) {
                    sd = 2;
                    mytie = 1;
                } else {
                    sd = margin < 0.0f ? 0 : 1;
                }
                side[pos] = sd;
                myR += sd == 1;
            }
        }
        for (int o = 32; o > 0; o >>= 1) myR += __shfl_xor(myR, o);
        const unsigned long long anyt = __ballot(mytie != 0);
        if (lane == 0) {
            atomicAdd(&cw_R, myR);
            if (anyt) atomicOr(&cw_tie, 1);
        }
        __syncthreads();
        if (tid == 0) {
            c->cnt[g] = cw_R;
            c->tie[g] = cw_tie;
        }
    }
    if (!coop_barrier(c, G, sh_nb, &cw_ok)) return -2;
    if (tid == 0) {
        int tie = 0, base = 0, tot = 0;
        for (int q = 0; q < G; ++q) {
            const int v = __hip_atomic_load(&c->cnt[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            tie |= __hip_atomic_load(&c->tie[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (q < g) base += v;
            tot += v;
        }
        cw_tie = tie;
        cw_base = base;
        cw_totR = tot;
    }
    __syncthreads();
    if (cw_tie) {  // every workgroup has read the flags before the walking one moves on and rewrites them
        if (!coop_barrier(c, G, sh_nb, &cw_ok)) return -2;
        return -1;
    }
    const int totR = cw_totR, split = m - totR;
    {  // rankR (rights before p) and the wrong-side lists P / Q of the closed form
        int running = cw_base;
        for (int cb = s0; cb < s1; cb += THREADS) {  // uniform
            const int p = cb + tid;
            const uint8_t sd = p < s1 ? side[p] : 0;
            const bool isR = sd == 1;
            const unsigned long long bal = __ballot(isR);
            if (lane == 0) cw_tot[wid] = __popcll(bal);
            __syncthreads();
            int wpre = 0, tot = 0;
            for (int w = 0; w < THREADS / RINND_WAVE; ++w) {
                wpre += w < wid ? cw_tot[w] : 0;
                tot += cw_tot[w];
            }
            if (p < s1) {
                const int r = running + wpre + __popcll(bal & lt);
                rankR[p] = r;
                const int rn = r + (isR ? 1 : 0);
                if (sd == 1 && p < split) P[r] = p;
                if (sd == 0 && p >= split) Q[(m - p - 1) - (totR - rn)] = p;
            }
            __syncthreads();
            running += tot;
        }
    }
    if (!coop_barrier(c, G, sh_nb, &cw_ok)) return -2;
    for (int p = s0 + tid; p < s1; p += THREADS) {  // destinations
        const uint8_t sd = side[p];
        const int r = rankR[p];
        int dest;
        if (sd == 0) {
            dest = p < split ? p : P[(m - p - 1) - (totR - r)];
        } else {
            dest = p <= split ? ((r == 0 ? m : Q[r - 1]) - 1) : p - 1;
        }
        tmp[dest] = idx[start + p];
    }
    if (!coop_barrier(c, G, sh_nb, &cw_ok)) return -2;
    for (int p = s0 + tid; p < s1; p += THREADS) idx[start + p] = tmp[p];
    if (!coop_barrier(c, G, sh_nb, &cw_ok)) return -2;
    return split;
}

// Sequential sums 0 + a[0] + a[1] + ... (and the same of b) in that order, as the reference's scalar loops
// add them, by one thread; the loads of eight terms go out before their additions.
__device__ __forceinline__ void seq_sum2(const float* a, const float* b, int n, float& sa, float& sb) {
    float x = 0.f, y = 0.f;
    int i = 0;
    for (; i + 8 <= n; i += 8) {
        float va[8], vb[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            va[u] = a[i + u];
            vb[u] = b[i + u];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            x += va[u];
            y += vb[u];
        }
    }
    for (; i < n; ++i) {
        x += a[i];
        y += b[i];
    }
    sa = x;
    sb = y;
}

template <int THREADS, bool WAVE_SUB, bool COOP = false>
__device__ void forest_tree_dfs(const ForestDFSArgs& a, const int t, const int64_t so) {
    extern __shared__ __align__(16) float fsm[];  // hp[dim], p1[dim], p2[dim], two product arrays q1[dim], q2[dim]
    __shared__ StackEnt stack[kForestStack];
    __shared__ int32_t wv_idx[kWaveSub], wv_tmp[kWaveSub], wv_rank[kWaveSub + 1], wv_P[kWaveSub], wv_Q[kWaveSub];
    __shared__ uint8_t wv_side[kWaveSub];
    __shared__ uint64_t rs[4];
    __shared__ int sh_sp, sh_ctr, sh_abort, sh_tie, sh_node, sh_split, sh_leaf, sh_scale;
    __shared__ int sh_i1, sh_i2;
    __shared__ float sh_off, sh_inv, sh_n1, sh_n2;
    __shared__ int wave_tot[THREADS / RINND_WAVE];
    __shared__ int sh_running;
    __shared__ int sh_seq, sh_nb;  // cooperative form: nodes published, barriers passed
    __shared__ int sh_ntie, sh_ties[kMaxTies];
    __shared__ int8_t sh_dec[kMaxTies];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wid = tid >> 6;
    const int64_t n = a.tree_n ? a.tree_n[t] : a.n;
    const float* const X = a.X + (a.tree_base ? a.tree_base[t] : 0) * a.ld;  // this tree's point set
    const int dim = a.dim;
    float* hp = fsm;
    float* p1s = fsm + dim;
    float* p2s = fsm + 2 * dim;
    float* q1s = fsm + 3 * dim;  // per-dimension products, summed in order by one thread
    float* q2s = fsm + 4 * dim;
    float* hyper = a.hyper[t];
    float* offs = a.offsets[t];
    int32_t* ch = a.children[t];
    int32_t* idx = a.indices[t];
    uint8_t* side = a.side + so;
    int32_t* rankR = a.rankR + so;
    int32_t* P = a.PQ + 2 * so;
    int32_t* Q = P + (n + 1);
    int32_t* tmp = a.tmp + so;
    const int64_t cap = a.cap_nodes[t];

    for (int64_t i = tid; i < n; i += blockDim.x) idx[i] = (int32_t)i;
    if (tid == 0) {
        for (int w = 0; w < 4; ++w) rs[w] = a.tree_states[t * 4 + w];
        stack[0] = StackEnt{0, (int32_t)n, 0, -1};
        sh_sp = 1;
        sh_ctr = 0;
        sh_abort = 0;
        sh_seq = 0;
        sh_nb = 0;
    }
    __syncthreads();

    // The subtree rooted at the stack's top entry (at most kWaveSub points, not a leaf), walked by wave 0 alone:
    // the same pops, node numbers, draws, hyperplanes and partitions as the workgroup walk below, with the
    // subtree's positions held in LDS.  Returns with the entry and everything it pushed popped.
    auto subtree_wave = [&]() {
        int sp = sh_sp;
        const int sp_end = sp - 1;
        const StackEnt root = stack[sp - 1];
        const int o0 = root.start, M = root.end - root.start;
        for (int i = lane; i < M; i += RINND_WAVE) wv_idx[i] = idx[o0 + i];
        wave_sync();
        const unsigned long long lt = lane ? (~0ull >> (64 - lane)) : 0ull;
        while (sp > sp_end) {
            const StackEnt c = stack[sp - 1];
            sp -= 1;
            const int cm = c.end - c.start, lo = c.start - o0;
            int node = 0, leaf = 0, abort = 0, i1 = 0, i2 = 0;
            if (lane == 0) {
                node = sh_ctr;
                if (node >= cap) {
                    abort = 1;
                    sh_abort = 1;
                } else {
                    sh_ctr = node + 1;
                    if (c.pslot >= 0) ch[c.pslot] = node;
                }
                leaf = (cm <= a.leaf_size || c.depth >= a.max_depth) ? 1 : 0;
                if (!leaf && !abort) {  // make_split (builder.rs:228-247)
                    Xoshiro r;
                    for (int w = 0; w < 4; ++w) r.s[w] = rs[w];
                    const uint64_t j1 = r.next_index((uint64_t)cm);
                    uint64_t j2 = r.next_index((uint64_t)cm);
                    while (j2 == j1 && cm > 1) j2 = r.next_index((uint64_t)cm);
                    for (int w = 0; w < 4; ++w) rs[w] = r.s[w];
                    i1 = (int)j1;
                    i2 = (int)j2;
                }
            }
            abort = __shfl(abort, 0);
            if (abort) break;
            node = __shfl(node, 0);
            leaf = __shfl(leaf, 0);
            if (leaf) {
                if (lane == 0) {
                    ch[2 * node] = -c.start;
                    ch[2 * node + 1] = -c.end;
                }
                continue;
            }
            i1 = __shfl(i1, 0);
            i2 = __shfl(i2, 0);
            const float* x1 = X + (int64_t)wv_idx[lo + i1] * a.ld;
            const float* x2 = X + (int64_t)wv_idx[lo + i2] * a.ld;
            for (int i = lane; i < dim; i += RINND_WAVE) {
                p1s[i] = x1[i];
                p2s[i] = x2[i];
            }
            wave_sync();
            float off = 0.f;
            if (!a.angular) {  // make_euclidean_split (builder.rs:249-279)
                for (int i = lane; i < dim; i += RINND_WAVE) {
                    const float h = p2s[i] - p1s[i];
                    hp[i] = h;
                    q1s[i] = h * h;
                    q2s[i] = h * ((p1s[i] + p2s[i]) * 0.5f);
                }
                wave_sync();
                float inv = 0.f;
                int scale = 0;
                if (lane == 0) {
                    float norm_sq, dot_nm;
                    seq_sum2(q1s, q2s, dim, norm_sq, dot_nm);
                    const float norm = sqrtf(norm_sq);
                    if (norm > 1e-8f) {
                        inv = 1.0f / norm;
                        off = -dot_nm * inv;
                        scale = 1;
                    }
                }
                scale = __shfl(scale, 0);
                inv = __shfl(inv, 0);
                off = __shfl(off, 0);
                if (scale)
                    for (int i = lane; i < dim; i += RINND_WAVE) hp[i] *= inv;
            } else {  // make_angular_split (builder.rs:281-308)
                float n1 = 0.f, n2 = 0.f;
                for (int i = lane; i < dim; i += RINND_WAVE) {
                    q1s[i] = p1s[i] * p1s[i];
                    q2s[i] = p2s[i] * p2s[i];
                }
                wave_sync();
                if (lane == 0) {
                    seq_sum2(q1s, q2s, dim, n1, n2);
                    n1 = sqrtf(n1);
                    n2 = sqrtf(n2);
                }
                n1 = __shfl(n1, 0);
                n2 = __shfl(n2, 0);
                const bool deg = n1 < 1e-8f || n2 < 1e-8f;
                for (int i = lane; i < dim; i += RINND_WAVE) hp[i] = deg ? 0.f : p2s[i] / n2 - p1s[i] / n1;
                wave_sync();
                if (!deg) {
                    float nn = 0.f;
                    for (int i = lane; i < dim; i += RINND_WAVE) q1s[i] = hp[i] * hp[i];
                    wave_sync();
                    if (lane == 0) {
                        float unused;
                        seq_sum2(q1s, q1s, dim, nn, unused);
                        nn = sqrtf(nn);
                    }
                    nn = __shfl(nn, 0);
                    if (nn > 1e-8f)
                        for (int i = lane; i < dim; i += RINND_WAVE) hp[i] /= nn;
                }
            }
            wave_sync();
            // margins, one lane per point: the eight AVX2 accumulators and their reduction order
            // (dot_product_avx2, builder.rs:29-60)
            const int chunks = dim >> 3;
            bool tie = false;
            for (int b = 0; b < cm; b += RINND_WAVE) {
                const int p = b + lane;
                if (p < cm) {
                    const float* x = X + (int64_t)wv_idx[lo + p] * a.ld;
                    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                    for (int cc = 0; cc < chunks; ++cc) {
                        const float4 u = *reinterpret_cast<const float4*>(x + cc * 8);
                        const float4 v = *reinterpret_cast<const float4*>(x + cc * 8 + 4);
                        acc[0] = __builtin_fmaf(u.x, hp[cc * 8 + 0], acc[0]);
                        acc[1] = __builtin_fmaf(u.y, hp[cc * 8 + 1], acc[1]);
                        acc[2] = __builtin_fmaf(u.z, hp[cc * 8 + 2], acc[2]);
                        acc[3] = __builtin_fmaf(u.w, hp[cc * 8 + 3], acc[3]);
                        acc[4] = __builtin_fmaf(v.x, hp[cc * 8 + 4], acc[4]);
                        acc[5] = __builtin_fmaf(v.y, hp[cc * 8 + 5], acc[5]);
                        acc[6] = __builtin_fmaf(v.z, hp[cc * 8 + 6], acc[6]);
                        acc[7] = __builtin_fmaf(v.w, hp[cc * 8 + 7], acc[7]);
                    }
                    float r = ((acc[0] + acc[4]) + (acc[1] + acc[5])) + ((acc[2] + acc[6]) + (acc[3] + acc[7]));
                    for (int i = chunks * 8; i < dim; ++i) r += x[i] * hp[i];
                    const float margin = r + off;
                    uint8_t sd;
                    if (fabsf(margin) < 1e-8f) {
                        sd = 2;
                        tie = true;
                    } else {
                        sd = margin < 0.0f ? 0 : 1;
                    }
                    wv_side[lo + p] = sd;
                }
            }
            tie = __ballot(tie) != 0ull;
            wave_sync();
            int split = 0;
            if (tie) {
                if (lane == 0) {  // partition_inplace verbatim (builder.rs:319-347)
                    Xoshiro r;
                    for (int w = 0; w < 4; ++w) r.s[w] = rs[w];
                    int left = 0, right = cm;
                    while (left < right) {
                        const uint8_t sd = wv_side[lo + left];
                        const bool goes_left = sd == 2 ? !r.next_bool() : sd == 0;
                        if (goes_left) {
                            ++left;
                        } else {
                            --right;
                            const int32_t ti = wv_idx[lo + left];
                            wv_idx[lo + left] = wv_idx[lo + right];
                            wv_idx[lo + right] = ti;
                            const uint8_t ts = wv_side[lo + left];
                            wv_side[lo + left] = wv_side[lo + right];
                            wv_side[lo + right] = ts;
                        }
                    }
                    for (int w = 0; w < 4; ++w) rs[w] = r.s[w];
                    split = left;
                }
                split = __shfl(split, 0);
                wave_sync();
            } else {  // the closed form of the two-pointer partition, as in the workgroup walk
                int running = 0;
                for (int b = 0; b < cm; b += RINND_WAVE) {
                    const int p = b + lane;
                    const bool isR = p < cm && wv_side[lo + p] == 1;
                    const unsigned long long bal = __ballot(isR);
                    if (p < cm) wv_rank[p] = running + __popcll(bal & lt);
                    running += __popcll(bal);
                }
                const int totR = running;
                split = cm - totR;
                if (lane == 0) wv_rank[cm] = totR;
                wave_sync();
                for (int p = lane; p < cm; p += RINND_WAVE) {
                    const uint8_t sd = wv_side[lo + p];
                    if (sd == 1 && p < split) wv_P[wv_rank[p]] = p;
                    if (sd == 0 && p >= split) wv_Q[(cm - p - 1) - (totR - wv_rank[p + 1])] = p;
                }
                wave_sync();
                for (int p = lane; p < cm; p += RINND_WAVE) {
                    const uint8_t sd = wv_side[lo + p];
                    int dest;
                    if (sd == 0) {
                        dest = p < split ? p : wv_P[(cm - p - 1) - (totR - wv_rank[p + 1])];
                    } else {
                        const int rr = wv_rank[p];
                        dest = p <= split ? ((rr == 0 ? cm : wv_Q[rr - 1]) - 1) : p - 1;
                    }
                    wv_tmp[dest] = wv_idx[lo + p];
                }
                wave_sync();
                for (int p = lane; p < cm; p += RINND_WAVE) wv_idx[lo + p] = wv_tmp[p];
                wave_sync();
            }
            if (split == 0 || split == cm) {  // degenerate split -> leaf (builder.rs:198-210)
                if (lane == 0) {
                    ch[2 * node] = -c.start;
                    ch[2 * node + 1] = -c.end;
                }
            } else {
                for (int i = lane; i < dim; i += RINND_WAVE) hyper[(int64_t)node * dim + i] = hp[i];
                if (sp + 2 > kForestStack) {
                    if (lane == 0) sh_abort = 2;
                    break;
                }
                if (lane == 0) {
                    offs[node] = off;
                    stack[sp] = StackEnt{c.start + split, c.end, c.depth + 1, 2 * node + 1};
                    stack[sp + 1] = StackEnt{c.start, c.start + split, c.depth + 1, 2 * node};
                }
                sp += 2;
            }
            wave_sync();
        }
        wave_sync();
        for (int i = lane; i < M; i += RINND_WAVE) idx[o0 + i] = wv_idx[i];
        if (lane == 0) sh_sp = sp_end;
    };

    for (;;) {
        if (sh_sp == 0 || sh_abort) break;
        const StackEnt e = stack[sh_sp - 1];
        __syncthreads();
        const int m = e.end - e.start;
        if constexpr (WAVE_SUB) {
            if (m <= kWaveSub && m > a.leaf_size && e.depth < a.max_depth) {  // uniform
                if (wid == 0) subtree_wave();
                __syncthreads();
                continue;
            }
        }
        if (tid == 0) {
            sh_sp -= 1;
            const int node = sh_ctr;
            if (node >= cap) {
                sh_abort = 1;
            } else {
                sh_ctr = node + 1;
                if (e.pslot >= 0) ch[e.pslot] = node;
            }
            sh_node = node;
            sh_leaf = (m <= a.leaf_size || e.depth >= a.max_depth) ? 1 : 0;
            sh_tie = 0;
            if (!sh_leaf && !sh_abort) {  // make_split (builder.rs:228-247)
                Xoshiro r;
                for (int w = 0; w < 4; ++w) r.s[w] = rs[w];
                const uint64_t i1 = r.next_index((uint64_t)m);
                uint64_t i2 = r.next_index((uint64_t)m);
                while (i2 == i1 && m > 1) i2 = r.next_index((uint64_t)m);
                for (int w = 0; w < 4; ++w) rs[w] = r.s[w];
                sh_i1 = (int)i1;
                sh_i2 = (int)i2;
            }
        }
        __syncthreads();
        if (sh_abort) break;
        const int node = sh_node;
        if (sh_leaf) {
            if (tid == 0) {
                ch[2 * node] = -e.start;
                ch[2 * node + 1] = -e.end;
            }
            continue;  // next iteration re-reads sh_sp after the barrier below
        }
        const float* x1 = X + (int64_t)idx[e.start + sh_i1] * a.ld;
        const float* x2 = X + (int64_t)idx[e.start + sh_i2] * a.ld;
        for (int i = tid; i < dim; i += blockDim.x) {
            p1s[i] = x1[i];
            p2s[i] = x2[i];
        }
        __syncthreads();
        if (!a.angular) {  // make_euclidean_split (builder.rs:249-279)
            for (int i = tid; i < dim; i += blockDim.x) {
                const float h = p2s[i] - p1s[i];
                hp[i] = h;
                q1s[i] = h * h;
                q2s[i] = h * ((p1s[i] + p2s[i]) * 0.5f);
            }
            __syncthreads();
            if (tid == 0) {
                float norm_sq, dot_nm;
                seq_sum2(q1s, q2s, dim, norm_sq, dot_nm);
                const float norm = sqrtf(norm_sq);
                if (norm > 1e-8f) {
                    const float inv = 1.0f / norm;
                    sh_inv = inv;
                    sh_off = -dot_nm * inv;
                    sh_scale = 1;
                } else {
                    sh_off = 0.f;
                    sh_scale = 0;
                }
            }
            __syncthreads();
            if (sh_scale)
                for (int i = tid; i < dim; i += blockDim.x) hp[i] *= sh_inv;
        } else {  // make_angular_split (builder.rs:281-308)
            for (int i = tid; i < dim; i += blockDim.x) {
                q1s[i] = p1s[i] * p1s[i];
                q2s[i] = p2s[i] * p2s[i];
            }
            __syncthreads();
            if (tid == 0) {
                float n1, n2;
                seq_sum2(q1s, q2s, dim, n1, n2);
                sh_n1 = sqrtf(n1);
                sh_n2 = sqrtf(n2);
                sh_off = 0.f;
            }
            __syncthreads();
            const bool deg = sh_n1 < 1e-8f || sh_n2 < 1e-8f;
            for (int i = tid; i < dim; i += blockDim.x)
                hp[i] = deg ? 0.f : p2s[i] / sh_n2 - p1s[i] / sh_n1;
            __syncthreads();
            if (!deg) {
                for (int i = tid; i < dim; i += blockDim.x) q1s[i] = hp[i] * hp[i];
                __syncthreads();
                if (tid == 0) {
                    float nn, unused;
                    seq_sum2(q1s, q1s, dim, nn, unused);
                    sh_n1 = sqrtf(nn);
                }
                __syncthreads();
                const float nn = sh_n1;
                if (nn > 1e-8f)
                    for (int i = tid; i < dim; i += blockDim.x) hp[i] /= nn;
            }
        }
        __syncthreads();
        bool coop_done = false;
        if constexpr (COOP) {
            if (m >= a.coop_min) {  // uniform: the tree's other workgroups join this node
                CoopCtl* cc = a.coop + t;
                float* chp = a.coop_hp + (int64_t)t * dim;
                for (int i = tid; i < dim; i += blockDim.x) chp[i] = hp[i];
                if (tid == 0) {
                    cc->start = e.start;
                    cc->end = e.end;
                    cc->off = sh_off;
                }
                __threadfence();
                __syncthreads();
                if (tid == 0) {
                    sh_seq += 1;
                    __hip_atomic_store(&cc->seq, sh_seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                }
                const int r = coop_node<THREADS>(X, a.ld, dim, hp, sh_off, e.start, m, 0, a.coop_groups, cc, idx, side,
                                                 rankR, P, Q, tmp, &sh_nb);
                if (r == -2) {  // uniform
                    if (tid == 0) sh_abort = 3;
                    __syncthreads();
                    break;
                }
                if (tid == 0) {
                    sh_tie = r == -1 ? 1 : 0;
                    sh_split = r;
                }
                __syncthreads();
                coop_done = true;
            }
        }
        // margins: two lanes per point, lane h holding the AVX2 accumulators 4h .. 4h+3 of the point's 8-wide
        // blocks (16-byte loads, a point's whole row in flight), combined in the AVX2 reduction order
        // (dot_product_avx2, builder.rs:29-60)
        if (!coop_done) {
            const float off = sh_off;
            const int g = tid >> 1, h = tid & 1;
            const int groups = blockDim.x >> 1;
            const int chunks = dim >> 3;
            int32_t nxt = e.start + g < e.end ? idx[e.start + g] : 0;  // the next point's id, one step ahead
            for (int base = e.start; base < e.end; base += groups) {
                const int pos = base + g;
                const bool act = pos < e.end;
                const int32_t cur = nxt;
                nxt = pos + groups < e.end ? idx[pos + groups] : 0;
                const float* x = X + (int64_t)cur * a.ld;
                float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
                if (act)
                    for (int c = 0; c < chunks; ++c) {
                        const float4 u = *reinterpret_cast<const float4*>(x + c * 8 + 4 * h);
                        const float4 w = *reinterpret_cast<const float4*>(hp + c * 8 + 4 * h);
                        a0 = __builtin_fmaf(u.x, w.x, a0);
                        a1 = __builtin_fmaf(u.y, w.y, a1);
                        a2 = __builtin_fmaf(u.z, w.z, a2);
                        a3 = __builtin_fmaf(u.w, w.w, a3);
                    }
                const float b0 = __shfl_xor(a0, 1), b1 = __shfl_xor(a1, 1);
                const float b2 = __shfl_xor(a2, 1), b3 = __shfl_xor(a3, 1);
                if (act && h == 0) {
                    float r = ((a0 + b0) + (a1 + b1)) + ((a2 + b2) + (a3 + b3));
                    for (int i = chunks * 8; i < dim; ++i) r += x[i] * hp[i];
                    const float margin = r + off;
                    uint8_t sd;
                    if (fabsf(margin) < 1e-8f) {
                        sd = 2;
                        sh_tie = 1;
                    } else {
                        sd = margin < 0.0f ? 0 : 1;
                    }
                    side[pos - e.start] = sd;
                }
            }
        }
        __syncthreads();
        // rankR[p] = number of right elements (side 1) before p, p in [0, m] (exclusive scan of 1-bit flags)
        auto rank_pass = [&]() {
            if (tid == 0) sh_running = 0;
            __syncthreads();
            for (int base = 0; base < m; base += blockDim.x) {
                const int p = base + tid;
                const bool isR = p < m && side[p] == 1;
                const unsigned long long bal = __ballot(isR);
                const unsigned long long lt = lane ? (~0ull >> (64 - lane)) : 0ull;
                const int lpre = __popcll(bal & lt);
                if (lane == 0) wave_tot[wid] = __popcll(bal);
                __syncthreads();
                int wpre = 0;
                for (int w = 0; w < wid; ++w) wpre += wave_tot[w];
                if (p < m) rankR[p] = sh_running + wpre + lpre;
                __syncthreads();
                if (tid == 0) {
                    int tot = 0;
                    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) tot += wave_tot[w];
                    sh_running += tot;
                }
                __syncthreads();
            }
            if (tid == 0) rankR[m] = sh_running;
            __syncthreads();
        };
        // the two-pointer partition of sides 0 / 1 in closed form
        auto closed_form = [&]() {
            rank_pass();
            const int totR = sh_running;
            const int split = m - totR;
            for (int p = tid; p < m; p += blockDim.x) {
                const uint8_t sd = side[p];
                if (sd == 1 && p < split) P[rankR[p]] = p;
                if (sd == 0 && p >= split) Q[(m - p - 1) - (totR - rankR[p + 1])] = p;
            }
            __syncthreads();
            for (int p = tid; p < m; p += blockDim.x) {
                const uint8_t sd = side[p];
                int dest;
                if (sd == 0) {
                    dest = p < split ? p : P[(m - p - 1) - (totR - rankR[p + 1])];
                } else {
                    const int r = rankR[p];
                    dest = p <= split ? ((r == 0 ? m : Q[r - 1]) - 1) : p - 1;
                }
                tmp[dest] = idx[e.start + p];
            }
            __syncthreads();
            for (int p = tid; p < m; p += blockDim.x) idx[e.start + p] = tmp[p];
            if (tid == 0) sh_split = split;
            __syncthreads();
        };
        if (sh_tie) {  // partition_inplace with margin ties (builder.rs:319-347)
            if (tid == 0) sh_ntie = 0;
            __syncthreads();
            for (int p = tid; p < m; p += blockDim.x)
                if (side[p] == 2) {
                    const int slot = atomicAdd(&sh_ntie, 1);
                    if (slot < kMaxTies) sh_ties[slot] = p;
                }
            __syncthreads();
            if (sh_ntie <= kMaxTies) {  // few ties: decide them in examination order, then the closed form
                rank_pass();
                if (tid == 0) {
                    const int T = sh_ntie;
                    for (int i = 1; i < T; ++i)  // ascending positions
                        for (int j = i; j > 0 && sh_ties[j - 1] > sh_ties[j]; --j) {
                            const int x = sh_ties[j];
                            sh_ties[j] = sh_ties[j - 1];
                            sh_ties[j - 1] = x;
                        }
                    tie_resolve(rankR, m, sh_ties, T, sh_dec, rs);
                    for (int i = 0; i < T; ++i) side[sh_ties[i]] = (uint8_t)sh_dec[i];
                }
                __syncthreads();
                closed_form();
            } else {  // many ties: replay the examination order element by element on wave 0
                if (wid == 0) {
                    const int left = tie_partition_wave(side, m, rs, rankR);
                    if (tid == 0) sh_split = left;
                }
                __syncthreads();
                for (int p = tid; p < m; p += blockDim.x) tmp[rankR[p]] = idx[e.start + p];
                __syncthreads();
                for (int p = tid; p < m; p += blockDim.x) idx[e.start + p] = tmp[p];
                __syncthreads();
            }
        } else if (!coop_done) {
            closed_form();
        }
        const int split = sh_split;
        if (split == 0 || split == m) {  // degenerate split -> leaf (builder.rs:198-210)
            if (tid == 0) {
                ch[2 * node] = -e.start;
                ch[2 * node + 1] = -e.end;
            }
        } else {
            for (int i = tid; i < dim; i += blockDim.x) hyper[(int64_t)node * dim + i] = hp[i];
            if (tid == 0) {
                offs[node] = sh_off;
                if (sh_sp + 2 > kForestStack) {
                    sh_abort = 2;
                } else {
                    stack[sh_sp] = StackEnt{e.start + split, e.end, e.depth + 1, 2 * node + 1};
                    stack[sh_sp + 1] = StackEnt{e.start, e.start + split, e.depth + 1, 2 * node};
                    sh_sp += 2;
                }
            }
        }
        __syncthreads();
    }
    __syncthreads();
    if (tid == 0) {
        a.n_nodes_out[t] = sh_ctr;
        a.status[t] = sh_abort;
        if constexpr (COOP) {  // the walk is over: the other workgroups leave
            __threadfence();
            __hip_atomic_store(&a.coop[t].done, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    __syncthreads();  // the workgroup's shared state is reused by its next tree
}

// Workgroup g >= 1 of a tree in the cooperative form: joins each node the walking workgroup publishes.
template <int THREADS>
__device__ void forest_coop_helper(const ForestDFSArgs& a, const int t, const int g) {
    extern __shared__ __align__(16) float fsm[];
    __shared__ int hs_go, hs_nb, hs_start, hs_end;
    __shared__ float hs_off;
    const int tid = threadIdx.x;
    const int64_t n = a.n;
    const int dim = a.dim;
    CoopCtl* cc = a.coop + t;
    const float* chp = a.coop_hp + (int64_t)t * dim;
    const int64_t so = (int64_t)t * (n + 1);
    int32_t* idx = a.indices[t];
    uint8_t* side = a.side + so;
    int32_t* rankR = a.rankR + so;
    int32_t* P = a.PQ + 2 * so;
    int32_t* Q = P + (n + 1);
    int32_t* tmp = a.tmp + so;
    float* hp = fsm;
    if (tid == 0) hs_nb = 0;
    int last = 0;
    for (;;) {
        if (tid == 0) {  // the next published node, or the end of the walk
            const uint64_t t0 = wall_clock64();
            int go = 0;
            for (;;) {  // relaxed polls, one acquire fence on success (as coop_wait)
                if (__hip_atomic_load(&cc->seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > last) {
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                    go = 1;
                    break;
                }
                if (__hip_atomic_load(&cc->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                    __hip_atomic_load(&cc->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    break;
                if (wall_clock64() - t0 > kCoopTimeoutTicks) {
                    __hip_atomic_store(&cc->abort, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(4);
            }
            if (go) {
                hs_start = __hip_atomic_load(&cc->start, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                hs_end = __hip_atomic_load(&cc->end, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                hs_off = __hip_atomic_load(&cc->off, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            hs_go = go;
        }
        __syncthreads();
        if (!hs_go) break;  // uniform
        ++last;
        for (int i = tid; i < dim; i += blockDim.x) hp[i] = chp[i];
        __syncthreads();
        const int r = coop_node<THREADS>(a.X, a.ld, dim, hp, hs_off, hs_start, hs_end - hs_start, g, a.coop_groups, cc,
                                         idx, side, rankR, P, Q, tmp, &hs_nb);
        __syncthreads();  // hp and the node fields are rewritten for the next node
        if (r == -2) break;
    }
}

// One workgroup of 1024 threads per tree.
__global__ __launch_bounds__(kForestThreads) void k_forest_dfs(ForestDFSArgs a) {
    forest_tree_dfs<kForestThreads, false>(a, blockIdx.x, (int64_t)blockIdx.x * (a.n + 1));
}
// Cooperative form: coop_groups workgroups per tree (the first walks the tree, the others join its large nodes).
__global__ __launch_bounds__(kForestThreads) void k_forest_dfs_coop(ForestDFSArgs a) {
    const int t = blockIdx.x / a.coop_groups, g = blockIdx.x - t * a.coop_groups;
    if (g == 0)
        forest_tree_dfs<kForestThreads, false, true>(a, t, (int64_t)t * (a.n + 1));
    else
        forest_coop_helper<kForestThreads>(a, t, g);
}
// Persistent workgroups of 256 threads pulling trees from a work queue, largest first (many trees: more of them
// in flight per CU).
template <int THREADS>
__global__ __launch_bounds__(THREADS) void k_forest_dfs_queue(ForestDFSArgs a) {
    __shared__ int sh_item;
    for (;;) {
        if (threadIdx.x == 0) sh_item = atomicAdd(a.queue, 1);
        __syncthreads();
        const int item = sh_item;
        __syncthreads();
        if (item >= a.n_trees) break;  // uniform: every workgroup ends once the queue is drained
        forest_tree_dfs<THREADS, true>(a, a.order[item], (int64_t)blockIdx.x * a.slot_stride);
    }
}
constexpr int kForestQueueThreads = 256;

int forest_queue_grid(int dim) {
    int per_cu = 0;
    const size_t lds = (size_t)5 * dim * sizeof(float);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_forest_dfs_queue<kForestQueueThreads>,
                                                     kForestQueueThreads, lds) != hipSuccess || per_cu < 1) {
        (void)hipGetLastError();
        per_cu = 1;
    }
    return per_cu * device_cus();
}

int forest_coop_groups(int dim, int n_trees) {
    int per_cu = 0;
    const size_t lds = (size_t)5 * dim * sizeof(float);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_forest_dfs_coop, kForestThreads, lds) != hipSuccess ||
        per_cu < 1) {
        (void)hipGetLastError();
        return 0;
    }
    const int G = std::min(kCoopMaxGroups, per_cu * device_cus() / std::max(n_trees, 1));
    return G >= 2 ? G : 0;
}

void launch_forest_dfs(hipStream_t s, const ForestDFSArgs& a) {
    const size_t lds = (size_t)5 * a.dim * sizeof(float);
    if (a.coop) {
        hipLaunchKernelGGL(k_forest_dfs_coop, dim3(a.n_trees * a.coop_groups), dim3(kForestThreads), lds, s, a);
    } else if (a.queue) {
        const unsigned grid = (unsigned)std::max(1, std::min(a.grid, a.n_trees));
        hipLaunchKernelGGL(k_forest_dfs_queue<kForestQueueThreads>, dim3(grid), dim3(kForestQueueThreads), lds, s, a);
    } else {
        hipLaunchKernelGGL(k_forest_dfs, dim3(a.n_trees), dim3(kForestThreads), lds, s, a);
    }
}

// ---- leaves in depth-first order -------------------------------------------------------------------
__global__ void k_leaf_mark(const int32_t* ch, int64_t n_nodes, int32_t* mark, int32_t* end_at) {
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n_nodes;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int32_t c0 = ch[2 * v];
        if (c0 <= 0) {
            const int32_t s = -c0, e = -ch[2 * v + 1];
            if (e > s) {
                mark[s] = 1;
                end_at[s] = e;
            }
        }
    }
}
__global__ void k_i32_to_u64(const int32_t* in, unsigned long long* out, int64_t n) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x)
        out[t] = (unsigned long long)in[t];
}
__global__ void k_leaf_emit(const int32_t* mark, const uint64_t* rank, const int32_t* end_at,
                            int64_t n, int64_t* leaves) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n;
         p += (int64_t)gridDim.x * blockDim.x)
        if (mark[p]) {
            const uint64_t r = rank[p];
            leaves[2 * r] = p;
            leaves[2 * r + 1] = end_at[p];
        }
}
void launch_leaf_list(hipStream_t s, const int32_t* children, int64_t n_nodes, int64_t n,
                      int32_t* leaf_mark, int64_t* leaf_rank, int32_t* leaf_end_at, int64_t* leaves,
                      void* scan_tmp, size_t scan_tmp_bytes) {
    HIP_TRY(hipMemsetAsync(leaf_mark, 0, (n + 1) * sizeof(int32_t), s));
    hipLaunchKernelGGL(k_leaf_mark, dim3(grid_for(n_nodes, 256)), dim3(256), 0, s, children, n_nodes,
                       leaf_mark, leaf_end_at);
    // widen marks into the second half of the rank buffer (2(n+1) words), scan into the first half;
    // rank[n] = number of leaves
    unsigned long long* wide = (unsigned long long*)(leaf_rank + (n + 1));
    hipLaunchKernelGGL(k_i32_to_u64, dim3(grid_for(n + 1, 256)), dim3(256), 0, s, leaf_mark, wide, n + 1);
    size_t bytes = scan_tmp_bytes;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(scan_tmp, bytes, wide, (unsigned long long*)leaf_rank,
                                             (int)(n + 1), s));
    hipLaunchKernelGGL(k_leaf_emit, dim3(grid_for(n, 256)), dim3(256), 0, s, leaf_mark,
                       (const uint64_t*)leaf_rank, leaf_end_at, n, leaves);
}

// Leaves of a virtual tree: tree number t of every index of a batch that has one, laid side by side.  Entry e
// places its tree's positions at pbase .. pbase + n (ids offset by the index's first row, base) and its leaves
// (start, end) at those positions; each position is in exactly one leaf of its own index's tree.
__global__ void k_vtree_build(const VTreeEnt* __restrict__ ents, int32_t* mark, int32_t* end_at, int32_t* vidx) {
    const VTreeEnt e = ents[blockIdx.y];
    const int64_t span = e.n_nodes > e.n ? e.n_nodes : e.n;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < span; v += (int64_t)gridDim.x * blockDim.x) {
        if (v < e.n_nodes) {
            const int32_t c0 = e.children[2 * v];
            if (c0 <= 0) {
                const int32_t st = -c0, en = -e.children[2 * v + 1];
                if (en > st) {
                    mark[e.pbase + st] = 1;
                    end_at[e.pbase + st] = (int32_t)(e.pbase + en);
                }
            }
        }
        if (v < e.n) vidx[e.pbase + v] = (int32_t)(e.base + e.indices[v]);
    }
}
void launch_leaf_list_batch(hipStream_t s, const VTreeEnt* ents_dev, int n_ents, int64_t max_span, int64_t n,
                            int32_t* leaf_mark, int64_t* leaf_rank, int32_t* leaf_end_at, int64_t* leaves,
                            int32_t* vidx, void* scan_tmp, size_t scan_tmp_bytes) {
    launch_zero(s, leaf_mark, (size_t)(n + 1) * sizeof(int32_t));
    if (n_ents > 0)
        hipLaunchKernelGGL(k_vtree_build, dim3(grid_for(max_span, 256, 512), n_ents), dim3(256), 0, s, ents_dev,
                           leaf_mark, leaf_end_at, vidx);
    unsigned long long* wide = (unsigned long long*)(leaf_rank + (n + 1));
    hipLaunchKernelGGL(k_i32_to_u64, dim3(grid_for(n + 1, 256)), dim3(256), 0, s, leaf_mark, wide, n + 1);
    size_t bytes = scan_tmp_bytes;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(scan_tmp, bytes, wide, (unsigned long long*)leaf_rank, (int)(n + 1), s));
    hipLaunchKernelGGL(k_leaf_emit, dim3(grid_for(n, 256)), dim3(256), 0, s, leaf_mark, (const uint64_t*)leaf_rank,
                       leaf_end_at, n, leaves);
}

__global__ void k_leaf_pair_counts(const int64_t* leaves, int64_t n_leaves, uint64_t* out) {
    for (int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; l <= n_leaves;
         l += (int64_t)gridDim.x * blockDim.x) {
        if (l == n_leaves) {
            out[l] = 0;
        } else {
            const uint64_t m = (uint64_t)(leaves[2 * l + 1] - leaves[2 * l]);
            out[l] = m * (m - 1) / 2;
        }
    }
}
void launch_leaf_pair_base(hipStream_t s, const int64_t* leaves, int64_t n_leaves,
                           uint64_t* pair_base_in, uint64_t* pair_base_out, void* scan_tmp,
                           size_t scan_tmp_bytes) {
    hipLaunchKernelGGL(k_leaf_pair_counts, dim3(grid_for(n_leaves + 1, 256)), dim3(256), 0, s, leaves,
                       n_leaves, pair_base_in);
    size_t bytes = scan_tmp_bytes;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(scan_tmp, bytes, (unsigned long long*)pair_base_in,
                                             (unsigned long long*)pair_base_out, (int)(n_leaves + 1), s));
}

// ---- emission helper: wave-aggregated append into U ---------------------------------------------
__device__ __forceinline__ void emit_update(bool keep, uint64_t key, int32_t p, int32_t q, float d,
                                            URec* U, unsigned long long* U_count, uint64_t U_cap,
                                            unsigned long long* cnt, int32_t* overflow) {
    const unsigned long long bal = __ballot(keep);
    if (bal == 0) return;
    const int lane = lane_id();
    const int total = __popcll(bal);
    const int leader = __ffsll((long long)bal) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(U_count, (unsigned long long)total);
    base = __shfl(base, leader);
    if (keep) {
        const unsigned long long lt = lane ? (~0ull >> (64 - lane)) : 0ull;
        const uint64_t pos = base + __popcll(bal & lt);
        if (pos < U_cap) {
            U[pos] = URec{key, p, q, d, 1};
            atomicAdd(&cnt[p], 1ull);
            atomicAdd(&cnt[q], 1ull);
        } else {
            *overflow = 1;
        }
    }
}

// ---- leaf initialisation pairs (initialize_from_leaves, mod.rs:239-266) ----------------------------
// One workgroup per leaf; pair (i, j), i < j, key = pair_base[leaf] + rank of (i, j) in the leaf's
// row-major upper triangle, so leaves keep forest order and pairs keep loop order.
__global__ __launch_bounds__(256) void k_leaf_pairs(int metric, const float* __restrict__ X,
                                                    int64_t ld, int dim,
                                                    const int32_t* __restrict__ tidx,
                                                    const int64_t* __restrict__ leaves,
                                                    int64_t n_leaves, const uint64_t* pair_base,
                                                    const float* __restrict__ hd, int k, URec* U,
                                                    unsigned long long* U_count, uint64_t U_cap,
                                                    unsigned long long* cnt, int32_t* overflow) {
    for (int64_t leaf = blockIdx.x; leaf < n_leaves; leaf += gridDim.x) {
        const int64_t s = leaves[2 * leaf], e = leaves[2 * leaf + 1];
        const int64_t m = e - s;
        const uint64_t base = pair_base[leaf];
        const int64_t total = m * m;
        for (int64_t t0 = 0; t0 < total; t0 += blockDim.x) {
            const int64_t t = t0 + threadIdx.x;
            const int64_t i = t / (m ? m : 1), j = t - i * m;
            bool keep = false;
            int32_t p = 0, q = 0;
            float d = 0.f;
            uint64_t key = 0;
            if (t < total && j > i) {
                p = tidx[s + i];
                q = tidx[s + j];
                d = distance_exact(metric, X + (int64_t)p * ld, X + (int64_t)q * ld, dim);
                const float thr = fmaxf(hd[(int64_t)p * k], hd[(int64_t)q * k]);
                keep = d < thr;
                key = base + (uint64_t)(i * m - i * (i + 1) / 2 + (j - i - 1));
            }
            emit_update(keep, key, p, q, d, U, U_count, U_cap, cnt, overflow);
        }
    }
}
void launch_leaf_pairs(hipStream_t s, int metric, const float* X, int64_t ld, int dim,
                       const int32_t* tree_indices, const int64_t* leaves, int64_t n_leaves,
                       const uint64_t* pair_base, const float* heap_dist, int k, URec* U,
                       unsigned long long* U_count, uint64_t U_cap, unsigned long long* cnt,
                       int32_t* overflow) {
    if (n_leaves == 0) return;
    hipLaunchKernelGGL(k_leaf_pairs, dim3(grid_for(n_leaves, 1, 1 << 16)), dim3(256), 0, s, metric, X,
                       ld, dim, tree_indices, leaves, n_leaves, pair_base, heap_dist, k, U, U_count,
                       U_cap, cnt, overflow);
}

// ---- scatter by destination -----------------------------------------------------------------------
__global__ void k_scatter(const URec* __restrict__ U, const unsigned long long* U_count,
                          const uint64_t* __restrict__ off, unsigned long long* cur, SegRec* seg) {
    const uint64_t total = *U_count;
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < total;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const URec u = U[t];
        if (u.a < 0) continue;
        const uint64_t sa = off[u.a] + atomicAdd(&cur[u.a], 1ull);
        seg[sa] = SegRec{u.key, u.b, u.v};
        if (u.two_sided) {
            const uint64_t sb = off[u.b] + atomicAdd(&cur[u.b], 1ull);
            seg[sb] = SegRec{u.key, u.a, u.v};
        }
    }
}
void launch_scatter(hipStream_t s, const URec* U, const unsigned long long* U_count,
                    const uint64_t* off, unsigned long long* cur, SegRec* seg) {
    hipLaunchKernelGGL(k_scatter, dim3(4096), dim3(256), 0, s, U, U_count, off, cur, seg);
}

// ---- per-destination sort + sequential apply ------------------------------------------------------
constexpr int kApplyWaves = 4;
constexpr int kSortCap = 512;           // records sorted in LDS by one wave
constexpr int kApplyScratch = 4096;     // per-wave heap scratch bytes
constexpr int kPerWave = kSortCap * 16 + kApplyScratch;
constexpr int kBigThreads = 1024;
constexpr int kBigChunk = 2048;         // records per LDS chunk in the large-group path

__device__ __forceinline__ int next_pow2(int64_t x) {
    int p = 1;
    while (p < x) p <<= 1;
    return p;
}

__device__ __forceinline__ bool rec_gt(const SegRec& a, const SegRec& b) { return a.key > b.key; }

// checked_flagged_push (neighbor_heap.rs:149-177) on an LDS row; returns true on insertion.
__device__ __forceinline__ bool heap_push(int32_t* hi, float* hd, uint8_t* hf, int k, int32_t x,
                                          float d, uint8_t flag) {
    if (d >= hd[0]) return false;
    for (int j = 0; j < k; ++j)
        if (hi[j] == x) return false;
    // replace root, sift_down (neighbor_heap.rs:207-243)
    int pos = 0, child = 1;
    while (child < k) {
        const int right = child + 1;
        if (right < k && hd[child] < hd[right]) child = right;
        if (d >= hd[child]) break;
        hd[pos] = hd[child];
        hi[pos] = hi[child];
        hf[pos] = hf[child];
        pos = child;
        child = 2 * pos + 1;
    }
    hd[pos] = d;
    hi[pos] = x;
    hf[pos] = flag;
    return true;
}

// Where a destination's sorted records live: one per lane (groups of <= 64) or in memory.
struct RecSrc {
    const SegRec* mem;
    int32_t rx;
    float rv;
    bool in_regs;
    __device__ __forceinline__ void get(int r, int32_t& x, float& v) const {
        if (in_regs) {
            x = __builtin_amdgcn_readlane(rx, r);
            v = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int32_t, rv), r));
        } else {
            const SegRec q = mem[r];
            x = q.x;
            v = q.v;
        }
    }
};

__device__ __forceinline__ float readlane_f(float x, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int32_t, x), l));
}

struct HeapApply {
    int32_t* idx;
    float* dist;
    uint8_t* flags;
    int k;
    unsigned long long* changes;
    __device__ static bool kRegs(const HeapApply& a) { return a.k <= RINND_WAVE; }
    // k <= 64: heap slot j lives in lane j; the duplicate scan is one ballot and the sift-down of
    // neighbor_heap.rs:207-243 walks the heap with uniform readlane reads.
    __device__ void apply_regs(int64_t v, const RecSrc& src, int L, int lane) const {
        const int64_t o = v * k;
        float hd = lane < k ? dist[o + lane] : __builtin_inff();
        int32_t hi = lane < k ? idx[o + lane] : -1;
        int32_t hf = lane < k ? flags[o + lane] : 0;
        unsigned long long ch = 0;
        for (int r = 0; r < L; ++r) {
            int32_t x;
            float d;
            src.get(r, x, d);
            if (d >= readlane_f(hd, 0)) continue;          // checked_flagged_push: reject d >= root
            if (__ballot(lane < k && hi == x)) continue;   // duplicate id
            int pos = 0;
            for (;;) {
                int child = 2 * pos + 1;
                if (child >= k) break;
                float dc = readlane_f(hd, child);
                if (child + 1 < k) {
                    const float dr = readlane_f(hd, child + 1);
                    if (dc < dr) {
                        child = child + 1;
                        dc = dr;
                    }
                }
                if (d >= dc) break;
                const int32_t ic = __builtin_amdgcn_readlane(hi, child);
                const int32_t fc = __builtin_amdgcn_readlane(hf, child);
                if (lane == pos) {
                    hd = dc;
                    hi = ic;
                    hf = fc;
                }
                pos = child;
            }
            if (lane == pos) {
                hd = d;
                hi = x;
                hf = 1;
            }
            ++ch;
        }
        if (lane < k) {
            idx[o + lane] = hi;
            dist[o + lane] = hd;
            flags[o + lane] = (uint8_t)hf;
        }
        if (lane == 0 && ch) atomicAdd(changes, ch);
    }
    __device__ void apply(int64_t v, const SegRec* recs, int64_t L, char* scratch, int lane) const {
        if (k <= RINND_WAVE) {
            apply_regs(v, RecSrc{recs, 0, 0.f, false}, (int)L, lane);
            return;
        }
        int32_t* hi = (int32_t*)scratch;
        float* hd = (float*)(hi + k);
        uint8_t* hf = (uint8_t*)(hd + k);
        const int64_t o = v * k;
        for (int j = lane; j < k; j += RINND_WAVE) {
            hi[j] = idx[o + j];
            hd[j] = dist[o + j];
            hf[j] = flags[o + j];
        }
        wave_sync();
        if (lane == 0) {
            unsigned long long ch = 0;
            for (int64_t r = 0; r < L; ++r) {
                const SegRec rec = recs[r];
                if (heap_push(hi, hd, hf, k, rec.x, rec.v, 1)) ++ch;
            }
            if (ch) atomicAdd(changes, ch);
        }
        wave_sync();
        for (int j = lane; j < k; j += RINND_WAVE) {
            idx[o + j] = hi[j];
            dist[o + j] = hd[j];
            flags[o + j] = hf[j];
        }
    }
};

template <class Ap>
__global__ __launch_bounds__(kApplyWaves * RINND_WAVE) void k_seg_apply(
    const uint64_t* __restrict__ off, const SegRec* __restrict__ seg, int64_t n_dest, Ap ap,
    int32_t* big_list, int32_t* big_count) {
    extern __shared__ __align__(16) char asm_[];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    SegRec* recs = (SegRec*)(asm_ + w * kPerWave);
    char* scratch = asm_ + w * kPerWave + kSortCap * 16;
    for (int64_t v = (int64_t)blockIdx.x * kApplyWaves + w; v < n_dest;
         v += (int64_t)gridDim.x * kApplyWaves) {
        const uint64_t o0 = off[v];
        const int64_t L = (int64_t)(off[v + 1] - o0);
        if (L == 0) continue;
        if (L > kSortCap) {
            if (lane == 0) big_list[atomicAdd(big_count, 1)] = (int32_t)v;
            continue;
        }
        if (L <= RINND_WAVE && Ap::kRegs(ap)) {
            // one record per lane, bitonic network across lanes (keys are unique; pads sort last)
            uint64_t key = ~0ull;
            int32_t rx = 0;
            float rv = 0.f;
            if (lane < L) {
                const SegRec q = seg[o0 + lane];
                key = q.key;
                rx = q.x;
                rv = q.v;
            }
            for (int kk = 2; kk <= RINND_WAVE; kk <<= 1)
                for (int j = kk >> 1; j > 0; j >>= 1) {
                    const uint64_t pk = __shfl_xor(key, j);
                    const int32_t px = __shfl_xor(rx, j);
                    const float pv = __shfl_xor(rv, j);
                    const bool up = (lane & kk) == 0;
                    const bool lower = (lane & j) == 0;
                    const bool gt = key > pk;
                    if (key != pk && (lower ? (gt == up) : (gt != up))) {
                        key = pk;
                        rx = px;
                        rv = pv;
                    }
                }
            ap.apply_regs(v, RecSrc{nullptr, rx, rv, true}, (int)L, lane);
            continue;
        }
        const int S = next_pow2(L);
        for (int i = lane; i < S; i += RINND_WAVE)
            recs[i] = i < L ? seg[o0 + i] : SegRec{~0ull, 0, 0.f};
        wave_sync();
        for (int kk = 2; kk <= S; kk <<= 1)
            for (int j = kk >> 1; j > 0; j >>= 1) {
                for (int i = lane; i < S; i += RINND_WAVE) {
                    const int ixj = i ^ j;
                    if (ixj > i) {
                        const bool up = (i & kk) == 0;
                        const SegRec A = recs[i], B = recs[ixj];
                        if (rec_gt(A, B) == up) {
                            recs[i] = B;
                            recs[ixj] = A;
                        }
                    }
                }
                wave_sync();
            }
        ap.apply(v, recs, L, scratch, lane);
    }
}

// Large groups: chunks sorted in LDS, merged pairwise in global memory (merge path), then applied.
template <class Ap>
__global__ __launch_bounds__(kBigThreads) void k_seg_apply_big(const uint64_t* __restrict__ off,
                                                               SegRec* seg, SegRec* seg2, Ap ap,
                                                               const int32_t* big_list,
                                                               const int32_t* big_count) {
    extern __shared__ __align__(16) char bsm[];
    SegRec* buf = (SegRec*)bsm;
    char* scratch = bsm + kBigChunk * 16;
    const int nb = *big_count;
    const int tid = threadIdx.x;
    for (int b = blockIdx.x; b < nb; b += gridDim.x) {
        const int64_t v = big_list[b];
        const uint64_t o0 = off[v];
        const int64_t L = (int64_t)(off[v + 1] - o0);
        SegRec* A = seg + o0;
        SegRec* B = seg2 + o0;
        for (int64_t c0 = 0; c0 < L; c0 += kBigChunk) {
            const int64_t cl = min((int64_t)kBigChunk, L - c0);
            const int S = next_pow2(cl);
            for (int i = tid; i < S; i += blockDim.x) buf[i] = i < cl ? A[c0 + i] : SegRec{~0ull, 0, 0.f};
            __syncthreads();
            for (int kk = 2; kk <= S; kk <<= 1)
                for (int j = kk >> 1; j > 0; j >>= 1) {
                    for (int i = tid; i < S; i += blockDim.x) {
                        const int ixj = i ^ j;
                        if (ixj > i) {
                            const bool up = (i & kk) == 0;
                            const SegRec x = buf[i], y = buf[ixj];
                            if (rec_gt(x, y) == up) {
                                buf[i] = y;
                                buf[ixj] = x;
                            }
                        }
                    }
                    __syncthreads();
                }
            for (int i = tid; i < cl; i += blockDim.x) A[c0 + i] = buf[i];
            __syncthreads();
        }
        SegRec* src = A;
        SegRec* dst = B;
        for (int64_t w = kBigChunk; w < L; w *= 2) {
            for (int64_t o = tid; o < L; o += blockDim.x) {
                const int64_t base = (o / (2 * w)) * (2 * w);
                const int64_t la = min(w, L - base);
                const int64_t lb = max((int64_t)0, min(w, L - base - w));
                const int64_t ol = o - base;
                const SegRec* Aa = src + base;
                const SegRec* Bb = src + base + w;
                int64_t lo = max((int64_t)0, ol - lb), hi = min(ol, la);
                while (lo < hi) {
                    const int64_t mid = (lo + hi) >> 1;
                    if (Aa[mid].key < Bb[ol - 1 - mid].key) lo = mid + 1;
                    else hi = mid;
                }
                const int64_t i = lo, j = ol - lo;
                const bool takeA = i < la && (j >= lb || Aa[i].key < Bb[j].key);
                dst[o] = takeA ? Aa[i] : Bb[j];
            }
            __syncthreads();
            SegRec* t = src;
            src = dst;
            dst = t;
        }
        if (tid < RINND_WAVE) ap.apply(v, src, L, scratch, tid);
        __syncthreads();
    }
}

template <class Ap>
static void run_apply(hipStream_t s, const uint64_t* off, SegRec* seg, SegRec* seg2, int64_t n_dest,
                      const Ap& ap, int32_t* big_list, int32_t* big_count) {
    HIP_TRY(hipMemsetAsync(big_count, 0, sizeof(int32_t), s));
    const unsigned grid = grid_for(n_dest, kApplyWaves, 1 << 16);
    hipLaunchKernelGGL(k_seg_apply<Ap>, dim3(grid), dim3(kApplyWaves * RINND_WAVE),
                       kApplyWaves * kPerWave, s, off, seg, n_dest, ap, big_list, big_count);
    hipLaunchKernelGGL(k_seg_apply_big<Ap>, dim3(256), dim3(kBigThreads),
                       kBigChunk * 16 + kApplyScratch, s, off, seg, seg2, ap, big_list, big_count);
}

// ---- lane-per-destination apply --------------------------------------------------------------------
// Stage 1 sorts every destination's records by key in place (wave per destination: a register
// bitonic network up to 64 records, an LDS network up to kSortCap, chunk sort + merge beyond).
// Stage 2 lists the destinations grouped by record count (power-of-two bins), so a wave's lanes
// have similar sequence lengths.  Stage 3 gives each lane one destination: its heap lives in an LDS
// column and its sorted records are pushed in order - the same sequential semantics as the
// reference, 64 destinations per wave.
__global__ __launch_bounds__(kApplyWaves * RINND_WAVE) void k_seg_sort(const uint64_t* __restrict__ off,
                                                                      SegRec* seg, int64_t n_dest,
                                                                      int32_t* big_list, int32_t* big_count) {
    extern __shared__ __align__(16) char ssm[];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    SegRec* recs = (SegRec*)(ssm + w * kSortCap * 16);
    for (int64_t v = (int64_t)blockIdx.x * kApplyWaves + w; v < n_dest; v += (int64_t)gridDim.x * kApplyWaves) {
        const uint64_t o0 = off[v];
        const int64_t L = (int64_t)(off[v + 1] - o0);
        if (L <= 1) continue;
        if (L > kSortCap) {
            if (lane == 0) big_list[atomicAdd(big_count, 1)] = (int32_t)v;
            continue;
        }
        if (L <= RINND_WAVE) {
            uint64_t key = ~0ull;
            int32_t rx = 0;
            float rv = 0.f;
            if (lane < L) {
                const SegRec q = seg[o0 + lane];
                key = q.key;
                rx = q.x;
                rv = q.v;
            }
            const int S = next_pow2(L);
            for (int kk = 2; kk <= S; kk <<= 1)
                for (int j = kk >> 1; j > 0; j >>= 1) {
                    const uint64_t pk = __shfl_xor(key, j);
                    const int32_t px = __shfl_xor(rx, j);
                    const float pv = __shfl_xor(rv, j);
                    const bool up = (lane & kk) == 0;
                    const bool lower = (lane & j) == 0;
                    const bool gt = key > pk;
                    if (key != pk && (lower ? (gt == up) : (gt != up))) {
                        key = pk;
                        rx = px;
                        rv = pv;
                    }
                }
            if (lane < L) seg[o0 + lane] = SegRec{key, rx, rv};
            continue;
        }
        const int S = next_pow2(L);
        for (int i = lane; i < S; i += RINND_WAVE) recs[i] = i < L ? seg[o0 + i] : SegRec{~0ull, 0, 0.f};
        wave_sync();
        for (int kk = 2; kk <= S; kk <<= 1)
            for (int j = kk >> 1; j > 0; j >>= 1) {
                for (int i = lane; i < S; i += RINND_WAVE) {
                    const int ixj = i ^ j;
                    if (ixj > i) {
                        const bool up = (i & kk) == 0;
                        const SegRec A = recs[i], B = recs[ixj];
                        if (rec_gt(A, B) == up) {
                            recs[i] = B;
                            recs[ixj] = A;
                        }
                    }
                }
                wave_sync();
            }
        for (int i = lane; i < L; i += RINND_WAVE) seg[o0 + i] = recs[i];
        wave_sync();
    }
}

__global__ __launch_bounds__(kBigThreads) void k_seg_sort_big(const uint64_t* __restrict__ off, SegRec* seg,
                                                              SegRec* seg2, const int32_t* big_list,
                                                              const int32_t* big_count) {
    extern __shared__ __align__(16) char bsm[];
    SegRec* buf = (SegRec*)bsm;
    const int nb = *big_count;
    const int tid = threadIdx.x;
    for (int b = blockIdx.x; b < nb; b += gridDim.x) {
        const int64_t v = big_list[b];
        const uint64_t o0 = off[v];
        const int64_t L = (int64_t)(off[v + 1] - o0);
        SegRec* A = seg + o0;
        SegRec* B = seg2 + o0;
        for (int64_t c0 = 0; c0 < L; c0 += kBigChunk) {
            const int64_t cl = min((int64_t)kBigChunk, L - c0);
            const int S = next_pow2(cl);
            for (int i = tid; i < S; i += blockDim.x) buf[i] = i < cl ? A[c0 + i] : SegRec{~0ull, 0, 0.f};
            __syncthreads();
            for (int kk = 2; kk <= S; kk <<= 1)
                for (int j = kk >> 1; j > 0; j >>= 1) {
                    for (int i = tid; i < S; i += blockDim.x) {
                        const int ixj = i ^ j;
                        if (ixj > i) {
                            const bool up = (i & kk) == 0;
                            const SegRec x = buf[i], y = buf[ixj];
                            if (rec_gt(x, y) == up) {
                                buf[i] = y;
                                buf[ixj] = x;
                            }
                        }
                    }
                    __syncthreads();
                }
            for (int i = tid; i < cl; i += blockDim.x) A[c0 + i] = buf[i];
            __syncthreads();
        }
        SegRec* src = A;
        SegRec* dst = B;
        for (int64_t w = kBigChunk; w < L; w *= 2) {
            for (int64_t o = tid; o < L; o += blockDim.x) {  // merge path of runs [base, base+w) and [base+w, ..)
                const int64_t base = (o / (2 * w)) * (2 * w);
                const int64_t la = min(w, L - base);
                const int64_t lb = max((int64_t)0, min(w, L - base - w));
                const int64_t ol = o - base;
                const SegRec* Aa = src + base;
                const SegRec* Bb = src + base + w;
                int64_t lo = max((int64_t)0, ol - lb), hi = min(ol, la);
                while (lo < hi) {
                    const int64_t mid = (lo + hi) >> 1;
                    if (Aa[mid].key < Bb[ol - 1 - mid].key) lo = mid + 1;
                    else hi = mid;
                }
                const int64_t i = lo, j = ol - lo;
                const bool takeA = i < la && (j >= lb || Aa[i].key < Bb[j].key);
                dst[o] = takeA ? Aa[i] : Bb[j];
            }
            __syncthreads();
            SegRec* t = src;
            src = dst;
            dst = t;
        }
        if (src != A)
            for (int64_t o = tid; o < L; o += blockDim.x) A[o] = src[o];
        __syncthreads();
    }
}

constexpr int kBins = 40;  // bin of L >= 1 is 1 + floor(log2 L); bin 0 (no records) is not listed
__device__ __forceinline__ int len_bin(uint64_t L) { return L ? 64 - __clzll((long long)L) : 0; }

// bins[0..kBins): counts, bins[kBins..2kBins): offsets, bins[2kBins..3kBins): cursors
__global__ __launch_bounds__(256) void k_bin_count(const uint64_t* __restrict__ off, int64_t n_dest, uint32_t* bins) {
    __shared__ uint32_t h[kBins];
    for (int i = threadIdx.x; i < kBins; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n_dest; v += (int64_t)gridDim.x * blockDim.x) {
        const int b = len_bin(off[v + 1] - off[v]);
        if (b) atomicAdd(&h[b], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kBins; i += blockDim.x)
        if (h[i]) atomicAdd(&bins[i], h[i]);
}
__global__ void k_bin_scan(uint32_t* bins) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        uint32_t acc = 0;
        for (int i = 0; i < kBins; ++i) {
            bins[kBins + i] = acc;
            bins[2 * kBins + i] = 0;
            acc += bins[i];
        }
    }
}
__global__ __launch_bounds__(256) void k_bin_scatter(const uint64_t* __restrict__ off, int64_t n_dest, uint32_t* bins,
                                                     int32_t* order) {
    __shared__ uint32_t h[kBins], base[kBins];
    for (int64_t v0 = (int64_t)blockIdx.x * blockDim.x; v0 < n_dest; v0 += (int64_t)gridDim.x * blockDim.x) {
        for (int i = threadIdx.x; i < kBins; i += blockDim.x) h[i] = 0;
        __syncthreads();
        const int64_t v = v0 + threadIdx.x;
        int b = 0;
        uint32_t r = 0;
        if (v < n_dest) {
            b = len_bin(off[v + 1] - off[v]);
            if (b) r = atomicAdd(&h[b], 1u);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < kBins; i += blockDim.x)
            base[i] = h[i] ? bins[kBins + i] + atomicAdd(&bins[2 * kBins + i], h[i]) : 0;
        __syncthreads();
        if (b) order[base[b] + r] = (int32_t)v;
        __syncthreads();
    }
}

// checked_flagged_push (neighbor_heap.rs:149-177) with flag = new on an LDS heap column
// (slot j of lane c at j * B + c); returns true on insertion.
__device__ __forceinline__ bool heap_push_lane(float* hd, int32_t* hi, uint8_t* hf, int B, int c, int k, int32_t x,
                                               float d) {
    if (d >= hd[c]) return false;
    bool dup = false;  // no early exit: the slot loads issue together
#pragma unroll 8
    for (int j = 0; j < k; ++j) dup |= hi[j * B + c] == x;
    if (dup) return false;
    int pos = 0;  // replace the root, sift down (neighbor_heap.rs:207-243); both children's slots load together
    for (;;) {
        const int l = 2 * pos + 1, r = l + 1;
        if (l >= k) break;
        const float dl = hd[l * B + c];
        const int32_t il = hi[l * B + c];
        const uint8_t fl = hf[l * B + c];
        float dr = dl;
        int32_t ir = il;
        uint8_t fr = fl;
        if (r < k) {
            dr = hd[r * B + c];
            ir = hi[r * B + c];
            fr = hf[r * B + c];
        }
        const bool right = r < k && dl < dr;  // the right child only if strictly larger
        const float dc = right ? dr : dl;
        if (d >= dc) break;
        hd[pos * B + c] = dc;
        hi[pos * B + c] = right ? ir : il;
        hf[pos * B + c] = right ? fr : fl;
        pos = right ? r : l;
    }
    hd[pos * B + c] = d;
    hi[pos * B + c] = x;
    hf[pos * B + c] = 1;
    return true;
}

__global__ __launch_bounds__(256) void k_heap_push_lanes(const int32_t* __restrict__ order, const uint32_t* bins,
                                                         const uint64_t* __restrict__ off,
                                                         const SegRec* __restrict__ seg, int32_t* idx, float* dist,
                                                         uint8_t* flags, int k, unsigned long long* changes) {
    extern __shared__ __align__(16) float hsm[];
    const int B = blockDim.x, c = threadIdx.x;
    float* hd = hsm;
    int32_t* hi = (int32_t*)(hd + B * k);
    uint8_t* hf = (uint8_t*)(hi + B * k);
    const int64_t total = (int64_t)bins[kBins + kBins - 1] + bins[kBins - 1];
    unsigned long long ch = 0;
    for (int64_t e0 = (int64_t)blockIdx.x * B; e0 < total; e0 += (int64_t)gridDim.x * B) {
        const int64_t e = total - 1 - (e0 + c);  // longest bins first: they start early instead of trailing
        if (e >= 0) {
            const int64_t v = order[e];
            const uint64_t o0 = off[v];
            const int64_t L = (int64_t)(off[v + 1] - o0);
            const int64_t hb = v * k;
            for (int j = 0; j < k; ++j) {
                hd[j * B + c] = dist[hb + j];
                hi[j * B + c] = idx[hb + j];
                hf[j * B + c] = flags[hb + j];
            }
            const SegRec* R = seg + o0;
            float root = hd[c];
            int64_t r = 0;
            for (; r + 8 <= L; r += 8) {  // eight records in flight per batch
                SegRec q[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) q[u] = R[r + u];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (q[u].v < root && heap_push_lane(hd, hi, hf, B, c, k, q[u].x, q[u].v)) {
                        ++ch;
                        root = hd[c];
                    }
            }
            for (; r < L; ++r) {
                const SegRec q = R[r];
                if (q.v < root && heap_push_lane(hd, hi, hf, B, c, k, q.x, q.v)) {
                    ++ch;
                    root = hd[c];
                }
            }
            for (int j = 0; j < k; ++j) {
                dist[hb + j] = hd[j * B + c];
                idx[hb + j] = hi[j * B + c];
                flags[hb + j] = hf[j * B + c];
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) ch += __shfl_xor(ch, o);
    if ((threadIdx.x & 63) == 0 && ch) atomicAdd(changes, ch);
}

// checked_heap_push_flat (candidates.rs:433-484) on an LDS column: swap sift toward the larger child.  The
// caller keeps the root priority rt in a register and calls this only when priority < rt (the heap's first
// test); rt is updated to the new root.  The duplicate scan runs only when dupchk is set (the caller knows
// when index cannot be in the heap).  Returns whether index was inserted.
__device__ __forceinline__ bool cand_push_lane(float* pri, int32_t* ix, int B, int c, int n, float priority,
                                               int32_t index, float& rt, bool dupchk) {
    if (dupchk) {
        bool dup = false;  // no early exit: the slot loads issue together
#pragma unroll 8
        for (int i = 0; i < n; ++i) dup |= ix[i * B + c] == index;
        if (dup) return false;
    }
    int pos = 0;
    float top = priority;  // the root after the sift: the larger of the value and the root's children
    for (;;) {  // both children's slots load together
        const int l = 2 * pos + 1, r = 2 * pos + 2;
        if (l >= n) break;
        const float pl = pri[l * B + c];
        const int32_t il = ix[l * B + c];
        float pr = pl;
        int32_t ir = il;
        if (r < n) {
            pr = pri[r * B + c];
            ir = ix[r * B + c];
        }
        int best = pos;
        float bp = priority;  // the sifted value occupies pos
        int32_t bi = 0;
        if (pl > bp) {
            best = l;
            bp = pl;
            bi = il;
        }
        if (r < n && pr > bp) {
            best = r;
            bp = pr;
            bi = ir;
        }
        if (pos == 0) top = bp;
        if (best == pos) break;
        pri[pos * B + c] = bp;
        ix[pos * B + c] = bi;
        pos = best;
    }
    pri[pos * B + c] = priority;
    ix[pos * B + c] = index;
    rt = top;
    return true;
}


static void sort_and_bin(hipStream_t s, const uint64_t* off, SegRec* seg, SegRec* seg2, int64_t n_dest,
                         int32_t* big_list, int32_t* big_count, uint32_t* bins) {
    HIP_TRY(hipMemsetAsync(big_count, 0, sizeof(int32_t), s));
    HIP_TRY(hipMemsetAsync(bins, 0, 3 * kBins * sizeof(uint32_t), s));
    hipLaunchKernelGGL(k_seg_sort, dim3(grid_for(n_dest, kApplyWaves, 1 << 16)), dim3(kApplyWaves * RINND_WAVE),
                       kApplyWaves * kSortCap * 16, s, off, seg, n_dest, big_list, big_count);
    hipLaunchKernelGGL(k_seg_sort_big, dim3(256), dim3(kBigThreads), kBigChunk * 16, s, off, seg, seg2, big_list,
                       big_count);
    // big_list is free again: it receives the destinations grouped by bin
    hipLaunchKernelGGL(k_bin_count, dim3(grid_for(n_dest, 256, 2048)), dim3(256), 0, s, off, n_dest, bins);
    hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(64), 0, s, bins);
    hipLaunchKernelGGL(k_bin_scatter, dim3(grid_for(n_dest, 256, 4096)), dim3(256), 0, s, off, n_dest, bins,
                       big_list);
}

// Lanes per workgroup (a multiple of 64, at most cap) for kernels that keep one heap column of `slots`
// entries per lane in LDS with a column stride of B + pad; 0 if 64 lanes do not fit in 64 KiB.
static int lanes_block(int slots, int bytes_per_slot, int cap, int pad = 0) {
    auto bytes = [&](int B) { return (int64_t)(B + pad) * slots * bytes_per_slot; };
    int B = cap;
    while (B > 64 && bytes(B) > 64 * 1024) B -= 64;
    return bytes(B) <= 64 * 1024 ? B : 0;
}

void launch_heap_apply(hipStream_t s, const uint64_t* off, SegRec* seg, SegRec* seg2,
                       int64_t n_dest, int32_t* idx, float* dist, uint8_t* flags, int k,
                       unsigned long long* changes, int32_t* big_list, int32_t* big_count, uint32_t* bins) {
    const int B = lanes_block(k, 9, 256);
    if (B == 0) {  // very large k: wave per destination
        run_apply(s, off, seg, seg2, n_dest, HeapApply{idx, dist, flags, k, changes}, big_list, big_count);
        return;
    }
    sort_and_bin(s, off, seg, seg2, n_dest, big_list, big_count, bins);
    const size_t lds = (size_t)round_up((int64_t)B * k * 2 + (B * k + 3) / 4, 4) * 4;
    hipLaunchKernelGGL(k_heap_push_lanes, dim3(grid_for(n_dest, B, 8192)), dim3(B), lds, s, big_list, bins, off, seg,
                       idx, dist, flags, k, changes);
}

// ---- candidates -----------------------------------------------------------------------------------
// row_on (optional): rows whose byte is 0 (indexes of a batch that have finished) count no edges
__global__ void k_row_valid(const int32_t* __restrict__ idx, int64_t n, int k, const uint8_t* __restrict__ row_on,
                            unsigned long long* cnt) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= n;
         i += (int64_t)gridDim.x * blockDim.x) {
        unsigned long long c = 0;
        if (i < n && (!row_on || row_on[i]))
            for (int j = 0; j < k; ++j) c += idx[i * k + j] >= 0;
        cnt[i] = c;
    }
}
// out[t] = in[min(t * stride, n)] for t < count
__global__ void k_gather_strided(const uint64_t* __restrict__ in, int64_t stride, int64_t n, int count, uint64_t* out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) out[t] = in[min((int64_t)t * stride, n)];
}
void launch_gather_strided(hipStream_t s, const uint64_t* in, int64_t stride, int64_t n, int count, uint64_t* out) {
    hipLaunchKernelGGL(k_gather_strided, dim3(grid_for(count, 64)), dim3(64), 0, s, in, stride, n, count, out);
}
void launch_row_valid_counts(hipStream_t s, const int32_t* idx, int64_t n, int k, uint64_t* cnt,
                             const uint8_t* row_on) {
    hipLaunchKernelGGL(k_row_valid, dim3(grid_for(n + 1, 256)), dim3(256), 0, s, idx, n, k, row_on,
                       (unsigned long long*)cnt);
}

// FastRng streams at arbitrary offsets: thread = chunk of kRngChunk draws, jumped to its start.
constexpr int kRngChunk = 2048;
__global__ void k_rng_streams(const uint64_t* __restrict__ tables, const RngStream* streams,
                              float* out) {
    const RngStream st = streams[blockIdx.y];
    const int64_t chunk = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t first = chunk * kRngChunk;
    if (first >= st.count) return;
    uint64_t s4[4] = {st.s[0], st.s[1], st.s[2], st.s[3]};
    xoshiro_jump(tables, s4, (uint64_t)first);
    Xoshiro r;
    for (int w = 0; w < 4; ++w) r.s[w] = s4[w];
    const int64_t last = min(first + kRngChunk, st.count);
    float* o = out + st.out_offset;
    for (int64_t t = first; t < last; ++t) o[t] = r.next_float();
}
void launch_rng_streams(hipStream_t s, const uint64_t* jump_tables, const RngStream* streams_dev,
                        int n_streams, int64_t max_count, float* out) {
    if (n_streams == 0 || max_count == 0) return;
    const int64_t chunks = ceil_div(max_count, kRngChunk);
    hipLaunchKernelGGL(k_rng_streams, dim3(grid_for(chunks, 64, 1 << 30), n_streams), dim3(64), 0, s,
                       jump_tables, streams_dev, out);
}


// ---- candidate build with a stable destination sort -----------------------------------------------
// A vertex v receives, in the reference's order, its own heap row's edges (forward pushes, keys
// v*k + j) and the reverse pushes of every edge i -> v (keys i*k + j'; T > 1: own row block first,
// then the other blocks ascending).  Forward pushes need no grouping: the push kernel reads v's row.
// Reverse pushes are written in edge order as (destination, source|flag, priority) and sorted by
// destination with a stable radix sort, so each destination's segment is ordered by source row -
// exactly the key order.  The push kernel merges the two streams.
__global__ void k_cand_gen_rev(const int32_t* __restrict__ idx, const uint8_t* __restrict__ flags, int64_t n, int k,
                               const uint64_t* __restrict__ rvp, const float* __restrict__ prio, int mode,
                               int64_t block_rows, const int64_t* __restrict__ block_prio_base, uint64_t counter_key,
                               const uint8_t* __restrict__ row_on, float* fpr, uint32_t* rkey, uint64_t* rval,
                               unsigned long long* cnt) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t r = rvp[i];
        const int64_t blk = mode == 1 ? i / block_rows : 0;
        const uint64_t blk_first = mode == 1 ? rvp[blk * block_rows] : 0;
        const bool on = !row_on || row_on[i];  // a finished index's rows send no reverse edges
        for (int j = 0; j < k; ++j) {
            const uint64_t e = (uint64_t)i * k + j;
            const int32_t nb = on ? idx[e] : -1;
            if (nb < 0) {
                rkey[e] = (uint32_t)n;  // sorts after every destination
                continue;
            }
            const uint32_t fbit = flags[e] ? 0x80000000u : 0u;
            float pf, pr;
            if (mode == 0) {  // candidates.rs:84-131: two draws per edge from the main stream
                pf = prio[2 * r];
                pr = prio[2 * r + 1];
            } else if (mode == 1) {  // candidates.rs:170-213: one draw per edge from the row block's stream
                pf = pr = prio[block_prio_base[blk] + (int64_t)(r - blk_first)];
            } else {  // counter draws
                const uint64_t h1 = mix64(counter_key ^ (2 * e));
                const uint64_t h2 = mix64(counter_key ^ (2 * e + 1));
                pf = (float)(uint32_t)(h1 >> 40) * (1.0f / 16777216.0f);
                pr = (float)(uint32_t)(h2 >> 40) * (1.0f / 16777216.0f);
            }
            fpr[e] = pf;
            rkey[e] = (uint32_t)nb;
            rval[e] = ((uint64_t)__builtin_bit_cast(uint32_t, pr) << 32) | ((uint32_t)i | fbit);
            atomicAdd(&cnt[nb], 1ull);
            ++r;
        }
    }
}

// The same for counter draws, one thread per heap entry (coalesced): the draws depend on the entry index only,
// and the destinations' segments are found in the sorted keys afterwards (no counting atomics).
__global__ __launch_bounds__(256) void k_cand_gen_rev_ctr(const int32_t* __restrict__ idx,
                                                          const uint8_t* __restrict__ flags, int64_t n, int k,
                                                          uint64_t counter_key, float* fpr, uint32_t* rkey,
                                                          uint64_t* rval) {
    const int64_t total = n * k;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int32_t nb = idx[e];
        if (nb < 0) {
            rkey[e] = (uint32_t)n;  // sorts after every destination
            continue;
        }
        const int64_t i = e / k;
        const uint32_t fbit = flags[e] ? 0x80000000u : 0u;
        const uint64_t h1 = mix64(counter_key ^ (2 * (uint64_t)e));
        const uint64_t h2 = mix64(counter_key ^ (2 * (uint64_t)e + 1));
        fpr[e] = (float)(uint32_t)(h1 >> 40) * (1.0f / 16777216.0f);
        rkey[e] = (uint32_t)nb;
        rval[e] = ((uint64_t)__builtin_bit_cast(uint32_t, (float)(uint32_t)(h2 >> 40) * (1.0f / 16777216.0f)) << 32) |
                  ((uint32_t)i | fbit);
    }
}

// Segment starts of keys sorted ascending (keys in [0, n]): off[v] = the first position whose key is >= v,
// for v in [0, n] (absent keys get the next present key's start; off[n] counts the keys below n).
__global__ __launch_bounds__(256) void k_seg_starts(const uint32_t* __restrict__ keys, int64_t N, int64_t n,
                                                    uint64_t* off) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p <= N; p += (int64_t)gridDim.x * blockDim.x) {
        const int64_t kprev = p > 0 ? (int64_t)keys[p - 1] : -1;
        const int64_t kcur = p < N ? (int64_t)keys[p] : n + 1;
        for (int64_t v = kprev + 1; v <= kcur && v <= n; ++v) off[v] = (uint64_t)p;
    }
}

// first position in [lo, hi) whose record's source row is >= bound (records sorted by source)
__device__ __forceinline__ uint64_t src_lower_bound(const uint64_t* rv, uint64_t lo, uint64_t hi, int64_t bound) {
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if ((int64_t)(uint32_t)(rv[mid] & 0x7fffffffu) < bound) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// FIDS (k <= 16 and ids below 2^26): the row's own ids are kept in registers.  A destination's reverse
// sources are distinct and differ from v, and v's row holds distinct ids, so an id can reach the same heap
// twice only if it is both a reverse source and in v's row: only then does a push scan the heap for a
// duplicate (pushed_* records the row slots whose id entered that heap).  A new-heap entry whose id is
// row slot j carries j + 1 in bits 26..30, which flags slot j old at the end (mark_old_flags) without
// searching the heap for every slot.
template <bool FIDS>
__global__ __launch_bounds__(128) void k_cand_push_rev(const int32_t* __restrict__ order, const uint32_t* bins,
                                                       const uint64_t* __restrict__ off, const uint64_t* __restrict__ rv,
                                                       const int32_t* __restrict__ idx, uint8_t* flags,
                                                       const float* __restrict__ fpr, int k, int mode, int64_t block_rows,
                                                       int32_t* new_idx, int32_t* old_idx, int C,
                                                       unsigned long long* row_pairs, unsigned long long* counts) {
    extern __shared__ __align__(16) float csm[];
    const int B = blockDim.x, c = threadIdx.x;
    const int S = B + 1;  // column stride (odd: the row stores are conflict-free)
    float* np = csm;
    int32_t* ni = (int32_t*)(np + S * C);
    float* op = (float*)(ni + S * C);
    int32_t* oi = (int32_t*)(op + S * C);
    const int64_t total = (int64_t)bins[kBins + kBins - 1] + bins[kBins - 1];
    unsigned long long st_rows = 0, st_pairs = 0;
    for (int64_t e0 = (int64_t)blockIdx.x * B; e0 < total; e0 += (int64_t)gridDim.x * B) {
        const int64_t e = total - 1 - (e0 + c);  // longest bins first: they start early instead of trailing
        const int32_t v = e >= 0 ? order[e] : -1;
        if (e >= 0) {
            const uint64_t o0 = off[v], o1 = off[v + 1];
            for (int j = 0; j < C; ++j) {
                np[j * S + c] = __builtin_inff();
                op[j * S + c] = __builtin_inff();
                ni[j * S + c] = -1;
                oi[j * S + c] = -1;
            }
            // One push path for both heaps (a lane picks its heap by an LDS offset), so a wave whose lanes
            // push into different heaps runs the sift once, not once per heap; both roots stay in registers.
            float rt_new = __builtin_inff(), rt_old = __builtin_inff();
            int32_t fid[16];
            uint32_t pushed_new = 0, pushed_old = 0;
            if (FIDS) {
#pragma unroll
                for (int j = 0; j < 16; ++j) fid[j] = j < k ? idx[(int64_t)v * k + j] : -1;
            }
            auto push = [&](uint32_t x, float pri) {
                const bool isnew = (int32_t)x < 0;  // "new" edge
                const float rn = rt_new, ro = rt_old;
                float rt = isnew ? rn : ro;
                if (pri >= rt) return;
                const int h = isnew ? 0 : 2 * S * C;
                int32_t id = (int32_t)(x & 0x7fffffffu);
                bool dupchk = true;
                int m = 0;  // 1 + the row slot holding id, or 0
                if (FIDS) {
#pragma unroll
                    for (int j = 0; j < 16; ++j) m = fid[j] == id ? j + 1 : m;
                    const uint32_t pn = pushed_new, po = pushed_old;  // values: no select of addresses
                    dupchk = m && (((isnew ? pn : po) >> (m - 1)) & 1u);
                    if (isnew) id |= m << 26;
                }
                const bool ins = cand_push_lane(np + h, ni + h, S, c, C, pri, id, rt, dupchk);
                if (FIDS) {  // selects, not branches: the pairs stay in registers
                    const uint32_t bit = ins && m ? 1u << (m - 1) : 0u;
                    pushed_new |= isnew ? bit : 0u;
                    pushed_old |= isnew ? 0u : bit;
                }
                rt_new = isnew ? rt : rt_new;
                rt_old = isnew ? rt_old : rt;
            };
            auto rev = [&](uint64_t a, uint64_t b) {  // a ring of four records in flight ahead of the push
                if (a >= b) return;
                uint64_t q0 = rv[a], q1 = rv[min(a + 1, b - 1)], q2 = rv[min(a + 2, b - 1)], q3 = rv[min(a + 3, b - 1)];
                for (uint64_t r = a; r < b; ++r) {
                    const uint64_t w = q0;
                    q0 = q1;
                    q1 = q2;
                    q2 = q3;
                    q3 = rv[min(r + 4, b - 1)];
                    push((uint32_t)w, __builtin_bit_cast(float, (uint32_t)(w >> 32)));
                }
            };
            auto fwd = [&]() {  // the row's own edges, with a ring of four entries in flight ahead of the push
                const int64_t f0 = (int64_t)v * k;
                int32_t i0 = idx[f0], i1 = idx[f0 + min(1, k - 1)], i2 = idx[f0 + min(2, k - 1)],
                        i3 = idx[f0 + min(3, k - 1)];
                uint8_t g0 = flags[f0], g1 = flags[f0 + min(1, k - 1)], g2 = flags[f0 + min(2, k - 1)],
                        g3 = flags[f0 + min(3, k - 1)];
                float p0 = fpr[f0], p1 = fpr[f0 + min(1, k - 1)], p2 = fpr[f0 + min(2, k - 1)],
                      p3 = fpr[f0 + min(3, k - 1)];
                for (int j = 0; j < k; ++j) {
                    const int32_t nb = i0;
                    const uint8_t fl = g0;
                    const float pr = p0;
                    i0 = i1, i1 = i2, i2 = i3;
                    g0 = g1, g1 = g2, g2 = g3;
                    p0 = p1, p1 = p2, p2 = p3;
                    const int64_t fn = f0 + min(j + 4, k - 1);
                    i3 = idx[fn];
                    g3 = flags[fn];
                    p3 = fpr[fn];
                    if (nb >= 0) push((uint32_t)nb | (fl ? 0x80000000u : 0u), pr);
                }
            };
            if (mode != 1) {  // key order = source row order; v's own row sits between sources < v and > v
                const uint64_t pv = src_lower_bound(rv, o0, o1, v + 1);
                rev(o0, pv);
                fwd();
                rev(pv, o1);
            } else {  // own row block first (merged with v's row), then the other blocks ascending
                const int64_t lo = (v / block_rows) * block_rows, hi = lo + block_rows;
                const uint64_t pa = src_lower_bound(rv, o0, o1, lo);
                const uint64_t pb = src_lower_bound(rv, pa, o1, hi);
                const uint64_t pv = src_lower_bound(rv, pa, pb, v + 1);
                rev(pa, pv);
                fwd();
                rev(pv, pb);
                rev(o0, pa);
                rev(pb, o1);
            }
            unsigned long long nn = 0, no = 0;
            for (int j = 0; j < C; ++j) {
                const int32_t e = ni[j * S + c];
                nn += e >= 0;
                no += oi[j * S + c] >= 0;
                if (FIDS && e >= (1 << 26)) {  // row slot (e >> 26) - 1 became a new candidate: flag it old
                    flags[(int64_t)v * k + (e >> 26) - 1] = 0;
                    ni[j * S + c] = e & ((1 << 26) - 1);
                }
            }
            const unsigned long long pr = nn * (nn - (nn > 0)) / 2 + nn * no;
            row_pairs[v] = pr;
            st_rows += nn + no;
            st_pairs += pr;
            // mark_old_flags (candidates.rs:354-417): the row's entries that became new candidates are
            // flagged old (FIDS: from the tags, in the loop above).  Only this lane reads or writes row v's
            // flags in this kernel.
            const int64_t f0 = (int64_t)v * k;
            for (int j = 0; j < (FIDS ? 0 : k); ++j) {
                const int32_t nb = idx[f0 + j];
                if (nb < 0 || !flags[f0 + j]) continue;
                for (int cc = 0; cc < C; ++cc)
                    if (ni[cc * S + c] == nb) {
                        flags[f0 + j] = 0;
                        break;
                    }
            }
        }
        wave_rows<false>(v, C, new_idx, ni, S, c & ~63);
        wave_rows<false>(v, C, old_idx, oi, S, c & ~63);
    }
    for (int o = 32; o > 0; o >>= 1) {
        st_rows += __shfl_xor(st_rows, o);
        st_pairs += __shfl_xor(st_pairs, o);
    }
    // one pair of atomics per workgroup: atomics on one address serialise in the L2
    __shared__ unsigned long long wst[2][16];
    if ((threadIdx.x & 63) == 0) {
        wst[0][threadIdx.x >> 6] = st_rows;
        wst[1][threadIdx.x >> 6] = st_pairs;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) {
            st_rows += wst[0][i];
            st_pairs += wst[1][i];
        }
        if (st_rows) {
            atomicAdd(&counts[0], st_rows);
            atomicAdd(&counts[1], st_pairs);
        }
    }
}

// destination lengths for the bins: reverse records plus the row's valid forward edges
__global__ __launch_bounds__(256) void k_bin_count2(const uint64_t* __restrict__ off, const uint64_t* __restrict__ off2,
                                                    int64_t n_dest, uint32_t* bins) {
    __shared__ uint32_t h[kBins];
    for (int i = threadIdx.x; i < kBins; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n_dest; v += (int64_t)gridDim.x * blockDim.x) {
        const int b = len_bin((off[v + 1] - off[v]) + (off2[v + 1] - off2[v]));
        if (b) atomicAdd(&h[b], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kBins; i += blockDim.x)
        if (h[i]) atomicAdd(&bins[i], h[i]);
}
__global__ __launch_bounds__(256) void k_bin_scatter2(const uint64_t* __restrict__ off, const uint64_t* __restrict__ off2,
                                                      int64_t n_dest, uint32_t* bins, int32_t* order) {
    __shared__ uint32_t h[kBins], base[kBins];
    for (int64_t v0 = (int64_t)blockIdx.x * blockDim.x; v0 < n_dest; v0 += (int64_t)gridDim.x * blockDim.x) {
        for (int i = threadIdx.x; i < kBins; i += blockDim.x) h[i] = 0;
        __syncthreads();
        const int64_t v = v0 + threadIdx.x;
        int b = 0;
        uint32_t r = 0;
        if (v < n_dest) {
            b = len_bin((off[v + 1] - off[v]) + (off2[v + 1] - off2[v]));
            if (b) r = atomicAdd(&h[b], 1u);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < kBins; i += blockDim.x)
            base[i] = h[i] ? bins[kBins + i] + atomicAdd(&bins[2 * kBins + i], h[i]) : 0;
        __syncthreads();
        if (b) order[base[b] + r] = (int32_t)v;
        __syncthreads();
    }
}

// Stable sort of (u32 destination, u64 record) pairs on the key's low end_bit bits.  rocprim's gfx950
// onesweep default for this pair type ranks 8-bit digits, three scatter passes for the 17..24-bit
// destinations of a 65k..16M-point graph; here the digit is widened to at most 11 bits so that keys of up
// to 22 bits take two passes (each pass reads and writes 12 B per record).  Block shape as the default.
template <unsigned BITS>
using PairSortConfig = rocprim::radix_sort_config<
    rocprim::default_config, rocprim::default_config,
    rocprim::radix_sort_onesweep_config<rocprim::kernel_config<1024, 8>, rocprim::kernel_config<1024, 8>, BITS,
                                        rocprim::block_radix_rank_algorithm::match>>;

template <unsigned BITS>
static hipError_t sort_pairs_bits(void* tmp, size_t& bytes, const uint32_t* kin, uint32_t* kout,
                                  const uint64_t* vin, uint64_t* vout, int64_t items, int end_bit, hipStream_t s) {
    return rocprim::radix_sort_pairs<PairSortConfig<BITS>>(tmp, bytes, kin, kout, vin, vout, (size_t)items, 0u,
                                                           (unsigned)end_bit, s);
}

static hipError_t sort_pairs_u32_u64(void* tmp, size_t& bytes, const uint32_t* kin, uint32_t* kout,
                                     const uint64_t* vin, uint64_t* vout, int64_t items, int end_bit,
                                     hipStream_t s) {
    const int passes = (end_bit + 10) / 11;
    const int bits = passes > 0 ? (end_bit + passes - 1) / passes : 8;  // <= 11
    switch (bits) {
        case 9: return sort_pairs_bits<9>(tmp, bytes, kin, kout, vin, vout, items, end_bit, s);
        case 10: return sort_pairs_bits<10>(tmp, bytes, kin, kout, vin, vout, items, end_bit, s);
        case 11: return sort_pairs_bits<11>(tmp, bytes, kin, kout, vin, vout, items, end_bit, s);
        default: return sort_pairs_bits<8>(tmp, bytes, kin, kout, vin, vout, items, end_bit, s);
    }
}

size_t cand_sort_tmp_bytes(int64_t items, int end_bit) {
    size_t b = 0;
    HIP_TRY(sort_pairs_u32_u64(nullptr, b, nullptr, nullptr, nullptr, nullptr, items, end_bit, nullptr));
    return b;
}

void launch_cand_build(hipStream_t s, const CandBuildArgs& a) {
    const int64_t nk = a.n * a.k;
    if (a.mode == 2) {
        hipLaunchKernelGGL(k_cand_gen_rev_ctr, dim3(grid_for(nk, 256, 1 << 16)), dim3(256), 0, s, a.idx, a.flags, a.n,
                           a.k, a.counter_key, a.fpr, a.rkey, a.rval);
    } else {
        launch_zero(s, a.cnt, (a.n + 1) * sizeof(unsigned long long));
        hipLaunchKernelGGL(k_cand_gen_rev, dim3(grid_for(a.n, 128)), dim3(128), 0, s, a.idx, a.flags, a.n, a.k, a.rvp,
                           a.prio, a.mode, a.block_rows, a.block_prio_base, a.counter_key, a.row_on, a.fpr, a.rkey,
                           a.rval, a.cnt);
    }
    int end_bit = 1;
    while (((int64_t)1 << end_bit) <= a.n) ++end_bit;  // destinations 0..n (n = no edge)
    size_t tb = a.sort_tmp_bytes;
    HIP_TRY(sort_pairs_u32_u64(a.sort_tmp, tb, a.rkey, a.rkey2, a.rval, a.rval2, nk, end_bit, s));
    if (a.mode == 2)
        hipLaunchKernelGGL(k_seg_starts, dim3(grid_for(nk + 1, 256)), dim3(256), 0, s, a.rkey2, nk, a.n, a.off);
    else
        scan_u64_exclusive(s, a.cnt, a.off, a.n + 1, a.scan_tmp, a.scan_tmp_bytes);
    launch_zero(s, a.bins, 3 * kBins * sizeof(uint32_t));
    hipLaunchKernelGGL(k_bin_count2, dim3(grid_for(a.n, 256, 2048)), dim3(256), 0, s, a.off, a.rvp, a.n, a.bins);
    hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(64), 0, s, a.bins);
    hipLaunchKernelGGL(k_bin_scatter2, dim3(grid_for(a.n, 256, 4096)), dim3(256), 0, s, a.off, a.rvp, a.n, a.bins,
                       a.order);
    const int B = lanes_block(a.C, 16, 128, 1);  // C <= 60: 64 lanes always fit
    if (a.k <= 16 && a.n < ((int64_t)1 << 26))
        hipLaunchKernelGGL(k_cand_push_rev<true>, dim3(grid_for(a.n, B, 8192)), dim3(B), (size_t)(B + 1) * a.C * 16, s,
                           a.order, a.bins, a.off, a.rval2, a.idx, a.flags, a.fpr, a.k, a.mode, a.block_rows, a.new_idx,
                           a.old_idx, a.C, a.row_pairs, a.counts);
    else
        hipLaunchKernelGGL(k_cand_push_rev<false>, dim3(grid_for(a.n, B, 8192)), dim3(B), (size_t)(B + 1) * a.C * 16, s,
                           a.order, a.bins, a.off, a.rval2, a.idx, a.flags, a.fpr, a.k, a.mode, a.block_rows, a.new_idx,
                           a.old_idx, a.C, a.row_pairs, a.counts);
}


// ---- local join, FP32 (update_iteration generation, mod.rs:383-447) --------------------------------
// A workgroup takes RPG rows at a time with TPR threads per row (TPR >= 64: a wave never spans two rows).
// Row i's new (C) and old (C) candidate rows are staged in LDS; every pair (j, kk) with kk in (j+1..C) over
// new and C+(0..C) over old is evaluated with the reference's arithmetic and kept if d is below the root
// of p or of q (the flags say which).  Kept pairs go to the row's own region of U (no global counter);
// the row's record count goes to krec.  The tile shape changes only the schedule, never a result.
constexpr int kJoinLdsBytes = 64 * 1024;

template <bool STAGED, int TPR, int RPG>
__global__ __launch_bounds__(TPR * RPG) void k_join_fp32(
    int metric, const float* __restrict__ X, int64_t ld, int dim,
    const int32_t* __restrict__ new_idx, const int32_t* __restrict__ old_idx, int C,
    int64_t row_begin, int64_t row_end, const float* __restrict__ hd, int k, JoinOut out) {
    static_assert(TPR % RINND_WAVE == 0, "a wave must not span two rows");
    extern __shared__ __align__(16) float jsm[];
    __shared__ int32_t ids[RPG][128];
    __shared__ float roots[RPG][128];
    __shared__ unsigned run[RPG], nrec[RPG];
    const int C2 = 2 * C;
    const int64_t ldl = ld + 4;  // padded LDS row stride
    const int lane = threadIdx.x & 63;
    const int g = threadIdx.x / TPR, gt = threadIdx.x - g * TPR;  // the thread's row in the group, its rank there
    const unsigned long long lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    unsigned long long kept = 0;  // threads gt == 0: kept pairs of their rows
    for (int64_t i0 = row_begin + (int64_t)blockIdx.x * RPG; i0 < row_end; i0 += (int64_t)gridDim.x * RPG) {
        for (int c = threadIdx.x; c < RPG * C2; c += blockDim.x) {
            const int gg = c / C2, cc = c - gg * C2;
            const int64_t i = i0 + gg;
            int32_t id = -1;
            if (i < row_end) id = cc < C ? new_idx[i * C + cc] : old_idx[i * C + (cc - C)];
            ids[gg][cc] = id;
            roots[gg][cc] = id >= 0 ? hd[(int64_t)id * k] : 0.f;
        }
        if (threadIdx.x < RPG) run[threadIdx.x] = nrec[threadIdx.x] = 0;
        __syncthreads();
        if (STAGED) {
            const int vec = (int)(ld >> 2);
            for (int t = threadIdx.x; t < RPG * C2 * vec; t += blockDim.x) {
                const int c = t / vec, e4 = t - c * vec;
                const int gg = c / C2;
                const int32_t id = ids[gg][c - gg * C2];
                float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
                if (id >= 0) v = *reinterpret_cast<const float4*>(X + (int64_t)id * ld + e4 * 4);
                *reinterpret_cast<float4*>(jsm + c * ldl + e4 * 4) = v;
            }
            __syncthreads();
        }
        const int64_t i = i0 + g;
        const bool row_ok = i < row_end;
        const int64_t r = i - row_begin;
        JRec* U = out.U + (row_ok ? out.rowoff[r] : 0);
        const float* rows = jsm + (int64_t)g * C2 * ldl;
        const int total = C * C2;
        for (int t0 = 0; t0 < total; t0 += TPR) {
            const int t = t0 + gt;
            bool keep = false, ep = false, eq = false;
            float d = 0.f;
            if (t < total) {
                const int j = t / C2, kk = t - j * C2;
                const int32_t p = ids[g][j], q = ids[g][kk];
                const bool valid = p >= 0 && q >= 0 && (kk < C ? kk > j : p != q);
                if (valid) {
                    const float* xp = STAGED ? rows + j * ldl : X + (int64_t)p * ld;
                    const float* xq = STAGED ? rows + kk * ldl : X + (int64_t)q * ld;
                    d = distance_exact(metric, xp, xq, dim);
                    ep = d < roots[g][j];
                    eq = d < roots[g][kk];
                    keep = ep || eq;
                }
            }
            const unsigned long long bal = __ballot(keep);
            const unsigned nr = (unsigned)(__popcll(__ballot(ep)) + __popcll(__ballot(eq)));
            if (bal) {
                const int leader = __ffsll((long long)bal) - 1;
                unsigned base = 0;
                if (lane == leader) {
                    base = atomicAdd(&run[g], (unsigned)__popcll(bal));
                    atomicAdd(&nrec[g], nr);
                }
                base = __shfl(base, leader);
                if (keep) U[base + __popcll(bal & lt)] = JRec{(uint32_t)t | (ep ? kEmitP : 0u) | (eq ? kEmitQ : 0u), d};
            }
        }
        __syncthreads();
        if (gt == 0 && row_ok) {
            out.kc[r] = (unsigned long long)run[g];
            out.krec[r] = (unsigned long long)nrec[g];
            kept += run[g];
        }
        __syncthreads();
    }
    if (gt == 0 && kept) atomicAdd(&out.kept[(blockIdx.x * RPG + g) % kKeptSlots], kept);
}

// Tile shapes (threads per row, rows per workgroup); shape 0 is the default.
struct JoinShape {
    int tpr, rpg;
};
static constexpr JoinShape kJoinShapes[] = {{256, 1}, {128, 1}, {128, 2}, {64, 2}, {64, 4}};
constexpr int kJoinShapeCount = (int)(sizeof(kJoinShapes) / sizeof(kJoinShapes[0]));

static bool join_shape_staged(int shape, int C, int64_t ld) {
    return (int64_t)kJoinShapes[shape].rpg * 2 * C * (ld + 4) * 4 <= kJoinLdsBytes;
}

template <int TPR, int RPG>
static void launch_join_fp32_shape(hipStream_t s, bool staged, int metric, const float* X, int64_t ld, int dim,
                                   const int32_t* new_idx, const int32_t* old_idx, int C, int64_t row_begin,
                                   int64_t row_end, const float* heap_dist, int k, const JoinOut& out) {
    const int64_t ldl = ld + 4;
    const size_t lds = staged ? (size_t)RPG * 2 * C * ldl * 4 : 16;
    const dim3 grid(grid_for(ceil_div(row_end - row_begin, (int64_t)RPG), 1, 1 << 18));
    if (staged)
        hipLaunchKernelGGL((k_join_fp32<true, TPR, RPG>), grid, dim3(TPR * RPG), lds, s, metric, X, ld, dim, new_idx,
                           old_idx, C, row_begin, row_end, heap_dist, k, out);
    else
        hipLaunchKernelGGL((k_join_fp32<false, TPR, RPG>), grid, dim3(TPR * RPG), lds, s, metric, X, ld, dim, new_idx,
                           old_idx, C, row_begin, row_end, heap_dist, k, out);
}

void launch_join_fp32(hipStream_t s, int metric, const float* X, int64_t ld, int dim,
                      const int32_t* new_idx, const int32_t* old_idx, int C, int64_t row_begin,
                      int64_t row_end, const float* heap_dist, int k, const JoinOut& out, int shape) {
    if (row_end <= row_begin) return;
    if (shape < 0 || shape >= kJoinShapeCount) shape = 0;
    // a multi-row shape is used only where all its rows' candidates fit in LDS
    if (shape != 0 && !join_shape_staged(shape, C, ld)) shape = 0;
    const bool staged = join_shape_staged(shape, C, ld);
    switch (shape) {
        case 1: launch_join_fp32_shape<128, 1>(s, staged, metric, X, ld, dim, new_idx, old_idx, C, row_begin, row_end, heap_dist, k, out); break;
        case 2: launch_join_fp32_shape<128, 2>(s, staged, metric, X, ld, dim, new_idx, old_idx, C, row_begin, row_end, heap_dist, k, out); break;
        case 3: launch_join_fp32_shape<64, 2>(s, staged, metric, X, ld, dim, new_idx, old_idx, C, row_begin, row_end, heap_dist, k, out); break;
        case 4: launch_join_fp32_shape<64, 4>(s, staged, metric, X, ld, dim, new_idx, old_idx, C, row_begin, row_end, heap_dist, k, out); break;
        default: launch_join_fp32_shape<256, 1>(s, staged, metric, X, ld, dim, new_idx, old_idx, C, row_begin, row_end, heap_dist, k, out); break;
    }
}

int join_fp32_shape_count() { return kJoinShapeCount; }
void join_fp32_shape_info(int shape, int* tpr, int* rpg) {
    if (shape < 0 || shape >= kJoinShapeCount) shape = 0;
    *tpr = kJoinShapes[shape].tpr;
    *rpg = kJoinShapes[shape].rpg;
}

// The join over rows [0, n) with the tile-shape tuner: every eligible shape runs, timed, on two slices of the
// rows (shapes in order, then in reverse order, so drift over the launch sequence cancels), and the rest of the
// rows go to the fastest per row.  Each row is joined once, whichever shape takes it: the results are those of
// any single shape.  Returns the fastest shape.
int launch_join_fp32_tuned(hipStream_t s, int metric, const float* X, int64_t ld, int dim, const int32_t* new_idx,
                           const int32_t* old_idx, int C, int64_t n, const float* heap_dist, int k, const JoinOut& out,
                           bool* tuned) {
    std::vector<int> shapes;
    for (int sh = 0; sh < kJoinShapeCount; ++sh)
        if (sh == 0 || join_shape_staged(sh, C, ld)) shapes.push_back(sh);
    const int V = (int)shapes.size();
    const int64_t slice = n / (4 * (int64_t)V);
    *tuned = false;
    if (V < 2 || slice < 16384) {  // too few rows to time: the default shape
        launch_join_fp32(s, metric, X, ld, dim, new_idx, old_idx, C, 0, n, heap_dist, k, out, 0);
        return 0;
    }
    *tuned = true;
    std::vector<hipEvent_t> ev(2 * V + 1);
    for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
    auto sub = [&](int64_t b) {  // the join's outputs for rows starting at b
        JoinOut o = out;
        o.rowoff += b;
        o.kc += b;
        o.krec += b;
        return o;
    };
    int64_t row = 0;
    HIP_TRY(hipEventRecord(ev[0], s));
    for (int pass = 0; pass < 2; ++pass)
        for (int v = 0; v < V; ++v) {
            const int sh = shapes[pass == 0 ? v : V - 1 - v];
            launch_join_fp32(s, metric, X, ld, dim, new_idx, old_idx, C, row, row + slice, heap_dist, k, sub(row), sh);
            row += slice;
            HIP_TRY(hipEventRecord(ev[1 + pass * V + v], s));
        }
    HIP_TRY(hipEventSynchronize(ev[2 * V]));
    std::vector<double> ms(V, 0.0);
    for (int pass = 0; pass < 2; ++pass)
        for (int v = 0; v < V; ++v) {
            float t = 0.f;
            HIP_TRY(hipEventElapsedTime(&t, ev[pass * V + v], ev[1 + pass * V + v]));
            ms[pass == 0 ? v : V - 1 - v] += t;
        }
    for (auto& e : ev) (void)hipEventDestroy(e);
    int best = 0;
    for (int v = 1; v < V; ++v)
        if (ms[v] < ms[best]) best = v;
    launch_join_fp32(s, metric, X, ld, dim, new_idx, old_idx, C, row, n, heap_dist, k, sub(row), shapes[best]);
    return shapes[best];
}


// ---- point relabelling (counter draws) --------------------------------------------------------------
__global__ void k_relabel_rows(const float* __restrict__ X, int64_t ld, int64_t n, const int32_t* __restrict__ perm,
                               int32_t* inv, float* __restrict__ Xr) {
    const int64_t vec = ld >> 2;
    const int64_t total = n * vec;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = t / vec, e = t - r * vec;
        const int32_t src = perm[r];
        *reinterpret_cast<float4*>(Xr + r * ld + e * 4) = *reinterpret_cast<const float4*>(X + (int64_t)src * ld + e * 4);
        if (e == 0) inv[src] = (int32_t)r;
    }
}
void launch_relabel_points(hipStream_t s, const float* X, int64_t ld, int64_t n, const int32_t* perm, int32_t* inv,
                           float* Xr) {
    hipLaunchKernelGGL(k_relabel_rows, dim3(grid_for(n * (ld >> 2), 256)), dim3(256), 0, s, X, ld, n, perm, inv, Xr);
}
// every tree's indices through the same map in one launch (blockIdx.y: the tree)
__global__ void k_map_ids_trees(int32_t* const* ids, int64_t count, const int32_t* __restrict__ map) {
    int32_t* p = ids[blockIdx.y];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = p[i];
        if (v >= 0) p[i] = map[v];
    }
}
void launch_map_ids_trees(hipStream_t s, int32_t* const* ids_dev, int n_trees, int64_t count, const int32_t* map) {
    if (count > 0 && n_trees > 0)
        hipLaunchKernelGGL(k_map_ids_trees, dim3(grid_for(count, 256, 2048), n_trees), dim3(256), 0, s, ids_dev, count,
                           map);
}
// row r of the relabelled heap is row perm[r] of the caller's, and its ids map through perm
__global__ void k_unpermute_heap(const int32_t* __restrict__ idx, const float* __restrict__ dist,
                                 const uint8_t* __restrict__ flags, const int32_t* __restrict__ perm, int64_t n, int k,
                                 int32_t* oidx, float* odist, uint8_t* oflags) {
    const int64_t total = n * k;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = t / k, j = t - r * k;
        const int64_t o = (int64_t)perm[r] * k + j;
        const int32_t v = idx[t];
        oidx[o] = v >= 0 ? perm[v] : v;
        odist[o] = dist[t];
        oflags[o] = flags[t];
    }
}
void launch_unpermute_heap(hipStream_t s, const int32_t* idx, const float* dist, const uint8_t* flags,
                           const int32_t* perm, int64_t n, int k, int32_t* oidx, float* odist, uint8_t* oflags) {
    hipLaunchKernelGGL(k_unpermute_heap, dim3(grid_for(n * k, 256)), dim3(256), 0, s, idx, dist, flags, perm, n, k,
                       oidx, odist, oflags);
}

// ---- join records in generation order + stable destination sort + lane apply ----------------------
// The records of each row's kept pairs in generation order (slot order; p's record before q's within a
// slot).  The row's kept pairs sit in its region in any order and their slots are unique: a group of G
// lanes per row marks the slots in two bitmasks over the C x 2C slots (records for p, records for q),
// scans the record counts of the 64-slot words, and then every kept pair writes its records at their ranks.
template <int G>
__global__ __launch_bounds__(256) void k_join_records(const int32_t* __restrict__ new_idx,
                                                      const int32_t* __restrict__ old_idx, int C, int64_t n,
                                                      JoinOut out, const uint64_t* __restrict__ koff, uint32_t* rkey,
                                                      uint64_t* rval) {
    // G lanes per row (64 / G rows per wave): a row keeps about as many pairs as it has candidates, so 16
    // lanes cover it in one pass when its slot masks fit in 16 words
    extern __shared__ __align__(16) char jrm[];
    const int lane = threadIdx.x & 63, gl = lane & (G - 1);
    const int grp = threadIdx.x / G;  // group within the block
    const int C2 = 2 * C, S = C * C2, W = (S + 63) >> 6;
    const int per_group = round_up(2 * W * 8 + W * 4 + C2 * 4, 16);
    unsigned long long* mp = (unsigned long long*)(jrm + grp * per_group);  // slots with a record for p
    unsigned long long* mq = mp + W;                                        // slots with a record for q
    uint32_t* pw = (uint32_t*)(mq + W);                                     // records before each word
    int32_t* ids = (int32_t*)(pw + W);
    constexpr int RPW = 64 / G;  // rows per wave
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t r0 = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * RPW; r0 < n;
         r0 += waves * RPW) {  // wave-uniform
        const int64_t i = r0 + lane / G;
        const bool valid = i < n;
        for (int c = gl; c < C2; c += G) ids[c] = !valid ? -1 : c < C ? new_idx[i * C + c] : old_idx[i * C + (c - C)];
        for (int q = gl; q < 2 * W; q += G) mp[q] = 0ull;
        wave_sync();
        const uint32_t nk = valid ? (uint32_t)out.kc[i] : 0u;
        const JRec* E = out.U + (valid ? out.rowoff[i] : 0);
        for (uint32_t e = gl; e < nk; e += G) {
            const uint32_t t = E[e].t;
            const uint32_t s0 = t & kSlotMask;
            if (t & kEmitP) atomicOr(&mp[s0 >> 6], 1ull << (s0 & 63));
            if (t & kEmitQ) atomicOr(&mq[s0 >> 6], 1ull << (s0 & 63));
        }
        wave_sync();
        uint32_t carry = 0;
        for (int q0 = 0; q0 < W; q0 += G) {  // exclusive scan of the words' record counts
            const int q = q0 + gl;
            const uint32_t pc = q < W ? (uint32_t)(__popcll(mp[q]) + __popcll(mq[q])) : 0u;
            uint32_t incl = pc;
            for (int o = 1; o < G; o <<= 1) {
                const uint32_t v = __shfl_up(incl, o, G);
                if (gl >= o) incl += v;
            }
            if (q < W) pw[q] = carry + incl - pc;
            carry += __shfl(incl, G - 1, G);
        }
        wave_sync();
        const uint64_t base = valid ? koff[i] : 0;
        for (uint32_t e = gl; e < nk; e += G) {
            const JRec rec = E[e];
            const uint32_t s0 = rec.t & kSlotMask;
            const int q = (int)(s0 >> 6);
            const unsigned long long below = (1ull << (s0 & 63)) - 1ull;
            uint64_t rank = base + pw[q] + (uint64_t)(__popcll(mp[q] & below) + __popcll(mq[q] & below));
            const int j = (int)(s0 / (uint32_t)C2), kk = (int)(s0 - (uint32_t)j * C2);
            const int32_t p = ids[j], qq = ids[kk];
            const uint32_t dbits = __builtin_bit_cast(uint32_t, rec.d);
            if (rec.t & kEmitP) {
                rkey[rank] = (uint32_t)p;
                rval[rank] = ((uint64_t)dbits << 32) | (uint32_t)qq;
                ++rank;
            }
            if (rec.t & kEmitQ) {
                rkey[rank] = (uint32_t)qq;
                rval[rank] = ((uint64_t)dbits << 32) | (uint32_t)p;
            }
        }
        wave_sync();
    }
}

// REG (k <= 16): the lane mirrors its heap's ids in registers (one slot per entry, unordered), so the
// duplicate test of a push below the root is a register compare; an accepted push replaces one occurrence
// of the evicted id by the new one.
template <bool REG>
__global__ __launch_bounds__(256) void k_heap_push_rec(const int32_t* __restrict__ order, const uint32_t* bins,
                                                       const uint64_t* __restrict__ off, const uint64_t* __restrict__ rv,
                                                       int32_t* idx, float* dist, uint8_t* flags, int k,
                                                       unsigned long long* changes, uint32_t* row_changes) {
    extern __shared__ __align__(16) float hsm2[];
    const int B = blockDim.x, c = threadIdx.x;
    const int S = B + 1;  // heap column stride (odd: the row staging is conflict-free)
    float* hd = hsm2;
    int32_t* hi = (int32_t*)(hd + S * k);
    uint8_t* hf = (uint8_t*)(hi + S * k);
    const int64_t total = (int64_t)bins[kBins + kBins - 1] + bins[kBins - 1];
    unsigned long long ch = 0;
    for (int64_t e0 = (int64_t)blockIdx.x * B; e0 < total; e0 += (int64_t)gridDim.x * B) {
        const int64_t e = total - 1 - (e0 + c);  // longest bins first: they start early instead of trailing
        const int32_t v = e >= 0 ? order[e] : -1;
        wave_heap_rows<true>(v, k, dist, idx, flags, hd, hi, hf, S, c & ~63);
        if (e >= 0) {
            const unsigned long long ch0 = ch;
            const uint64_t o0 = off[v];
            const int64_t L = (int64_t)(off[v + 1] - o0);
            const uint64_t* R = rv + o0;
            float root = hd[c];
            int32_t hid[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) hid[j] = REG && j < k ? hi[j * S + c] : -1;
            auto offer = [&](uint64_t q) {
                const float d = __builtin_bit_cast(float, (uint32_t)(q >> 32));
                const int32_t x = (int32_t)(uint32_t)q;
                if (d >= root) return;
                if (REG) {
                    bool dup = false;
#pragma unroll
                    for (int j = 0; j < 16; ++j) dup |= hid[j] == x;
                    if (dup) return;
                    int32_t ev;
                    root = heap_push_col_new(hd, hi, hf, S, c, k, x, d, ev);
                    bool done = false;
#pragma unroll
                    for (int j = 0; j < 16; ++j) {
                        const bool hit = !done && j < k && hid[j] == ev;
                        hid[j] = hit ? x : hid[j];
                        done |= hit;
                    }
                    ++ch;
                } else if (heap_push_lane(hd, hi, hf, S, c, k, x, d)) {
                    ++ch;
                    root = hd[c];
                }
            };
            int64_t r = 0;
            for (; r + 8 <= L; r += 8) {  // eight records in flight per batch
                uint64_t q[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) q[u] = R[r + u];
#pragma unroll
                for (int u = 0; u < 8; ++u) offer(q[u]);
            }
            for (; r < L; ++r) offer(R[r]);
            if (row_changes) row_changes[v] = (uint32_t)(ch - ch0);
        }
        wave_heap_rows<false>(v, k, dist, idx, flags, hd, hi, hf, S, c & ~63);
    }
    for (int o = 32; o > 0; o >>= 1) ch += __shfl_xor(ch, o);
    // one atomic per workgroup: atomics on one address serialise in the L2
    __shared__ unsigned long long wch[16];
    if ((threadIdx.x & 63) == 0) wch[threadIdx.x >> 6] = ch;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) ch += wch[i];
        if (ch) atomicAdd(changes, ch);
    }
}

// The same ordered apply for heaps too large for a lane's LDS column (64 columns of k entries do not fit):
// one wave per destination, its heap row contiguous in LDS.  The wave tests each record's id against
// the whole row at once (a ballot) and lane 0 sifts, so the pushes stay in the records' order.
constexpr int kWaveApplyWaves = 4;
__global__ __launch_bounds__(kWaveApplyWaves * RINND_WAVE) void k_heap_push_rec_wave(
    const int32_t* __restrict__ order, const uint32_t* bins, const uint64_t* __restrict__ off,
    const uint64_t* __restrict__ rv, int32_t* idx, float* dist, uint8_t* flags, int k, unsigned long long* changes,
    uint32_t* row_changes) {
    extern __shared__ __align__(16) float wsm[];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    float* hd = wsm + (size_t)w * k;
    int32_t* hi = (int32_t*)(wsm + (size_t)kWaveApplyWaves * k) + (size_t)w * k;
    uint8_t* hf = (uint8_t*)(wsm + (size_t)2 * kWaveApplyWaves * k) + (size_t)w * k;
    const int64_t total = (int64_t)bins[kBins + kBins - 1] + bins[kBins - 1];
    unsigned long long ch = 0;
    for (int64_t e = (int64_t)blockIdx.x * kWaveApplyWaves + w; e < total; e += (int64_t)gridDim.x * kWaveApplyWaves) {
        const int32_t v = order[e];
        const int64_t row = (int64_t)v * k;
        for (int j = lane; j < k; j += RINND_WAVE) {
            hd[j] = dist[row + j];
            hi[j] = idx[row + j];
            hf[j] = flags[row + j];
        }
        wave_sync();
        const unsigned long long ch0 = ch;
        const uint64_t o0 = off[v];
        const int64_t L = (int64_t)(off[v + 1] - o0);
        for (int64_t r = 0; r < L; ++r) {
            const uint64_t q = rv[o0 + r];
            const float d = __builtin_bit_cast(float, (uint32_t)(q >> 32));
            const int32_t x = (int32_t)(uint32_t)q;
            if (d >= hd[0]) continue;  // uniform: every lane reads the same root
            bool dup = false;
            for (int j = lane; j < k; j += RINND_WAVE) dup |= hi[j] == x;
            if (__ballot(dup)) continue;
            if (lane == 0) {
                int32_t ev;
                (void)heap_push_col_new(hd, hi, hf, 1, 0, k, x, d, ev);
                ++ch;
            }
            wave_sync();
        }
        wave_sync();
        for (int j = lane; j < k; j += RINND_WAVE) {
            dist[row + j] = hd[j];
            idx[row + j] = hi[j];
            flags[row + j] = hf[j];
        }
        if (row_changes && lane == 0) row_changes[v] = (uint32_t)(ch - ch0);
        wave_sync();
    }
    if (lane == 0 && ch) atomicAdd(changes, ch);
}

size_t join_sort_tmp_bytes(int64_t items, int end_bit) {
    size_t b = 0;
    HIP_TRY(sort_pairs_u32_u64(nullptr, b, nullptr, nullptr, nullptr, nullptr, items, end_bit, nullptr));
    return b;
}

void launch_join_apply(hipStream_t s, const JoinApplyArgs& a) {
    const int C2 = 2 * a.C, S = a.C * C2, W = (S + 63) >> 6;
    const size_t per_group = (size_t)round_up(2 * W * 8 + W * 4 + C2 * 4, 16);
    if (W <= 16)
        hipLaunchKernelGGL(k_join_records<16>, dim3(grid_for(ceil_div(a.n, (int64_t)16), 1, 1 << 18)), dim3(256),
                           16 * per_group, s, a.new_idx, a.old_idx, a.C, a.n, a.out, a.koff, a.rkey, a.rval);
    else
        hipLaunchKernelGGL(k_join_records<64>, dim3(grid_for(ceil_div(a.n, (int64_t)4), 1, 1 << 18)), dim3(256),
                           4 * per_group, s, a.new_idx, a.old_idx, a.C, a.n, a.out, a.koff, a.rkey, a.rval);
    int end_bit = 1;
    while (((int64_t)1 << end_bit) < a.n) ++end_bit;
    if (a.n_records > 0) {
        size_t tb = a.sort_tmp_bytes;
        HIP_TRY(sort_pairs_u32_u64(a.sort_tmp, tb, a.rkey, a.rkey2, a.rval, a.rval2, (int64_t)a.n_records, end_bit,
                                   s));
    }
    hipLaunchKernelGGL(k_seg_starts, dim3(grid_for((int64_t)a.n_records + 1, 256)), dim3(256), 0, s, a.rkey2,
                       (int64_t)a.n_records, a.n, a.off);
    // destinations binned by record count, longest first
    launch_zero(s, a.bins, 3 * kBins * sizeof(uint32_t));
    hipLaunchKernelGGL(k_bin_count, dim3(grid_for(a.n, 256, 2048)), dim3(256), 0, s, a.off, a.n, a.bins);
    hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(64), 0, s, a.bins);
    hipLaunchKernelGGL(k_bin_scatter, dim3(grid_for(a.n, 256, 4096)), dim3(256), 0, s, a.off, a.n, a.bins, a.order);
    const int B = lanes_block(a.k, 9, 256, 1);
    if (B == 0) {  // large k: a wave per destination
        const size_t lds = (size_t)kWaveApplyWaves * round_up((int64_t)a.k * 9, 16);
        hipLaunchKernelGGL(k_heap_push_rec_wave, dim3(grid_for(a.n, kWaveApplyWaves, 8192)),
                           dim3(kWaveApplyWaves * RINND_WAVE), lds, s, a.order, a.bins, a.off, a.rval2, a.idx, a.dist,
                           a.flags, a.k, a.changes, a.row_changes);
        return;
    }
    const size_t lds = (size_t)round_up((int64_t)(B + 1) * a.k * 2 + ((B + 1) * a.k + 3) / 4, 4) * 4;
    if (a.k <= 16)
        hipLaunchKernelGGL(k_heap_push_rec<true>, dim3(grid_for(a.n, B, 8192)), dim3(B), lds, s, a.order, a.bins,
                           a.off, a.rval2, a.idx, a.dist, a.flags, a.k, a.changes, a.row_changes);
    else
        hipLaunchKernelGGL(k_heap_push_rec<false>, dim3(grid_for(a.n, B, 8192)), dim3(B), lds, s, a.order, a.bins,
                           a.off, a.rval2, a.idx, a.dist, a.flags, a.k, a.changes, a.row_changes);
}

// ---- sort_all (neighbor_heap.rs:293-320): stable ascending sort per row -------------------------
__global__ __launch_bounds__(256) void k_sort_rows(int32_t* idx, float* dist, uint8_t* flags,
                                                   int64_t n, int k) {
    __shared__ uint64_t keys[4][256];
    __shared__ int32_t si[4][256];
    __shared__ float sd[4][256];
    __shared__ uint8_t sf[4][256];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int S = 1;
    while (S < k) S <<= 1;
    for (int64_t row = (int64_t)blockIdx.x * 4 + w; row < n; row += (int64_t)gridDim.x * 4) {
        const int64_t o = row * k;
        for (int j = lane; j < S; j += 64) {
            if (j < k) {
                si[w][j] = idx[o + j];
                sd[w][j] = dist[o + j];
                sf[w][j] = flags[o + j];
                keys[w][j] = ((uint64_t)float_order_key(sd[w][j]) << 32) | (uint32_t)j;
            } else {
                keys[w][j] = ~0ull;
            }
        }
        wave_sync();
        for (int kk = 2; kk <= S; kk <<= 1)
            for (int j = kk >> 1; j > 0; j >>= 1) {
                for (int i = lane; i < S; i += 64) {
                    const int ixj = i ^ j;
                    if (ixj > i) {
                        const bool up = (i & kk) == 0;
                        const uint64_t A = keys[w][i], B = keys[w][ixj];
                        if ((A > B) == up) {
                            keys[w][i] = B;
                            keys[w][ixj] = A;
                        }
                    }
                }
                wave_sync();
            }
        for (int j = lane; j < k; j += 64) {
            const int src = (int)(keys[w][j] & 0xffffffffu);
            idx[o + j] = si[w][src];
            dist[o + j] = sd[w][src];
            flags[o + j] = sf[w][src];
        }
        wave_sync();
    }
}
// The same for rows of at most 64 entries, in registers: S (k rounded up to a power of two) lanes per row,
// 64 / S rows per wave, the bitonic network on (distance order key, slot) by lane exchanges; each lane then
// takes the entry its rank names from the lane that holds it.
__global__ __launch_bounds__(256) void k_sort_rows_reg(int32_t* idx, float* dist, uint8_t* flags, int64_t n, int k,
                                                       int S) {
    const int lane = threadIdx.x & 63;
    const int G = 64 / S, j = lane % S, base = lane - j;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t row0 = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * G; row0 < n;
         row0 += waves * G) {  // wave-uniform
        const int64_t row = row0 + lane / S;
        const bool ok = row < n && j < k;
        int32_t vi = 0;
        float vd = 0.f;
        uint32_t vf = 0u;
        uint64_t key = ~0ull;
        if (ok) {
            vi = idx[row * k + j];
            vd = dist[row * k + j];
            vf = flags[row * k + j];
            key = ((uint64_t)float_order_key(vd) << 32) | (uint32_t)j;
        }
        for (int kk = 2; kk <= S; kk <<= 1)
            for (int jj = kk >> 1; jj > 0; jj >>= 1) {
                const uint32_t lo = __shfl_xor((uint32_t)key, jj), hi = __shfl_xor((uint32_t)(key >> 32), jj);
                const uint64_t other = ((uint64_t)hi << 32) | lo;
                const bool take_min = ((j & jj) == 0) == ((j & kk) == 0);
                key = take_min ? (other < key ? other : key) : (other > key ? other : key);
            }
        const int src = base + (int)(key & 0x3fu);  // rank j's entry sits in lane base + its slot
        const int32_t oi = __shfl(vi, src);
        const float od = __shfl(vd, src);
        const uint32_t of = __shfl(vf, src);
        if (ok) {
            idx[row * k + j] = oi;
            dist[row * k + j] = od;
            flags[row * k + j] = (uint8_t)of;
        }
    }
}

void launch_sort_rows(hipStream_t s, int32_t* idx, float* dist, uint8_t* flags, int64_t n, int k) {
    if (n == 0) return;
    int S = 1;
    while (S < k) S <<= 1;
    if (S <= 64)
        hipLaunchKernelGGL(k_sort_rows_reg, dim3(grid_for(ceil_div(n, (int64_t)(64 / S)), 4, 1 << 16)), dim3(256), 0, s,
                           idx, dist, flags, n, k, S);
    else
        hipLaunchKernelGGL(k_sort_rows, dim3(grid_for(n, 4, 1 << 16)), dim3(256), 0, s, idx, dist, flags, n, k);
}

// distance corrections (core/index.rs:669-673)
__global__ void k_correct(float* d, int64_t count, int corr) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < count;
         t += (int64_t)gridDim.x * blockDim.x) {
        const float x = d[t];
        float y = x;
        if (corr == RINND_CORR_SQRT) y = sqrtf(x);
        else if (corr == RINND_CORR_ALT_COSINE) y = 1.0f - powf(2.0f, -x);
        else if (corr == RINND_CORR_ALT_INNER_PRODUCT) y = x >= 3.40282347e+38f ? 0.0f : -1.0f / x;
        d[t] = y;
    }
}
void launch_correct(hipStream_t s, float* dist, int64_t count, int corr) {
    if (count == 0 || corr == RINND_CORR_NONE) return;
    hipLaunchKernelGGL(k_correct, dim3(grid_for(count, 256)), dim3(256), 0, s, dist, count, corr);
}

// ---- scans ---------------------------------------------------------------------------------------
size_t scan_u64_tmp_bytes(int64_t n) {
    size_t bytes = 0;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (unsigned long long*)nullptr,
                                             (unsigned long long*)nullptr, (int)n));
    return bytes;
}
void scan_u64_exclusive(hipStream_t s, const unsigned long long* in, uint64_t* out, int64_t n,
                        void* tmp, size_t tmp_bytes) {
    size_t bytes = tmp_bytes;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, (unsigned long long*)out, (int)n, s));
}

}  // namespace rinnd
