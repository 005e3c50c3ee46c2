// forest_level.hip - RP forest with counter-based draws, built level-synchronously over all trees.
//
// Same tree algorithm as TreeBuilder (core/tree/builder.rs:172-348): a node is a leaf when
// m <= leaf_size or depth >= max_depth; otherwise it splits by the bisector of two random members
// (euclidean) or by the normalised difference of their directions (angular), a point goes left iff
// its margin is < 0 (ties |margin| < 1e-8 decided by a random bit per point), and a degenerate split (all
// points on one side) makes the node a leaf.  The unit normal is rounded to int8 codes q times a power of two
// s (the largest coordinate takes 7 bits; the stored f32 hyperplane holds exactly q s, and the side tests read
// the one-byte codes: a quarter of the f32 plane's bytes); the plane passes through the midpoint of the two
// members.  Since s is a power of two, the margin x . (q s) + off is s (x . q + off / s) exactly: the side
// tests sum x . q and compare with -off / s (ties below 1e-8 / s).  The random draws come from mix64 of the tree key
// and node id (the two members) or of the tree key, node id and point (a tie) instead of a depth-first
// FastRng stream, so every node of a level is independent and one level is processed for all trees at once:
//   k_lvl_planes   one workgroup per active node: split points, hyperplane, offset (or leaf)
//   k_lvl_margins  16 lanes per point, all trees: the point's entry in each tree advanced from the last
//                  level's side and the children table, then the side of the tree's current hyperplane,
//                  kept as one bit per (tree, point) in per-point 16-bit masks
//   flag chunks    left flags in (tree, position) order as bytes (dense levels: gathered from the masks
//                  through each tree's order) and their count per chunk of 256; an exclusive scan of the
//                  chunk counts gives the prefix of any item with at most 256 bytes read
//   k_lvl_counts   left size per node; degenerate splits become leaves; children counted per tree
//   k_lvl_assign   child node ids (breadth-first per tree), the next level's active list and each
//                  entry's first child entry
//   k_lvl_scatter  stable partition of every split node's positions
// Dense levels write nothing indexed by position from the point-major side tests (no scattered writes);
// sparse levels test position-major and write position-ordered flags.  Once the active entries average at
// most 1,024 of the tree x point positions, flag chunks, their scan, k_lvl_counts and k_lvl_scatter give way
// to one pass per entry (k_lvl_part_wave: a wave per entry of at most 1,024 members; k_lvl_part_entry: a
// workgroup per larger one) that reads the entry's flags, counts them and partitions the entry's positions
// in the same stable order.
// Node ids are breadth-first per tree (root 0); leaves hold their [start, end) range of
// tree.indices exactly as in the depth-first layout.
#include <hipcub/hipcub.hpp>

#include "internal.h"

namespace rinnd {

struct LvlNode {
    int32_t tree, node, start, end, depth, parent;  // parent: the entry one level up (-1 for the roots)
};

// What the side tests of an active entry read: its node id (-1 if it does not split) and offset.
struct PInfo {
    int32_t node;
    float off;  // the plane's offset over its scale: the side test compares sum x q + off with 0
    float tie;  // 1e-8 over the scale: |sum x q + off| below it is a margin tie
};
// The same for the two children of an entry of the last level (left child first), with the child's entry
// (-1 for both when the parent did not split), so that a point moves to its next entry and finds its next
// node with one 16-byte load.
struct CInfo {
    int32_t node;
    float off;
    int32_t entry;
    float tie;
};

struct LvlArgs {
    const float* X;
    int64_t ld;
    int dim;
    int64_t n;
    int64_t leaf_size, max_depth;
    int angular;
    const uint64_t* keys;  // per tree
    float** hyper;         // per tree, cap x dim (the codes times the plane's power-of-two scale)
    float** offsets;       // per tree, cap
    int32_t** children;    // per tree, cap x 2
    int8_t** hyper_q;      // per tree, cap x dim int8 codes q of the planes read by the side tests
    int n_trees;
    const int8_t* hq0;     // the same codes as one array: tree t's node v at hq0 + (t * cap + v) * dim
    int64_t cap;
};

__device__ __forceinline__ float bf16_round(float f) {  // round to nearest even, back to f32
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return __builtin_bit_cast(float, u & 0xffff0000u);
}
__device__ __forceinline__ float q8(uint32_t w, int b) { return (float)(int8_t)(w >> (8 * b)); }  // code b of w

static inline unsigned lgrid(int64_t work, int block, int64_t cap = 1 << 20) {
    const int64_t g = ceil_div(work, (int64_t)block);
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

__device__ __forceinline__ float block_max(float v, float* red) {  // order-independent
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    float t = 0.f;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t = fmaxf(t, red[i]);
    return t;
}

// The power of two s with max |h| / s in [64, 128) (1 for a zero plane): codes round(h / s) fit in [-127, 127].
__device__ __forceinline__ float plane_scale(float mx) {
    if (!(mx > 0.f)) return 1.f;
    int e = 0;
    (void)frexpf(mx, &e);  // mx = m 2^e, m in [0.5, 1)
    return ldexpf(1.f, e - 7);
}
__device__ __forceinline__ int8_t plane_code(float h, float sc) {
    return (int8_t)fminf(127.f, fmaxf(-127.f, rintf(h / sc)));
}

__device__ __forceinline__ float block_sum(float v, float* red) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    float t = 0.f;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += red[i];
    return t;
}

__global__ __launch_bounds__(256) void k_lvl_planes(LvlArgs a, const LvlNode* nodes, int64_t na,
                                                    const int32_t* const* perm, uint8_t* split, PInfo* pinfo,
                                                    CInfo* cinfo, unsigned long long* active_pairs) {
    __shared__ float red[4];
    unsigned long long act = 0;  // members of split nodes: the (tree, point) pairs tested this level
    for (int64_t e = blockIdx.x; e < na; e += gridDim.x) {
        const LvlNode nd = nodes[e];
        const int64_t m = nd.end - nd.start;
        int32_t* ch = a.children[nd.tree];
        if (m <= a.leaf_size || nd.depth >= a.max_depth) {
            if (threadIdx.x == 0) {
                ch[2 * nd.node] = -nd.start;
                ch[2 * nd.node + 1] = -nd.end;
                split[e] = 0;
                pinfo[e] = PInfo{-1, 0.f, 0.f};
                if (nd.parent >= 0) cinfo[2 * nd.parent + (e & 1)] = CInfo{-1, 0.f, (int32_t)e, 0.f};
            }
            continue;  // uniform across the block
        }
        const uint64_t key = a.keys[nd.tree];
        const uint64_t h1 = mix64(key ^ mix64(2 * (uint64_t)nd.node + 1));
        const uint64_t h2 = mix64(key ^ mix64(2 * (uint64_t)nd.node + 2));
        const int64_t i1 = (int64_t)(h1 % (uint64_t)m);
        int64_t i2 = (int64_t)(h2 % (uint64_t)(m - 1));
        if (i2 >= i1) ++i2;
        const float* p1 = a.X + (int64_t)perm[nd.tree][nd.start + i1] * a.ld;
        const float* p2 = a.X + (int64_t)perm[nd.tree][nd.start + i2] * a.ld;
        float* hp = a.hyper[nd.tree] + (int64_t)nd.node * a.dim;
        int8_t* hq = a.hyper_q[nd.tree] + (int64_t)nd.node * a.dim;
        // the unit normal h (or the raw difference below the norm threshold), its scale and codes; the stored
        // plane is q s exactly
        float sc, off = 0.f;
        if (!a.angular) {
            float nsq = 0.f;
            for (int i = threadIdx.x; i < a.dim; i += blockDim.x) {
                const float h = p2[i] - p1[i];
                nsq += h * h;
            }
            nsq = block_sum(nsq, red);
            const float norm = sqrtf(nsq);
            const bool scale = norm > 1e-8f;
            const float inv = scale ? 1.0f / norm : 0.f;
            float mx = 0.f;
            for (int i = threadIdx.x; i < a.dim; i += blockDim.x)
                mx = fmaxf(mx, fabsf(scale ? (p2[i] - p1[i]) * inv : p2[i] - p1[i]));
            sc = plane_scale(block_max(mx, red));
            // the plane passes through the midpoint of the two members
            float dm = 0.f;
            for (int i = threadIdx.x; i < a.dim; i += blockDim.x) {
                const int8_t q = plane_code(scale ? (p2[i] - p1[i]) * inv : p2[i] - p1[i], sc);
                const float h = (float)q * sc;
                hp[i] = h;
                hq[i] = q;
                dm += h * ((p1[i] + p2[i]) * 0.5f);
            }
            dm = block_sum(dm, red);
            off = scale ? -dm : 0.f;
        } else {
            float n1 = 0.f, n2 = 0.f;
            for (int i = threadIdx.x; i < a.dim; i += blockDim.x) {
                n1 += p1[i] * p1[i];
                n2 += p2[i] * p2[i];
            }
            n1 = sqrtf(block_sum(n1, red));
            n2 = sqrtf(block_sum(n2, red));
            const bool deg = n1 < 1e-8f || n2 < 1e-8f;
            float nn = 0.f;
            for (int i = threadIdx.x; i < a.dim; i += blockDim.x) {
                const float h = deg ? 0.f : p2[i] / n2 - p1[i] / n1;
                nn += h * h;
            }
            nn = sqrtf(block_sum(nn, red));
            float mx = 0.f;
            for (int i = threadIdx.x; i < a.dim; i += blockDim.x) {
                const float h = deg ? 0.f : p2[i] / n2 - p1[i] / n1;
                mx = fmaxf(mx, fabsf(nn > 1e-8f ? h / nn : h));
            }
            sc = plane_scale(block_max(mx, red));
            for (int i = threadIdx.x; i < a.dim; i += blockDim.x) {
                const float h = deg ? 0.f : p2[i] / n2 - p1[i] / n1;
                const int8_t q = plane_code(nn > 1e-8f ? h / nn : h, sc);
                hp[i] = (float)q * sc;
                hq[i] = q;
            }
        }
        if (threadIdx.x == 0) {
            const float off_s = off / sc, tie_s = 1e-8f / sc;
            a.offsets[nd.tree][nd.node] = off;
            pinfo[e] = PInfo{nd.node, off_s, tie_s};
            if (nd.parent >= 0) cinfo[2 * nd.parent + (e & 1)] = CInfo{nd.node, off_s, (int32_t)e, tie_s};
        }
        if (threadIdx.x == 0) split[e] = 1;
        act += (unsigned long long)m;
    }
    if (threadIdx.x == 0 && act) atomicAdd(active_pairs, act);
}

// r = x . q + off / s of a plane q s (s a power of two): the point goes left iff r < 0; |r| < tie (= 1e-8 / s, a
// margin below 1e-8) is a tie, decided by a random bit per (tree, node, point).
__device__ __forceinline__ uint32_t side_of(float margin, float tie, uint64_t key, int32_t node, int64_t point) {
    if (fabsf(margin) < tie)  // ties: a random bit per (tree, node, point)
        return (uint32_t)(mix64(key ^ mix64(((uint64_t)(uint32_t)node << 32) ^ (uint64_t)point)) & 1);
    return margin < 0.0f ? 1u : 0u;
}

// The side-test info of point i in tree t at this level.  At level 0 (cinfo null) its entry is tree t's root
// and the info that entry's PInfo.  Later, its entry one level up (nodeOf) and the side bit the last level
// left in the point's mask word choose one of that entry's two CInfo (children entries are consecutive, the
// left one first); the point's entry moves to the child's.  nodeOf is point-major (point i's entries of the
// n_trees trees are consecutive, one segment for the lanes that test the point) and rewritten in place (the
// lane that owns (t, i) is its only reader and writer).
__device__ __forceinline__ PInfo lvl_info(int64_t n, int n_trees, const PInfo* __restrict__ pinfo,
                                          const CInfo* __restrict__ cinfo, int32_t* nodeOf, const uint16_t* mask16,
                                          int t, int64_t i) {
    const int64_t ti = i * n_trees + t;
    const int32_t e = nodeOf[ti];
    if (e < 0) return PInfo{-1, 0.f, 0.f};
    if (!cinfo) return pinfo[e];
    const bool left = (mask16[(int64_t)(t >> 4) * n + i] >> (t & 15)) & 1u;
    const CInfo c = cinfo[2 * (int64_t)e + (left ? 0 : 1)];
    nodeOf[ti] = c.entry;
    return PInfo{c.entry >= 0 ? c.node : -1, c.off, c.tie};
}

// Side tests, point-major, 16 lanes per point (any dim).  First the lanes fetch, in parallel, the point's
// entry, node and offset for every tree (lane l owns trees l, l+16, ...); then the point's row, held in
// registers (CH float4 per lane, dim <= 64 * CH), is tested against the split trees' hyperplanes, four
// trees per batch so their hyperplane loads are in flight together.  The left bits of trees 16 j .. 16 j + 15
// go to the point's mask word j.
template <int CH>
__global__ __launch_bounds__(256) void k_lvl_margins(LvlArgs a, const PInfo* __restrict__ pinfo, int n_trees,
                                                     const CInfo* __restrict__ cinfo, int32_t* nodeOf,
                                                     uint16_t* mask16) {
    constexpr int kMaxT = 64;  // trees per point handled by the lane tables (4 per lane)
    const int gl = threadIdx.x & 15;
    const int gbase = threadIdx.x & 48;  // first lane of this 16-lane group within the wave
    const int64_t groups = ((int64_t)gridDim.x * blockDim.x) >> 4;
    for (int64_t i = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 4; i < a.n; i += groups) {
        // per-lane tables for trees t = gl + 16 j
        int32_t tnode[kMaxT / 16];
        float toff[kMaxT / 16], ttie[kMaxT / 16];
        uint32_t mask = 0;  // bit j: tree gl + 16 j splits this point's node
#pragma unroll
        for (int j = 0; j < kMaxT / 16; ++j) {
            const int t = gl + 16 * j;
            tnode[j] = 0;
            toff[j] = 0.f;
            ttie[j] = 0.f;
            if (t < n_trees) {
                const PInfo pi = lvl_info(a.n, n_trees, pinfo, cinfo, nodeOf, mask16, t, i);
                if (pi.node >= 0) {
                    tnode[j] = pi.node;
                    toff[j] = pi.off;
                    ttie[j] = pi.tie;
                    mask |= 1u << j;
                }
            }
        }
        float4 xr[CH > 0 ? CH : 1];
        const float* x = a.X + i * a.ld;
        if (CH > 0) {
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const int d0 = 4 * gl + 64 * c;
                xr[c] = d0 < a.dim ? *reinterpret_cast<const float4*>(x + d0) : make_float4(0.f, 0.f, 0.f, 0.f);
            }
        }
        uint32_t lbits = 0;  // bit j: tree gl + 16 j sends the point left
        // walk the split trees of this point in batches of four
        for (int j = 0; j < kMaxT / 16; ++j) {
            // trees t = l + 16 j over the group's lanes l whose bit j is set
            uint64_t wm = __ballot((mask >> j) & 1u);
            uint32_t gm = (uint32_t)((wm >> gbase) & 0xffffu);
            while (gm) {
                int tl[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    tl[u] = gm ? __ffs((int)gm) - 1 : -1;
                    if (gm) gm &= gm - 1;
                }
                float acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (tl[u] < 0) continue;
                    const int t = tl[u] + 16 * j;
                    const int32_t node = __shfl(tnode[j], gbase + tl[u]);
                    const int8_t* hq = a.hyper_q[t] + (int64_t)node * a.dim;
                    if (CH > 0) {
#pragma unroll
                        for (int c = 0; c < CH; ++c) {
                            const int d0 = 4 * gl + 64 * c;
                            if (d0 < a.dim) {
                                const uint32_t hv = *reinterpret_cast<const uint32_t*>(hq + d0);
                                acc[u] += xr[c].x * q8(hv, 0) + xr[c].y * q8(hv, 1) + xr[c].z * q8(hv, 2) +
                                          xr[c].w * q8(hv, 3);
                            }
                        }
                    } else {
                        for (int d = gl; d < a.dim; d += 16) acc[u] += x[d] * (float)hq[d];
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    for (int o = 8; o > 0; o >>= 1) acc[u] += __shfl_xor(acc[u], o);  // inside the 16-lane group
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (tl[u] != gl) continue;  // the lane that owns tree t keeps its side
                    const int t = gl + 16 * j;
                    lbits |= side_of(acc[u] + toff[j], ttie[j], a.keys[t], tnode[j], i) << j;
                }
            }
        }
        for (int j = 0; 16 * j < n_trees; ++j) {  // words of batches where some tree splits this point's node
            const uint32_t any = (uint32_t)((__ballot((mask >> j) & 1u) >> gbase) & 0xffffu);
            const uint32_t bits = (uint32_t)((__ballot((lbits >> j) & 1u) >> gbase) & 0xffffu);
            if (gl == 0 && any) mask16[(int64_t)j * a.n + i] = (uint16_t)bits;
        }
    }
}

template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xf, 0xf, false));
}
// Lane u's value in every lane of its 16-lane DPP row (row_newbcast; u is a constant once unrolled).
__device__ __forceinline__ int row_bcast(int v, int u) {
    switch (u) {
        case 0: return dpp_i<0x150>(v);
        case 1: return dpp_i<0x151>(v);
        case 2: return dpp_i<0x152>(v);
        case 3: return dpp_i<0x153>(v);
        case 4: return dpp_i<0x154>(v);
        case 5: return dpp_i<0x155>(v);
        case 6: return dpp_i<0x156>(v);
        case 7: return dpp_i<0x157>(v);
        case 8: return dpp_i<0x158>(v);
        case 9: return dpp_i<0x159>(v);
        case 10: return dpp_i<0x15A>(v);
        case 11: return dpp_i<0x15B>(v);
        case 12: return dpp_i<0x15C>(v);
        case 13: return dpp_i<0x15D>(v);
        case 14: return dpp_i<0x15E>(v);
        default: return dpp_i<0x15F>(v);
    }
}
// Sum over the 16 lanes of a DPP row (rotations by 8, 4, 2, 1), in every lane of the row.
__device__ __forceinline__ float row16_sum(float v) {
    v += dpp_f<0x128>(v);
    v += dpp_f<0x124>(v);
    v += dpp_f<0x122>(v);
    v += dpp_f<0x121>(v);
    return v;
}

// Dense levels, point-major: a 16-lane group (one DPP row) per point holds the point's row in registers
// (lane l: dims 8 (l + 16 r) .. +8).  Trees go in batches of 16: lane u advances tree u's entry and
// fetches its PInfo, the group broadcasts each tree's node from its lane (row_newbcast), reads that tree's
// plane codes with one coalesced 8-byte load per lane, and reduces the 16 partial dot products with row
// rotations; lane u keeps tree u's margin, and the batch's 16 sides become the point's mask word.
// Needs dim % 8 == 0 and dim <= 128 R.
template <int R>
__global__ __launch_bounds__(256) void k_lvl_margins_pm(LvlArgs a, const PInfo* __restrict__ pinfo, int n_trees,
                                                        const CInfo* __restrict__ cinfo, int32_t* nodeOf,
                                                        uint16_t* mask16) {
    const int gl = threadIdx.x & 15;
    const int64_t groups = ((int64_t)gridDim.x * blockDim.x) >> 4;
    const int64_t tstride = a.cap * a.dim;
    for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4; i < a.n; i += groups) {  // group-uniform
        float4 xr[2 * R];
        const float* x = a.X + i * a.ld;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int d0 = 8 * (gl + 16 * r);
            if (d0 < a.dim) {
                xr[2 * r] = *reinterpret_cast<const float4*>(x + d0);
                xr[2 * r + 1] = *reinterpret_cast<const float4*>(x + d0 + 4);
            } else {
                xr[2 * r] = xr[2 * r + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
        }
        for (int tb = 0; tb < n_trees; tb += 16) {
            const int tl = tb + gl;
            int32_t nodeL = -1;
            float offL = 0.f, tieL = 0.f;
            if (tl < n_trees) {
                const PInfo pi = lvl_info(a.n, n_trees, pinfo, cinfo, nodeOf, mask16, tl, i);
                nodeL = pi.node;
                offL = pi.off;
                tieL = pi.tie;
            }
            // group-uniform: no tree of the batch splits this point's node (its mask word is not read again)
            if (!((__ballot(nodeL >= 0) >> (threadIdx.x & 48)) & 0xffffull)) continue;
            float mine = 0.f;
            constexpr int UB = R >= 4 ? 2 : (R == 1 ? 16 : 8 / R);  // planes in flight per batch (all 16 for dim <= 128)
#pragma unroll
            for (int ub = 0; ub < 16; ub += UB) {  // UB planes in flight, then their dot products
                uint2 h[UB][R];
#pragma unroll
                for (int u = 0; u < UB; ++u) {
                    // trees past n_trees and nodes that do not split read a valid plane whose result is unused
                    const int32_t nodeU = row_bcast(nodeL, ub + u);
                    const int tu = tb + ub + u < n_trees ? tb + ub + u : n_trees - 1;
                    const uint2* hb = reinterpret_cast<const uint2*>(
                        a.hq0 + (int64_t)tu * tstride + (int64_t)(nodeU > 0 ? nodeU : 0) * a.dim);
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const int c = gl + 16 * r;
                        h[u][r] = hb[8 * c < a.dim ? c : 0];  // chunks past dim meet zero row values
                    }
                }
#pragma unroll
                for (int u = 0; u < UB; ++u) {
                    float acc = 0.f;
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const uint2 hv = h[u][r];
                        acc = fmaf(xr[2 * r].x, q8(hv.x, 0), acc);
                        acc = fmaf(xr[2 * r].y, q8(hv.x, 1), acc);
                        acc = fmaf(xr[2 * r].z, q8(hv.x, 2), acc);
                        acc = fmaf(xr[2 * r].w, q8(hv.x, 3), acc);
                        acc = fmaf(xr[2 * r + 1].x, q8(hv.y, 0), acc);
                        acc = fmaf(xr[2 * r + 1].y, q8(hv.y, 1), acc);
                        acc = fmaf(xr[2 * r + 1].z, q8(hv.y, 2), acc);
                        acc = fmaf(xr[2 * r + 1].w, q8(hv.y, 3), acc);
                    }
                    acc = row16_sum(acc);
                    if (gl == ub + u) mine = acc;
                }
            }
            const uint32_t left = nodeL >= 0 ? side_of(mine + offL, tieL, a.keys[tl], nodeL, i) : 0u;
            const uint64_t bits = __ballot(left != 0);
            if (gl == 0) mask16[(int64_t)(tb >> 4) * a.n + i] = (uint16_t)(bits >> (threadIdx.x & 48));
        }
    }
}

// Shallow dense levels on MFMA: while the level's active entries (over all trees) number at most NE, a
// workgroup stages every split entry's plane codes in LDS once (as bf16, exactly), and each wave multiplies blocks of 16 points by
// all of them: the point's f32 row split into bf16 hi + lo parts (x = hi + lo up to 2^-16 relative), 16x16x32
// bf16 MFMAs with f32 accumulation, the 16 x na margins staged in LDS.  Then the wave's 64 lanes take
// (point p = lane % 16, trees lane / 16 + 4 i): the point's entry advances as in lvl_info, its margin is read
// at that entry, and the sides become the point's 16-bit mask words (assembled from ballots).  The margins are
// the same sums as the VALU kernels' in another order; a point within rounding of a plane may take the other
// side, which any rounding order allows.  Needs dim % 32 == 0, dim <= 256, n_trees <= 64.
template <int KS>  // dim / 32
__global__ __launch_bounds__(256) void k_lvl_margins_mfma(LvlArgs a, const LvlNode* __restrict__ cur, int na,
                                                          const PInfo* __restrict__ pinfo, int n_trees,
                                                          const CInfo* __restrict__ cinfo, int32_t* nodeOf,
                                                          uint16_t* mask16) {
    extern __shared__ __align__(16) uint16_t msm[];
    const int na16 = (na + 15) & ~15;
    const int pld = KS * 32 + 8;  // staged plane stride (bf16): 16-byte aligned, rows on different banks
    float* eoff = (float*)(msm + (int64_t)na16 * pld);   // per entry: offset over the plane's scale
    float* etie = eoff + na16;                           // per entry: tie threshold over the scale
    int32_t* enode = (int32_t*)(etie + na16);            // per entry: node id, -1 if it does not split
    float* csm = (float*)(enode + na16);                 // per wave: na16 x 16 margins (entry-major)
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    const int mrow = lane & 15, kqi = lane >> 4;
    // stage the planes' codes of the split entries as bf16 (exact; zeros elsewhere)
    const int v8 = KS * 4;  // 8-code chunks per plane
    for (int i = t; i < na16 * v8; i += blockDim.x) {
        const int e = i / v8, c = i - e * v8;
        uint2 v = make_uint2(0u, 0u);
        if (e < na) {
            const PInfo pi = pinfo[e];
            if (pi.node >= 0)
                v = *reinterpret_cast<const uint2*>(a.hq0 + ((int64_t)cur[e].tree * a.cap + pi.node) * a.dim + 8 * c);
        }
        uint32_t o[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t w = j < 2 ? v.x : v.y;
            const int b = 2 * (j & 1);
            o[j] = (__builtin_bit_cast(uint32_t, q8(w, b)) >> 16) | (__builtin_bit_cast(uint32_t, q8(w, b + 1)) & 0xffff0000u);
        }
        *reinterpret_cast<uint4*>(msm + e * pld + 8 * c) = make_uint4(o[0], o[1], o[2], o[3]);
    }
    for (int e = t; e < na16; e += blockDim.x) {
        const PInfo pi = e < na ? pinfo[e] : PInfo{-1, 0.f, 0.f};
        enode[e] = pi.node;
        eoff[e] = pi.off;
        etie[e] = pi.tie;
    }
    __syncthreads();
    float* cw = csm + (int64_t)w * na16 * 16;
    const int64_t nblk = (a.n + 15) >> 4;
    const int64_t wstride = (int64_t)gridDim.x * (blockDim.x >> 6);
    // software pipeline: the next block's rows and (first 16 trees) entries and mask word are loaded while the
    // current block is evaluated
    float4 rf[KS][2];
    int32_t pe0[4];  // entry of tree mrow for point 4 ii + kqi of the block
    uint16_t pmw[4];  // and that point's first mask word
    auto prefetch = [&](int64_t blk) {
        const int64_t pr0 = (blk << 4) + mrow;
        const bool ok = blk < nblk && pr0 < a.n;
        const int64_t pr = ok ? pr0 : 0;  // every address stays in bounds
        const float* x = a.X + pr * a.ld;
#pragma unroll
        for (int kk = 0; kk < KS; ++kk) {
            rf[kk][0] = ok ? *reinterpret_cast<const float4*>(x + 32 * kk + 8 * kqi) : make_float4(0.f, 0.f, 0.f, 0.f);
            rf[kk][1] = ok ? *reinterpret_cast<const float4*>(x + 32 * kk + 8 * kqi + 4) : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int ii = 0; ii < 4; ++ii) {  // the decisions' layout: four points of 16 trees, one segment each
            const int64_t qp = (blk << 4) + 4 * ii + kqi;
            const bool qok = blk < nblk && qp < a.n;
            pe0[ii] = qok && mrow < n_trees ? nodeOf[qp * n_trees + mrow] : -1;
            pmw[ii] = qok && cinfo ? mask16[qp] : (uint16_t)0;
        }
    };
    int64_t blk = (int64_t)blockIdx.x * (blockDim.x >> 6) + w;
    prefetch(blk);
    for (; blk < nblk; blk += wstride) {  // wave-uniform
        const int64_t p0 = blk << 4;
        // A operands: row p0 + mrow, dims 32 kk + 8 kqi .. + 7, as bf16 hi and lo parts
        v8bf ah[KS], al[KS];
#pragma unroll
        for (int kk = 0; kk < KS; ++kk) {
            const float f[8] = {rf[kk][0].x, rf[kk][0].y, rf[kk][0].z, rf[kk][0].w,
                                rf[kk][1].x, rf[kk][1].y, rf[kk][1].z, rf[kk][1].w};
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const float h = bf16_round(f[j]);
                ah[kk][j] = __builtin_bit_cast(__bf16, (uint16_t)(__builtin_bit_cast(uint32_t, h) >> 16));
                const float l = bf16_round(f[j] - h);
                al[kk][j] = __builtin_bit_cast(__bf16, (uint16_t)(__builtin_bit_cast(uint32_t, l) >> 16));
            }
        }
        int32_t ce0[4];
        uint16_t cmw[4];
#pragma unroll
        for (int ii = 0; ii < 4; ++ii) {
            ce0[ii] = pe0[ii];
            cmw[ii] = pmw[ii];
        }
        prefetch(blk + wstride);
        wave_sync();  // the previous block's margins are read
        for (int ct = 0; ct < (na16 >> 4); ++ct) {
            v4f acc = {0.f, 0.f, 0.f, 0.f};
            const uint16_t* brow = msm + (16 * ct + mrow) * pld + 8 * kqi;
#pragma unroll
            for (int kk = 0; kk < KS; ++kk) {
                const v8bf b = *reinterpret_cast<const v8bf*>(brow + 32 * kk);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[kk], b, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[kk], b, acc, 0, 0, 0);
            }
            // rows 4 kqi .. + 3 (points), column 16 ct + mrow (entry)
            *reinterpret_cast<float4*>(cw + (16 * ct + mrow) * 16 + 4 * kqi) = make_float4(acc[0], acc[1], acc[2], acc[3]);
        }
        wave_sync();
        // sides: in step ii lane (point p0 + 4 ii + kqi, tree 16 j + mrow), so that the 16 lanes of a point read
        // and write its entries as one segment; the point's mask word j is bits 16 kqi .. + 15 of the ballot
        for (int j = 0; 16 * j < n_trees; ++j) {  // uniform
#pragma unroll
            for (int ii = 0; ii < 4; ++ii) {
                const int pq = 4 * ii + kqi;  // the point's row in the block
                const int64_t p = p0 + pq;
                const bool pv = p < a.n;
                const int tr = 16 * j + mrow;
                uint32_t left = 0u;
                if (pv && tr < n_trees) {
                    const int64_t ti = p * n_trees + tr;
                    const int32_t e0 = j == 0 ? ce0[ii] : nodeOf[ti];
                    int32_t e = -1, node = -1;
                    float off = 0.f, tie = 0.f;
                    if (e0 >= 0) {
                        if (!cinfo) {
                            e = e0;
                            node = enode[e];
                            off = eoff[e];
                            tie = etie[e];
                        } else {
                            const uint32_t mw = j == 0 ? cmw[ii] : mask16[(int64_t)j * a.n + p];
                            const bool l0 = (mw >> (tr & 15)) & 1u;
                            const CInfo c = cinfo[2 * (int64_t)e0 + (l0 ? 0 : 1)];
                            nodeOf[ti] = c.entry;
                            e = c.entry;
                            node = c.entry >= 0 ? c.node : -1;
                            off = c.off;
                            tie = c.tie;
                        }
                    }
                    if (node >= 0) left = side_of(cw[e * 16 + pq] + off, tie, a.keys[tr], node, p);
                }
                const uint64_t bal = __ballot(left != 0u);
                if (mrow == 0 && pv) mask16[(int64_t)j * a.n + p] = (uint16_t)(bal >> (16 * kqi));
            }
        }
    }
}

// Sparse levels, position-major: a wave reads 64 consecutive (tree, position) items of the act table,
// and its four 16-lane groups test the active ones four at a time: the member's row is gathered through
// perm and the 16 lanes split the dimensions; flags are written at the positions.  Needs dim % 4 == 0.
// cscan_prev (after an entry-parallel level): act holds 2 e + side of the position's entry e one level up,
// whose children are entries cscan_prev[e] (left) and cscan_prev[e] + 1.
__global__ __launch_bounds__(256) void k_lvl_margins_pos(LvlArgs a, const PInfo* __restrict__ pinfo, int n_trees,
                                                         const int32_t* const* act, const int32_t* const* perm,
                                                         const uint32_t* __restrict__ cscan_prev,
                                                         uint8_t* __restrict__ flags8) {
    const int lane = threadIdx.x & 63, gl = lane & 15, g = lane >> 4;
    const int64_t total = (int64_t)n_trees * a.n;
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t w0 = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64; w0 < total;
         w0 += nw * 64) {  // wave-uniform
        const int64_t item = w0 + lane;
        int32_t node = -1, t = 0, pos = 0;
        float off = 0.f, tie = 0.f;
        if (item < total) {
            t = (int32_t)(item / a.n);
            pos = (int32_t)(item - (int64_t)t * a.n);
            int32_t e = act[t][pos];
            if (e >= 0 && cscan_prev) e = (int32_t)cscan_prev[e >> 1] + (e & 1);  // parent entry, side
            if (e >= 0) {
                const PInfo pi = pinfo[e];
                node = pi.node;
                off = pi.off;
                tie = pi.tie;
            }
        }
        uint64_t m = __ballot(node >= 0);
        while (m) {  // wave-uniform
            uint64_t mg = m;
            for (int k = 0; k < g && mg; ++k) mg &= mg - 1;  // group g takes the g-th remaining item
            const int src = mg ? __ffsll((unsigned long long)mg) - 1 : -1;
            for (int k = 0; k < 4 && m; ++k) m &= m - 1;
            const int sl = src < 0 ? 0 : src;
            const int32_t nd = __shfl(node, sl), tt = __shfl(t, sl), ps = __shfl(pos, sl);
            const float of = __shfl(off, sl), ti = __shfl(tie, sl);
            float acc = 0.f;
            int32_t pt = 0;
            if (src >= 0) {
                pt = perm[tt][ps];
                const float* x = a.X + (int64_t)pt * a.ld;
                const int8_t* hq = a.hyper_q[tt] + (int64_t)nd * a.dim;
                for (int d0 = 4 * gl; d0 < a.dim; d0 += 64) {
                    const float4 xv = *reinterpret_cast<const float4*>(x + d0);
                    const uint32_t hv = *reinterpret_cast<const uint32_t*>(hq + d0);
                    acc = fmaf(xv.x, q8(hv, 0), acc);
                    acc = fmaf(xv.y, q8(hv, 1), acc);
                    acc = fmaf(xv.z, q8(hv, 2), acc);
                    acc = fmaf(xv.w, q8(hv, 3), acc);
                }
            }
            for (int o = 8; o > 0; o >>= 1) acc += __shfl_xor(acc, o);  // inside the 16-lane group
            if (src >= 0 && gl == 0)
                flags8[(int64_t)tt * a.n + ps] = (uint8_t)side_of(acc + of, ti, a.keys[tt], nd, pt);
        }
    }
}

constexpr int kChunk = 256;  // items per flag chunk: one wave, four consecutive items per lane

// The four items item0 .. item0 + 3 of a lane as (tree, position); VEC4 (n % 4 == 0): one tree, and the
// per-tree arrays are 16-byte aligned at pos.
struct Item4 {
    int t[4];
    int64_t pos[4];
    bool ok[4];
};
template <bool VEC4>
__device__ __forceinline__ Item4 item4(int64_t item0, int64_t n, int64_t total) {
    Item4 r;
    if (VEC4) {
        const int t = (int)(item0 / n);
        const int64_t p = item0 - (int64_t)t * n;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            r.t[k] = t;
            r.pos[k] = p + k;
            r.ok[k] = item0 < total;
        }
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int64_t it = item0 + k;
            r.ok[k] = it < total;
            r.t[k] = r.ok[k] ? (int)(it / n) : 0;
            r.pos[k] = r.ok[k] ? it - (int64_t)r.t[k] * n : 0;
        }
    }
    return r;
}

// Left flags (bytes 0/1, one per (tree, position) item) of one level and their count per chunk.  Dense
// levels gather each item's side from the mask word of the point at that position; sparse levels find the
// flags already written by the side tests (mask16 null).  Flags of positions outside split nodes are
// stale, which is harmless: every prefix difference that is used spans one split node's items only.
template <bool VEC4>
__global__ __launch_bounds__(256) void k_lvl_flag_chunks(int n_trees, int64_t n, const uint16_t* __restrict__ mask16,
                                                         const int32_t* const* perm, uint8_t* flags8,
                                                         uint32_t* chunk_cnt) {
    const int64_t total = (int64_t)n_trees * n;
    const int64_t nchunks = ceil_div(total, (int64_t)kChunk);
    const int lane = threadIdx.x & 63;
    if (blockIdx.x == 0 && threadIdx.x == 0) chunk_cnt[nchunks] = 0;  // the scan's total
    for (int64_t c = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); c < nchunks; c += (int64_t)gridDim.x * 4) {
        const int64_t item0 = c * kChunk + 4 * lane;
        uint32_t f[4] = {0u, 0u, 0u, 0u};
        if (mask16) {
            const Item4 it = item4<VEC4>(item0, n, total);
            int32_t pt[4] = {0, 0, 0, 0};
            if (VEC4) {
                if (it.ok[0]) {
                    const int4 v = *reinterpret_cast<const int4*>(perm[it.t[0]] + it.pos[0]);
                    pt[0] = v.x, pt[1] = v.y, pt[2] = v.z, pt[3] = v.w;
                }
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (it.ok[k]) pt[k] = perm[it.t[k]][it.pos[k]];
            }
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (it.ok[k]) f[k] = (mask16[(int64_t)(it.t[k] >> 4) * n + pt[k]] >> (it.t[k] & 15)) & 1u;
            if (item0 + 3 < total) {
                *reinterpret_cast<uint32_t*>(flags8 + item0) = f[0] | f[1] << 8 | f[2] << 16 | f[3] << 24;
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (item0 + k < total) flags8[item0 + k] = (uint8_t)f[k];
            }
        } else if (item0 + 3 < total) {
            const uint32_t v = *reinterpret_cast<const uint32_t*>(flags8 + item0);
            f[0] = v & 0xff, f[1] = (v >> 8) & 0xff, f[2] = (v >> 16) & 0xff, f[3] = v >> 24;
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (item0 + k < total) f[k] = flags8[item0 + k];
        }
        uint32_t cnt = f[0] + f[1] + f[2] + f[3];
        for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
        if (lane == 0) chunk_cnt[c] = cnt;
    }
}

// Number of left flags among items [0, x): the chunk's offset plus the flags of the chunk before x.
__device__ __forceinline__ uint32_t lvl_prefix(int64_t x, const uint32_t* __restrict__ chunk_off,
                                               const uint8_t* __restrict__ flags8) {
    const int64_t c = x / kChunk;
    const int r = (int)(x - c * kChunk);
    uint32_t p = chunk_off[c];
    const uint32_t* w = reinterpret_cast<const uint32_t*>(flags8 + c * kChunk);
    int k = 0;
    for (; k + 4 <= (r >> 2); k += 4) {  // bytes are 0 or 1: one set bit per left flag
        const uint4 v = *reinterpret_cast<const uint4*>(w + k);
        p += __popc(v.x) + __popc(v.y) + __popc(v.z) + __popc(v.w);
    }
    for (; k < (r >> 2); ++k) p += __popc(w[k]);
    if (r & 3) p += __popc(w[r >> 2] & ((1u << (8 * (r & 3))) - 1u));
    return p;
}

__global__ void k_lvl_counts(LvlArgs a, const LvlNode* nodes, int64_t na, const uint32_t* chunk_off,
                             const uint8_t* flags8, uint8_t* split, uint32_t* nchild, uint32_t* leftsz,
                             uint32_t* lstart) {
    if (blockIdx.x == 0 && threadIdx.x == 0) nchild[na] = 0;  // the scan's total
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < na; e += (int64_t)gridDim.x * blockDim.x) {
        const LvlNode nd = nodes[e];
        uint32_t c = 0;
        if (split[e]) {
            const int64_t off = (int64_t)nd.tree * a.n;
            const uint32_t p0 = lvl_prefix(off + nd.start, chunk_off, flags8);
            const uint32_t L = lvl_prefix(off + nd.end, chunk_off, flags8) - p0;
            const uint32_t m = (uint32_t)(nd.end - nd.start);
            leftsz[e] = L;
            lstart[e] = p0;
            if (L == 0 || L == m) {  // degenerate split -> leaf with a zero hyperplane (builder.rs:198-210)
                split[e] = 0;
                a.children[nd.tree][2 * nd.node] = -nd.start;
                a.children[nd.tree][2 * nd.node + 1] = -nd.end;
                float* hp = a.hyper[nd.tree] + (int64_t)nd.node * a.dim;
                for (int i = 0; i < a.dim; ++i) hp[i] = 0.f;
                a.offsets[nd.tree][nd.node] = 0.f;
            } else {
                c = 2;
            }
        }
        nchild[e] = c;
    }
}

// Per tree: first child id of the next level (the tree's node count so far) and the exclusive prefix of
// the children counts over the trees (the active list is grouped by tree).
// Tree t's entries of the active list are [start[t], start[t+1]) (the list is grouped by tree), so its
// children count is a difference of the children scan - no atomics in k_lvl_counts.  cum (the next list's
// start per tree, with the total at n_trees) becomes the next level's start.
__global__ void k_lvl_tree_base(int n_trees, const uint32_t* start, const uint32_t* cscan, unsigned* tree_children,
                                int32_t* node_count, int32_t* base, uint32_t* cum) {
    if (threadIdx.x >= 64 || blockIdx.x != 0) return;  // one wave: a lane per tree, 64 trees per step
    const int lane = threadIdx.x;
    uint32_t carry = 0;
    for (int t0 = 0; t0 < n_trees; t0 += 64) {
        const int t = t0 + lane;
        const uint32_t c = t < n_trees ? cscan[start[t + 1]] - cscan[start[t]] : 0u;
        uint32_t incl = c;  // inclusive scan of the children counts over the lanes
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t v = __shfl_up(incl, o);
            if (lane >= o) incl += v;
        }
        if (t < n_trees) {
            tree_children[t] = c;
            base[t] = node_count[t];
            cum[t] = carry + incl - c;
            node_count[t] += (int32_t)c;
        }
        carry += __shfl(incl, 63);
    }
    if (lane == 0) cum[n_trees] = carry;
}

// An entry that does not split gets its two CInfo marked (no child entry); the children of the others
// fill theirs in the next level's k_lvl_planes.
__global__ void k_lvl_assign(const LvlNode* nodes, int64_t na, const uint8_t* split, const uint32_t* cscan,
                             const uint32_t* leftsz, const int32_t* tree_base_id, const uint32_t* tree_cum,
                             int32_t** children, LvlNode* next, CInfo* cinfo) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < na; e += (int64_t)gridDim.x * blockDim.x) {
        if (!split[e]) {
            cinfo[2 * e] = CInfo{-1, 0.f, -1, 0.f};
            cinfo[2 * e + 1] = CInfo{-1, 0.f, -1, 0.f};
            continue;
        }
        const LvlNode nd = nodes[e];
        const uint32_t s = cscan[e];  // the active list stays grouped by tree, so s - tree_cum is per tree
        const int32_t left = tree_base_id[nd.tree] + (int32_t)(s - tree_cum[nd.tree]);
        children[nd.tree][2 * nd.node] = left;
        children[nd.tree][2 * nd.node + 1] = left + 1;
        const int32_t mid = nd.start + (int32_t)leftsz[e];
        next[s] = LvlNode{nd.tree, left, nd.start, mid, nd.depth + 1, (int32_t)e};
        next[s + 1] = LvlNode{nd.tree, left + 1, mid, nd.end, nd.depth + 1, (int32_t)e};
    }
}

// act values: >= 0 active entry; -2 became a leaf one level ago (the perm buffers already agree, the
// act buffers are synced now); -1 settled in both buffers (never touched again).
// One wave per chunk, four consecutive items per lane: the prefix of an item is its chunk's offset plus a
// wave scan of the lanes' flag counts; lstart holds the prefix at each split node's first item.
template <bool VEC4>
__global__ __launch_bounds__(256) void k_lvl_scatter(const LvlNode* nodes, int n_trees, int64_t n, const uint8_t* split,
                                                     const uint32_t* cscan, const uint32_t* leftsz,
                                                     const uint32_t* lstart, const uint32_t* chunk_off,
                                                     const uint8_t* flags8, const int32_t* const* perm,
                                                     int32_t* const* perm2, int32_t* const* act,
                                                     int32_t* const* act2) {
    const int64_t total = (int64_t)n_trees * n;
    const int64_t nchunks = ceil_div(total, (int64_t)kChunk);
    const int lane = threadIdx.x & 63;
    for (int64_t c = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); c < nchunks; c += (int64_t)gridDim.x * 4) {
        const int64_t item0 = c * kChunk + 4 * lane;
        const Item4 it = item4<VEC4>(item0, n, total);
        uint32_t f[4] = {0u, 0u, 0u, 0u};
        int32_t ev[4] = {-1, -1, -1, -1}, pv[4] = {0, 0, 0, 0};
        if (VEC4) {
            if (it.ok[0]) {
                const uint32_t fv = *reinterpret_cast<const uint32_t*>(flags8 + item0);
                const int4 av = *reinterpret_cast<const int4*>(act[it.t[0]] + it.pos[0]);
                const int4 p4 = *reinterpret_cast<const int4*>(perm[it.t[0]] + it.pos[0]);
                f[0] = fv & 0xff, f[1] = (fv >> 8) & 0xff, f[2] = (fv >> 16) & 0xff, f[3] = fv >> 24;
                ev[0] = av.x, ev[1] = av.y, ev[2] = av.z, ev[3] = av.w;
                pv[0] = p4.x, pv[1] = p4.y, pv[2] = p4.z, pv[3] = p4.w;
            }
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (it.ok[k]) {
                    f[k] = flags8[item0 + k];
                    ev[k] = act[it.t[k]][it.pos[k]];
                    pv[k] = perm[it.t[k]][it.pos[k]];
                }
        }
        const uint32_t cnt = f[0] + f[1] + f[2] + f[3];
        uint32_t incl = cnt;  // inclusive scan over the wave's lanes
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t v = __shfl_up(incl, o);
            if (lane >= o) incl += v;
        }
        uint32_t before = chunk_off[c] + incl - cnt;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int32_t e = ev[k];
            const int t = it.t[k];
            const int64_t pos = it.pos[k];
            const uint32_t fk = f[k];
            const uint32_t b = before;
            before += fk;
            if (e == -1) continue;  // also the items past total
            if (e == -2) {
                act[t][pos] = -1;
                act2[t][pos] = -1;
                continue;
            }
            if (!split[e]) {  // became a leaf at this level
                perm2[t][pos] = pv[k];
                act2[t][pos] = -2;
                continue;
            }
            const LvlNode nd = nodes[e];
            const uint32_t lr = b - lstart[e];  // left members before pos
            const int64_t np = fk ? nd.start + lr : nd.start + leftsz[e] + (pos - nd.start - lr);
            perm2[t][np] = pv[k];
            act2[t][np] = (int32_t)cscan[e] + (fk ? 0 : 1);
        }
    }
}

// ---- entry-parallel partition (the levels with many active entries) -------------------------------------
// One workgroup per active entry: the entry's left flags (dense levels: the bit the side tests left in each
// member's mask word; sparse levels: the flags k_lvl_margins_pos wrote at the positions), their count (a
// degenerate split becomes a leaf, as in k_lvl_counts), then the stable partition of the entry's positions
// into the other buffer - the same order k_lvl_scatter gives, without the flag chunks or their scan.  Moved
// positions get 2 e + side in the other act buffer (their next entry is cscan[e] + side, known once the
// children are scanned); an entry that does not split copies its positions and marks them settled (-1) in
// both act buffers.  Entries of at most kEntryStage members keep their ids and flags in LDS between the
// passes; larger ones read the flags from flags8 (dense levels write them there first).
constexpr int kEntryStage = 4096;
constexpr int kWaveStage = 1024;  // entries of at most this many members: one wave each (k_lvl_part_wave)

__device__ __forceinline__ uint32_t block_sum_u32(uint32_t v, uint32_t* red) {  // 256 threads
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
}
// exclusive prefix of v over the 256 threads in thread order; *total gets the sum
__device__ __forceinline__ uint32_t block_excl_u32(uint32_t v, uint32_t* red, uint32_t* total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t incl = v;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(incl, o);
        if (lane >= o) incl += u;
    }
    __syncthreads();
    if (lane == 63) red[4 + w] = incl;
    __syncthreads();
    uint32_t before = 0;
    for (int i = 0; i < w; ++i) before += red[4 + i];
    *total = red[4] + red[5] + red[6] + red[7];
    return before + incl - v;
}

template <bool FLAGS>
__global__ __launch_bounds__(256) void k_lvl_part_entry(LvlArgs a, const LvlNode* __restrict__ nodes, int64_t na,
                                                        const int32_t* const* perm, int32_t* const* perm2,
                                                        int32_t* const* act, int32_t* const* act2,
                                                        const uint16_t* __restrict__ mask16, uint8_t* split,
                                                        uint32_t* nchild, uint32_t* leftsz, uint8_t* flags8) {
    __shared__ int32_t s_pt[kEntryStage];
    __shared__ uint8_t s_side[kEntryStage];
    __shared__ uint32_t red[8];
    const int tid = threadIdx.x;
    if (blockIdx.x == 0 && tid == 0) nchild[na] = 0;  // the scan's total
    for (int64_t e = blockIdx.x; e < na; e += gridDim.x) {
        const LvlNode nd = nodes[e];
        const int t = nd.tree;
        const int m = nd.end - nd.start;
        if (m <= kWaveStage) continue;  // uniform: k_lvl_part_wave's
        const int32_t* P = perm[t] + nd.start;
        int32_t* P2 = perm2[t] + nd.start;
        int32_t* A = act[t] + nd.start;
        int32_t* A2 = act2[t] + nd.start;
        if (!split[e]) {  // uniform: a leaf since k_lvl_planes; settled in both buffers
            for (int i = tid; i < m; i += 256) {
                P2[i] = P[i];
                A[i] = -1;
                A2[i] = -1;
            }
            if (tid == 0) nchild[e] = 0;
            continue;
        }
        const bool staged = m <= kEntryStage;
        uint8_t* F = flags8 + (int64_t)t * a.n + nd.start;
        uint32_t cnt = 0;
        const uint16_t* mk = mask16 + (int64_t)(t >> 4) * a.n;
        const int sh = t & 15;
        for (int i0 = tid; i0 < m; i0 += 1024) {  // four members per thread in flight
            int32_t pt[4];
            uint8_t sd[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int i = i0 + 256 * k;
                pt[k] = i < m ? P[i] : 0;
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int i = i0 + 256 * k;
                sd[k] = i < m ? (FLAGS ? F[i] : (uint8_t)((mk[pt[k]] >> sh) & 1u)) : (uint8_t)0;
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int i = i0 + 256 * k;
                if (i < m) {
                    if (staged) {
                        s_pt[i] = pt[k];
                        s_side[i] = sd[k];
                    } else if (!FLAGS) {
                        F[i] = sd[k];
                    }
                }
                cnt += sd[k];
            }
        }
        const uint32_t L = block_sum_u32(cnt, red);  // also orders pass 1's writes before pass 2's reads
        if (L == 0 || L == (uint32_t)m) {  // uniform: degenerate split -> leaf with a zero hyperplane
            if (tid == 0) {
                split[e] = 0;
                nchild[e] = 0;
                a.children[t][2 * nd.node] = -nd.start;
                a.children[t][2 * nd.node + 1] = -nd.end;
                a.offsets[t][nd.node] = 0.f;
            }
            float* hp = a.hyper[t] + (int64_t)nd.node * a.dim;
            for (int i = tid; i < a.dim; i += 256) hp[i] = 0.f;
            for (int i = tid; i < m; i += 256) {
                P2[i] = P[i];
                A[i] = -1;
                A2[i] = -1;
            }
            __syncthreads();  // the staged arrays are rewritten by the next entry
            continue;
        }
        if (tid == 0) {
            nchild[e] = 2;
            leftsz[e] = L;
        }
        const int32_t code = (int32_t)(2 * e);  // the next level's entry: 2 e + side (left 0, right 1)
        uint32_t lb = 0, rb = 0;  // left / right members placed by the earlier tiles
        for (int base = 0; base < m; base += 1024) {
            int32_t pv[4];
            uint32_t f[4];
            const int i0 = base + 4 * tid;
            uint32_t c = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int i = i0 + k;
                f[k] = 0u;
                pv[k] = 0;
                if (i < m) {
                    pv[k] = staged ? s_pt[i] : P[i];
                    f[k] = staged ? s_side[i] : F[i];
                }
                c += f[k];
            }
            uint32_t tileL = 0;
            const uint32_t lx = block_excl_u32(c, red, &tileL);  // lefts of this tile before this thread
            const int valid_before = max(0, min(4 * tid, m - base));
            uint32_t lp = lb + lx, rp = rb + (uint32_t)valid_before - lx;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (i0 + k >= m) break;
                if (f[k]) {
                    P2[lp] = pv[k];
                    A2[lp++] = code;
                } else {
                    P2[L + rp] = pv[k];
                    A2[L + rp++] = code + 1;
                }
            }
            const uint32_t tile_items = (uint32_t)min(1024, m - base);
            lb += tileL;
            rb += tile_items - tileL;
        }
        __syncthreads();  // the staged arrays are rewritten by the next entry
    }
}

// The same for the entries of at most kWaveStage members, one wave each (four per workgroup): no workgroup
// barriers; the stable partition takes 64 members a step, ranks from ballots.
template <bool FLAGS>
__global__ __launch_bounds__(256) void k_lvl_part_wave(LvlArgs a, const LvlNode* __restrict__ nodes, int64_t na,
                                                       const int32_t* const* perm, int32_t* const* perm2,
                                                       int32_t* const* act, int32_t* const* act2,
                                                       const uint16_t* __restrict__ mask16,
                                                       const uint8_t* __restrict__ flags8, uint8_t* split,
                                                       uint32_t* nchild, uint32_t* leftsz) {
    __shared__ int32_t s_pt[4][kWaveStage];
    __shared__ uint8_t s_side[4][kWaveStage];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int32_t* const spt = s_pt[w];
    uint8_t* const ssd = s_side[w];
    if (blockIdx.x == 0 && threadIdx.x == 0) nchild[na] = 0;  // the scan's total
    for (int64_t e = (int64_t)blockIdx.x * 4 + w; e < na; e += (int64_t)gridDim.x * 4) {  // wave-uniform
        const LvlNode nd = nodes[e];
        const int t = nd.tree;
        const int m = nd.end - nd.start;
        if (m > kWaveStage) continue;  // k_lvl_part_entry's
        const int32_t* P = perm[t] + nd.start;
        int32_t* P2 = perm2[t] + nd.start;
        int32_t* A = act[t] + nd.start;
        int32_t* A2 = act2[t] + nd.start;
        if (!split[e]) {  // a leaf since k_lvl_planes; settled in both buffers
            for (int i = lane; i < m; i += 64) {
                P2[i] = P[i];
                A[i] = -1;
                A2[i] = -1;
            }
            if (lane == 0) nchild[e] = 0;
            continue;
        }
        const uint8_t* F = flags8 + (int64_t)t * a.n + nd.start;
        const uint16_t* mk = mask16 + (int64_t)(t >> 4) * a.n;
        const int sh = t & 15;
        uint32_t cnt = 0;
        for (int i = lane; i < m; i += 64) {
            const int32_t pt = P[i];
            const uint8_t sd = FLAGS ? F[i] : (uint8_t)((mk[pt] >> sh) & 1u);
            spt[i] = pt;
            ssd[i] = sd;
            cnt += sd;
        }
        for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
        const uint32_t L = cnt;
        __threadfence_block();
        wave_sync();
        if (L == 0 || L == (uint32_t)m) {  // wave-uniform: degenerate split -> leaf with a zero hyperplane
            if (lane == 0) {
                split[e] = 0;
                nchild[e] = 0;
                a.children[t][2 * nd.node] = -nd.start;
                a.children[t][2 * nd.node + 1] = -nd.end;
                a.offsets[t][nd.node] = 0.f;
            }
            float* hp = a.hyper[t] + (int64_t)nd.node * a.dim;
            for (int i = lane; i < a.dim; i += 64) hp[i] = 0.f;
            for (int i = lane; i < m; i += 64) {
                P2[i] = spt[i];
                A[i] = -1;
                A2[i] = -1;
            }
            wave_sync();
            continue;
        }
        if (lane == 0) {
            nchild[e] = 2;
            leftsz[e] = L;
        }
        const int32_t code = (int32_t)(2 * e);  // the next level's entry: 2 e + side (left 0, right 1)
        const uint64_t below = (1ull << lane) - 1ull;
        uint32_t lb = 0, rb = 0;
        for (int i0 = 0; i0 < m; i0 += 64) {
            const int i = i0 + lane;
            const bool ok = i < m;
            const uint32_t sd = ok ? ssd[i] : 0u;
            const int32_t pt = ok ? spt[i] : 0;
            const uint64_t lm = __ballot(ok && sd), rm = __ballot(ok && !sd);
            if (ok) {
                if (sd) {
                    const uint32_t d = lb + (uint32_t)__popcll(lm & below);
                    P2[d] = pt;
                    A2[d] = code;
                } else {
                    const uint32_t d = L + rb + (uint32_t)__popcll(rm & below);
                    P2[d] = pt;
                    A2[d] = code + 1;
                }
            }
            lb += (uint32_t)__popcll(lm);
            rb += (uint32_t)__popcll(rm);
        }
        wave_sync();  // the staged arrays are rewritten by the next entry
    }
}

// The first entry-parallel level: positions that became leaves one level earlier (act -2, see k_lvl_scatter)
// are settled in both buffers, since the entry-parallel levels never visit them.
__global__ void k_lvl_act_settle(int n_trees, int64_t n, int32_t* const* act, int32_t* const* act2) {
    const int64_t total = (int64_t)n_trees * n;
    for (int64_t item = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; item < total;
         item += (int64_t)gridDim.x * blockDim.x) {
        const int t = (int)(item / n);
        const int64_t pos = item - (int64_t)t * n;
        if (act[t][pos] == -2) {
            act[t][pos] = -1;
            act2[t][pos] = -1;
        }
    }
}

__global__ void k_lvl_init(int n_trees, int64_t n, int32_t* const* perm, int32_t* const* act, int32_t* nodeOf) {
    const int64_t total = (int64_t)n_trees * n;
    for (int64_t item = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; item < total;
         item += (int64_t)gridDim.x * blockDim.x) {
        const int t = (int)(item / n);
        const int64_t pos = item - (int64_t)t * n;
        perm[t][pos] = (int32_t)pos;
        act[t][pos] = t;  // level 0: the root of tree t is active entry t
        nodeOf[pos * n_trees + t] = t;
    }
}

static void launch_margins_pm(hipStream_t s, const LvlArgs& a, const PInfo* pinfo, int T, const CInfo* cinfo,
                              int32_t* nodeOf, uint16_t* mask16) {
    const unsigned grid = lgrid(a.n * 16, 256, 1 << 16);
    if (a.dim <= 128)
        hipLaunchKernelGGL(k_lvl_margins_pm<1>, dim3(grid), dim3(256), 0, s, a, pinfo, T, cinfo, nodeOf, mask16);
    else if (a.dim <= 256)
        hipLaunchKernelGGL(k_lvl_margins_pm<2>, dim3(grid), dim3(256), 0, s, a, pinfo, T, cinfo, nodeOf, mask16);
    else if (a.dim <= 512)
        hipLaunchKernelGGL(k_lvl_margins_pm<4>, dim3(grid), dim3(256), 0, s, a, pinfo, T, cinfo, nodeOf, mask16);
    else
        hipLaunchKernelGGL(k_lvl_margins_pm<8>, dim3(grid), dim3(256), 0, s, a, pinfo, T, cinfo, nodeOf, mask16);
}

constexpr int kMfmaEntries = 128;  // the MFMA side tests stage at most this many planes per workgroup
static bool mfma_levels_ok(int dim, int T) { return dim % 32 == 0 && dim <= 256 && T <= 64; }
static void launch_margins_mfma(hipStream_t s, const LvlArgs& a, const LvlNode* cur, int na, const PInfo* pinfo, int T,
                                const CInfo* cinfo, int32_t* nodeOf, uint16_t* mask16) {
    const int na16 = (na + 15) & ~15;
    const size_t lds = (size_t)na16 * (a.dim + 8) * 2 + (size_t)na16 * 12 + 4 * (size_t)na16 * 16 * 4;
    const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(4, (160 * 1024) / lds));
    const int64_t nblk = (a.n + 15) >> 4;
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(nblk, (int64_t)4),
                                                                           (int64_t)device_cus() * per_cu));
    switch (a.dim / 32) {
        case 1: hipLaunchKernelGGL(k_lvl_margins_mfma<1>, dim3(grid), dim3(256), lds, s, a, cur, na, pinfo, T, cinfo, nodeOf, mask16); break;
        case 2: hipLaunchKernelGGL(k_lvl_margins_mfma<2>, dim3(grid), dim3(256), lds, s, a, cur, na, pinfo, T, cinfo, nodeOf, mask16); break;
        case 3: hipLaunchKernelGGL(k_lvl_margins_mfma<3>, dim3(grid), dim3(256), lds, s, a, cur, na, pinfo, T, cinfo, nodeOf, mask16); break;
        case 4: hipLaunchKernelGGL(k_lvl_margins_mfma<4>, dim3(grid), dim3(256), lds, s, a, cur, na, pinfo, T, cinfo, nodeOf, mask16); break;
        case 5: hipLaunchKernelGGL(k_lvl_margins_mfma<5>, dim3(grid), dim3(256), lds, s, a, cur, na, pinfo, T, cinfo, nodeOf, mask16); break;
        case 6: hipLaunchKernelGGL(k_lvl_margins_mfma<6>, dim3(grid), dim3(256), lds, s, a, cur, na, pinfo, T, cinfo, nodeOf, mask16); break;
        case 7: hipLaunchKernelGGL(k_lvl_margins_mfma<7>, dim3(grid), dim3(256), lds, s, a, cur, na, pinfo, T, cinfo, nodeOf, mask16); break;
        default: hipLaunchKernelGGL(k_lvl_margins_mfma<8>, dim3(grid), dim3(256), lds, s, a, cur, na, pinfo, T, cinfo, nodeOf, mask16); break;
    }
}

bool build_forest_level(hipStream_t s, const float* X, int64_t ld, int dim, int64_t n, int64_t leaf_size,
                        int64_t max_depth, int angular, const std::vector<uint64_t>& keys, rinnd_forest& F,
                        const KTimer& margins_timer, Workspace& W) {
    const int T = (int)F.trees.size();
    const int64_t total = (int64_t)T * n;
    if (T == 0) return true;
    RINND_CHECK_ARG(total < (int64_t)INT32_MAX - 1, "n_trees x n must stay below 2^31 for the level forest builder");
    RINND_CHECK_ARG(T <= 64, "the level forest builder handles at most 64 trees per call");
    const int64_t cap = F.trees[0].cap_nodes;
    std::vector<int32_t*> perm_h(T), perm2_h(T), act_h(T), act2_h(T), ch_h(T);
    std::vector<float*> hp_h(T), of_h(T);
    WBuf<int32_t> perm2(W, kWsForest + 0, total), act(W, kWsForest + 1, total), act2(W, kWsForest + 2, total);
    WBuf<int32_t> nodeOf(W, kWsForest + 3, total);  // dense levels: each (point, tree)'s entry, point-major
    const int NB = (T + 15) / 16;                      // mask words per point
    WBuf<uint16_t> mask16(W, kWsForest + 27, (size_t)NB * n);
    WBuf<int8_t> hyper_q(W, kWsForest + 4, (size_t)T * cap * dim);
    for (int t = 0; t < T; ++t) {
        perm_h[t] = F.trees[t].indices.p;
        perm2_h[t] = perm2.p + (int64_t)t * n;
        act_h[t] = act.p + (int64_t)t * n;
        act2_h[t] = act2.p + (int64_t)t * n;
        ch_h[t] = F.trees[t].children.p;
        hp_h[t] = F.trees[t].hyper.p;
        of_h[t] = F.trees[t].offsets.p;
    }
    // pointer tables [perm, perm2, act, act2] for the two ping-pong phases
    std::vector<int32_t*> tab(8 * (size_t)T);
    for (int t = 0; t < T; ++t) {
        tab[0 * T + t] = perm_h[t];
        tab[1 * T + t] = perm2_h[t];
        tab[2 * T + t] = act_h[t];
        tab[3 * T + t] = act2_h[t];
        tab[4 * T + t] = perm2_h[t];
        tab[5 * T + t] = perm_h[t];
        tab[6 * T + t] = act2_h[t];
        tab[7 * T + t] = act_h[t];
    }
    WBuf<int32_t*> d_tab(W, kWsForest + 5, tab.size()), d_ch(W, kWsForest + 6, T);
    WBuf<float*> d_hp(W, kWsForest + 7, T), d_of(W, kWsForest + 8, T);
    WBuf<uint64_t> d_keys(W, kWsForest + 9, T);
    HIP_TRY(hipMemcpyAsync(d_tab.p, tab.data(), tab.size() * sizeof(int32_t*), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_ch.p, ch_h.data(), T * sizeof(int32_t*), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_hp.p, hp_h.data(), T * sizeof(float*), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_of.p, of_h.data(), T * sizeof(float*), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_keys.p, keys.data(), T * 8, hipMemcpyHostToDevice, s));
    std::vector<int8_t*> hb_h(T);
    for (int t = 0; t < T; ++t) hb_h[t] = hyper_q.p + (size_t)t * cap * dim;
    WBuf<int8_t*> d_hb(W, kWsForest + 10, T);
    HIP_TRY(hipMemcpyAsync(d_hb.p, hb_h.data(), T * sizeof(int8_t*), hipMemcpyHostToDevice, s));
    const LvlArgs a{X, ld, dim, n, std::max<int64_t>(leaf_size, 1), max_depth, angular,
                    d_keys.p, d_hp.p, d_of.p, d_ch.p, d_hb.p, T, hyper_q.p, cap};

    const int64_t nchunks = ceil_div(total, (int64_t)kChunk);
    WBuf<uint8_t> flags8(W, kWsForest + 11, (size_t)nchunks * kChunk);
    // four consecutive items per lane in one tree, 16-byte aligned in every per-tree array
    bool vec4 = n % 4 == 0 && ((uintptr_t)perm2.p & 15) == 0 && ((uintptr_t)act.p & 15) == 0 &&
                ((uintptr_t)act2.p & 15) == 0;
    for (int t = 0; t < T; ++t) vec4 = vec4 && ((uintptr_t)perm_h[t] & 15) == 0;
    WBuf<uint32_t> chunk_cnt(W, kWsForest + 12, nchunks + 1), chunk_off(W, kWsForest + 29, nchunks + 1);
    size_t scan_bytes = 0;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, chunk_cnt.p, chunk_off.p, (int)(nchunks + 1)));
    WBuf<char> scan_tmp(W, kWsForest + 13, std::max<size_t>(scan_bytes, 16));
    // a level holds at most n nodes per tree (every active node has at least one member)
    const int64_t max_nodes = std::max<int64_t>(1, std::min<int64_t>(cap, n)) * (int64_t)T;
    WBuf<LvlNode> curb(W, kWsForest + 14, max_nodes + 1), nxtb(W, kWsForest + 15, max_nodes + 1);
    WBuf<uint8_t> split(W, kWsForest + 16, max_nodes + 1);
    WBuf<PInfo> pinfo(W, kWsForest + 25, max_nodes + 1);
    WBuf<uint32_t> nchild(W, kWsForest + 17, max_nodes + 1), cscan(W, kWsForest + 18, max_nodes + 1),
        leftsz(W, kWsForest + 19, max_nodes + 1);
    WBuf<CInfo> cinfo(W, kWsForest + 28, 2 * (size_t)(max_nodes + 1));
    WBuf<uint32_t> lstart(W, kWsForest + 30, max_nodes + 1);
    // level status read back once per level: [tested pairs so far (u64) | children per tree (u32 x T)]
    WBuf<unsigned long long> lstat(W, kWsForest + 21, 1 + (T + 1) / 2);
    unsigned long long* active_pairs = lstat.p;
    unsigned* tree_children = (unsigned*)(lstat.p + 1);
    HIP_TRY(hipMemsetAsync(active_pairs, 0, 8, s));
    LvlNode* cur = curb.p;
    LvlNode* nxt = nxtb.p;
    unsigned long long pairs_before = 0;
    WBuf<int32_t> d_base(W, kWsForest + 22, T), d_count(W, kWsForest + 20, T);
    WBuf<uint32_t> d_cum(W, kWsForest + 23, T + 1), d_start(W, kWsForest + 26, T + 1);
    size_t cscan_bytes = 0;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, cscan_bytes, nchild.p, cscan.p, (int)(max_nodes + 1)));
    WBuf<char> cscan_tmp(W, kWsForest + 24, std::max<size_t>(cscan_bytes, 16));

    std::vector<LvlNode> roots(T);
    for (int t = 0; t < T; ++t) roots[t] = LvlNode{t, 0, 0, (int32_t)n, 0, -1};
    HIP_TRY(hipMemcpyAsync(cur, roots.data(), T * sizeof(LvlNode), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_lvl_init, dim3(lgrid(total, 256)), dim3(256), 0, s, T, n, d_tab.p, d_tab.p + 2 * T,
                       nodeOf.p);
    std::vector<int64_t> node_count(T, 1);
    {
        const std::vector<int32_t> ones(T, 1);
        std::vector<uint32_t> first(T + 1);  // level 0: root of tree t is entry t
        for (int t = 0; t <= T; ++t) first[t] = (uint32_t)t;
        HIP_TRY(hipMemcpyAsync(d_count.p, ones.data(), T * 4, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(d_start.p, first.data(), (T + 1) * 4, hipMemcpyHostToDevice, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    // the level status is copied to pinned memory and read while assign and scatter run
    const size_t lst_bytes = (1 + (T + 1) / 2) * sizeof(unsigned long long);
    unsigned long long* lst = (unsigned long long*)W.status.ensure(lst_bytes);
    const hipEvent_t lst_ev = W.status_ev.get();
    int level = 0;
    uint32_t* st_cur = d_start.p;  // first active entry of each tree (and the total), this level
    uint32_t* st_next = d_cum.p;   // the same for the next level
    int64_t na = T;
    int phase = 0;
    // Side tests: point-major (every point's row read once for all trees) while most (tree, point) pairs
    // are tested; position-major (rows gathered per tested pair, no point -> position table) once the
    // tested pairs of a level fall to kSparsePairs per point, which they never exceed again.
    constexpr int64_t kSparsePairs = 2;
    const bool point_major_dpp = dim % 8 == 0 && dim <= 1024;
    const bool mfma_levels = mfma_levels_ok(dim, T);  // shallow dense levels: planes in LDS, MFMA side tests
    bool sparse = false;
    // Entry-parallel partition (k_lvl_part_wave / k_lvl_part_entry) from the first level whose entries average
    // at most kEntryItems of the T x n positions on (below that the chunked scatter over all positions is
    // faster); act then holds each position's 2 e + side relative to the last level's entries.
    constexpr int64_t kEntryItems = 1024;
    bool entry_mode = false;
    while (na > 0) {
        int32_t* const* P = d_tab.p + (size_t)phase * 4 * T;  // perm, perm2, act, act2
        hipLaunchKernelGGL(k_lvl_planes, dim3(lgrid(na, 1, 4096)), dim3(256), 0, s, a, cur, na,
                           (const int32_t* const*)P, split.p, pinfo.p, cinfo.p, active_pairs);
        const bool prev_entry = entry_mode;  // act holds 2 e + side codes of the last level's entries
        entry_mode = entry_mode || na * kEntryItems >= total;
        if (entry_mode && !prev_entry)
            hipLaunchKernelGGL(k_lvl_act_settle, dim3(lgrid(total, 256, 1 << 16)), dim3(256), 0, s, T, n, P + 2 * T,
                               P + 3 * T);
        margins_timer.begin(s);
        const unsigned mg = lgrid(n * 16, 256, 1 << 16);
        const CInfo* cinfo_prev = level > 0 ? cinfo.p : nullptr;
        if (sparse)
            hipLaunchKernelGGL(k_lvl_margins_pos, dim3(lgrid(total, 256, 1 << 16)), dim3(256), 0, s, a, pinfo.p, T,
                               (const int32_t* const*)(P + 2 * T), (const int32_t* const*)P,
                               prev_entry ? cscan.p : nullptr, flags8.p);
        else if (mfma_levels && na <= kMfmaEntries)
            launch_margins_mfma(s, a, cur, (int)na, pinfo.p, T, cinfo_prev, nodeOf.p, mask16.p);
        else if (point_major_dpp)
            launch_margins_pm(s, a, pinfo.p, T, cinfo_prev, nodeOf.p, mask16.p);
        else if (dim % 4 == 0 && dim <= 64)
            hipLaunchKernelGGL(k_lvl_margins<1>, dim3(mg), dim3(256), 0, s, a, pinfo.p, T, cinfo_prev, nodeOf.p,
                               mask16.p);
        else if (dim % 4 == 0 && dim <= 128)
            hipLaunchKernelGGL(k_lvl_margins<2>, dim3(mg), dim3(256), 0, s, a, pinfo.p, T, cinfo_prev, nodeOf.p,
                               mask16.p);
        else if (dim % 4 == 0 && dim <= 256)
            hipLaunchKernelGGL(k_lvl_margins<4>, dim3(mg), dim3(256), 0, s, a, pinfo.p, T, cinfo_prev, nodeOf.p,
                               mask16.p);
        else
            hipLaunchKernelGGL(k_lvl_margins<0>, dim3(mg), dim3(256), 0, s, a, pinfo.p, T, cinfo_prev, nodeOf.p,
                               mask16.p);
        margins_timer.stop(s);
        const unsigned cgrid = lgrid(nchunks, 4, 1 << 16);  // four chunks (waves) per workgroup
        if (entry_mode) {
            const unsigned wg = lgrid(na, 4, 1 << 16), eg = lgrid(na, 1, 2048);
            if (sparse) {  // flags at the positions, from k_lvl_margins_pos
                hipLaunchKernelGGL(k_lvl_part_wave<true>, dim3(wg), dim3(256), 0, s, a, cur, na,
                                   (const int32_t* const*)P, P + T, P + 2 * T, P + 3 * T, mask16.p, flags8.p, split.p,
                                   nchild.p, leftsz.p);
                hipLaunchKernelGGL(k_lvl_part_entry<true>, dim3(eg), dim3(256), 0, s, a, cur, na,
                                   (const int32_t* const*)P, P + T, P + 2 * T, P + 3 * T, mask16.p, split.p, nchild.p,
                                   leftsz.p, flags8.p);
            } else {  // flags in the members' mask words
                hipLaunchKernelGGL(k_lvl_part_wave<false>, dim3(wg), dim3(256), 0, s, a, cur, na,
                                   (const int32_t* const*)P, P + T, P + 2 * T, P + 3 * T, mask16.p, flags8.p, split.p,
                                   nchild.p, leftsz.p);
                hipLaunchKernelGGL(k_lvl_part_entry<false>, dim3(eg), dim3(256), 0, s, a, cur, na,
                                   (const int32_t* const*)P, P + T, P + 2 * T, P + 3 * T, mask16.p, split.p, nchild.p,
                                   leftsz.p, flags8.p);
            }
        } else {
            if (vec4)
                hipLaunchKernelGGL(k_lvl_flag_chunks<true>, dim3(cgrid), dim3(256), 0, s, T, n,
                                   sparse ? nullptr : mask16.p, (const int32_t* const*)P, flags8.p, chunk_cnt.p);
            else
                hipLaunchKernelGGL(k_lvl_flag_chunks<false>, dim3(cgrid), dim3(256), 0, s, T, n,
                                   sparse ? nullptr : mask16.p, (const int32_t* const*)P, flags8.p, chunk_cnt.p);
            size_t b = scan_bytes;
            HIP_TRY(hipcub::DeviceScan::ExclusiveSum(scan_tmp.p, b, chunk_cnt.p, chunk_off.p, (int)(nchunks + 1), s));
            hipLaunchKernelGGL(k_lvl_counts, dim3(lgrid(na, 256)), dim3(256), 0, s, a, cur, na, chunk_off.p,
                               flags8.p, split.p, nchild.p, leftsz.p, lstart.p);
        }
        size_t cb = cscan_bytes;
        HIP_TRY(hipcub::DeviceScan::ExclusiveSum(cscan_tmp.p, cb, nchild.p, cscan.p, (int)(na + 1), s));
        hipLaunchKernelGGL(k_lvl_tree_base, dim3(1), dim3(64), 0, s, T, st_cur, cscan.p, tree_children, d_count.p,
                           d_base.p, st_next);
        HIP_TRY(hipMemcpyAsync(lst, lstat.p, lst_bytes, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipEventRecord(lst_ev, s));
        // children ids, the next active list and the partition need nothing from the host
        hipLaunchKernelGGL(k_lvl_assign, dim3(lgrid(na, 256)), dim3(256), 0, s, cur, na, split.p, cscan.p,
                           leftsz.p, d_base.p, st_next, d_ch.p, nxt, cinfo.p);
        if (entry_mode) {
        } else if (vec4)
            hipLaunchKernelGGL(k_lvl_scatter<true>, dim3(cgrid), dim3(256), 0, s, cur, T, n, split.p, cscan.p,
                               leftsz.p, lstart.p, chunk_off.p, flags8.p, (const int32_t* const*)P, P + T, P + 2 * T,
                               P + 3 * T);
        else
            hipLaunchKernelGGL(k_lvl_scatter<false>, dim3(cgrid), dim3(256), 0, s, cur, T, n, split.p, cscan.p,
                               leftsz.p, lstart.p, chunk_off.p, flags8.p, (const int32_t* const*)P, P + T, P + 2 * T,
                               P + 3 * T);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventSynchronize(lst_ev));
        const unsigned long long pairs_now = lst[0];
        const unsigned* tc = reinterpret_cast<const unsigned*>(lst + 1);
        bool sparse_next = sparse;
        {
            // algorithmic bytes: every point row once, each split node's plane codes once (a byte a
            // coordinate), and per tested (tree, point) its node and position reads and its flag write
            const double tested = (double)(pairs_now - pairs_before);
            pairs_before = pairs_now;
            margins_timer.collect(std::min<double>(tested, (double)n) * dim * 4.0 + (double)na * dim +
                                  tested * 12.0);
            sparse_next = sparse || (dim % 4 == 0 && tested <= (double)(kSparsePairs * n));
        }
        int64_t next_na = 0;
        bool over = false;
        for (int t = 0; t < T; ++t) {  // the host's copy of the node counts k_lvl_tree_base keeps
            node_count[t] += tc[t];
            next_na += tc[t];
            if (node_count[t] > cap) over = true;
        }
        if (over) {  // the next level's nodes would not fit: the kernels above stayed inside this level's
            HIP_TRY(hipStreamSynchronize(s));
            return false;
        }
        sparse = sparse_next;
        std::swap(cur, nxt);
        std::swap(st_cur, st_next);
        phase ^= 1;
        na = next_na;
        ++level;
    }
    // after an odd number of levels the final order sits in the second buffer
    for (int t = 0; t < T; ++t) {
        if (phase == 1 && n > 0) HIP_TRY(hipMemcpyAsync(perm_h[t], perm2_h[t], n * 4, hipMemcpyDeviceToDevice, s));
        F.trees[t].n_nodes = node_count[t];
    }
    HIP_TRY(hipStreamSynchronize(s));  // host vectors used by the copies above go out of scope
    return true;
}

}  // namespace rinnd
