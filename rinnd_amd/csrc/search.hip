// search.hip - batched greedy graph search (core/search/greedy.rs:70-241) on the GPU.
//
// One wavefront per query, persistent over the batch.  Per query the wave keeps in LDS:
//   - the result heap: a bounded std::collections::BinaryHeap<(OrderedFloat<f32>, i32)> replayed
//     with Rust's sift_up / sift_down_to_bottom so that into_sorted's tie order (a stable sort of
//     the heap's internal layout by distance, candidate_heap.rs:140-148) is reproduced;
//   - the seed heap: a min-heap over (distance, id); ids are unique (visited set), so its pop order
//     is fixed by the (distance, id) order alone.  It spills to a per-slot global region when its
//     LDS capacity is exceeded;
//   - the visited set: an epoch byte per vertex in a per-slot global array (VisitedSet,
//     core/visited.rs:8-63, with an 8-bit epoch).
// A popped vertex's CSR neighbours are checked, marked and measured by the lanes in parallel; the
// insertion decisions, which depend on the bound updated after every insert, are then replayed
// sequentially by lane 0 in CSR order (greedy.rs:176-200).
#include "search.h"

namespace rinnd {

#ifndef RINND_SEED_LDS
#define RINND_SEED_LDS 2048
#endif
constexpr int kSeedLds = RINND_SEED_LDS;  // seed-heap entries a wave keeps in LDS before it spills to memory
constexpr int kMaxK = 512;
constexpr int kMaxDeg = 256;
constexpr int kFrontier = 256;
constexpr int kQLds = 2048;

__device__ __forceinline__ int ord_cmp(float ad, int32_t ai, float bd, int32_t bi) {
    if (ad < bd) return -1;
    if (ad > bd) return 1;
    return ai < bi ? -1 : (ai > bi ? 1 : 0);
}

// ---- Rust std BinaryHeap (max-heap) ----
__device__ void bh_push(float* hd, int32_t* hi, int& len, float d, int32_t x) {
    int pos = len++;
    while (pos > 0) {  // sift_up(0, pos)
        const int parent = (pos - 1) / 2;
        if (ord_cmp(d, x, hd[parent], hi[parent]) <= 0) break;
        hd[pos] = hd[parent];
        hi[pos] = hi[parent];
        pos = parent;
    }
    hd[pos] = d;
    hi[pos] = x;
}
__device__ void bh_pop(float* hd, int32_t* hi, int& len) {
    // item = data.pop(); if !empty { swap(item, data[0]); sift_down_to_bottom(0) }
    --len;
    if (len == 0) return;
    const float ed = hd[len];
    const int32_t ei = hi[len];
    const int end = len;
    int pos = 0, child = 1;
    while (end >= 2 && child <= end - 2) {
        child += ord_cmp(hd[child], hi[child], hd[child + 1], hi[child + 1]) <= 0 ? 1 : 0;
        hd[pos] = hd[child];
        hi[pos] = hi[child];
        pos = child;
        child = 2 * pos + 1;
    }
    if (child == end - 1) {
        hd[pos] = hd[child];
        hi[pos] = hi[child];
        pos = child;
    }
    while (pos > 0) {  // sift_up(0, pos)
        const int parent = (pos - 1) / 2;
        if (ord_cmp(ed, ei, hd[parent], hi[parent]) <= 0) break;
        hd[pos] = hd[parent];
        hi[pos] = hi[parent];
        pos = parent;
    }
    hd[pos] = ed;
    hi[pos] = ei;
}
// BoundedHeap::push (candidate_heap.rs:97-113)
__device__ bool bounded_push(float* hd, int32_t* hi, int& len, int k, float d, int32_t x) {
    if (len < k) {
        bh_push(hd, hi, len, d, x);
        return true;
    }
    if (len > 0 && d < hd[0]) {
        bh_pop(hd, hi, len);
        bh_push(hd, hi, len, d, x);
        return true;
    }
    return false;
}

// ---- seed min-heap ----
__device__ void mh_push(float* hd, int32_t* hi, int64_t& len, float d, int32_t x) {
    int64_t pos = len++;
    while (pos > 0) {
        const int64_t parent = (pos - 1) >> 1;
        if (ord_cmp(hd[parent], hi[parent], d, x) <= 0) break;
        hd[pos] = hd[parent];
        hi[pos] = hi[parent];
        pos = parent;
    }
    hd[pos] = d;
    hi[pos] = x;
}
__device__ void mh_pop(float* hd, int32_t* hi, int64_t& len, float& d, int32_t& x) {
    d = hd[0];
    x = hi[0];
    --len;
    if (len == 0) return;
    const float ld = hd[len];
    const int32_t li = hi[len];
    int64_t pos = 0;
    for (;;) {
        int64_t c = 2 * pos + 1;
        if (c >= len) break;
        if (c + 1 < len && ord_cmp(hd[c + 1], hi[c + 1], hd[c], hi[c]) < 0) ++c;
        if (ord_cmp(ld, li, hd[c], hi[c]) <= 0) break;
        hd[pos] = hd[c];
        hi[pos] = hi[c];
        pos = c;
    }
    hd[pos] = ld;
    hi[pos] = li;
}

// Per-wave LDS besides the dynamic part (the query row and the result heap, sized by dim and k).  The
// best-first frontier of the tree descent and the neighbour batch of the seeding/traversal are never live
// together, so they share storage.
struct WaveState {
    float sh_d[kSeedLds];
    int32_t sh_i[kSeedLds];
    union {
        struct {
            int32_t nb_i[kMaxDeg];
            float nb_d[kMaxDeg];
            uint8_t nb_ok[kMaxDeg];
        };
        struct {
            float fr_p[kFrontier];
            int32_t fr_n[kFrontier];
        };
    };
    int64_t leaves[64 * 2];
    int rh_len;
    int64_t sh_len;
    int n_leaves;
    int done;
    int32_t v;
    float bound;
    int global_mode;
};

__device__ __forceinline__ float corrected(int corr, float x) {
    if (corr == RINND_CORR_SQRT) return sqrtf(x);
    if (corr == RINND_CORR_ALT_COSINE) return 1.0f - powf(2.0f, -x);
    if (corr == RINND_CORR_ALT_INNER_PRODUCT) return x >= 3.40282347e+38f ? 0.0f : -1.0f / x;
    return x;
}

// The metric evaluated from its parts (q.x, |q|^2, |x|^2, |q - x|^2), with the same formulas as
// distance_exact: used for the FP8 candidate distances, which are approximate by construction.
__device__ __forceinline__ float distance_from_parts(int metric, float dot, float qq, float xx, float d2) {
    const float FMAX = 3.40282347e+38f;
    switch (metric) {
        case RINND_DIST_SQEUCLIDEAN: return d2;
        case RINND_DIST_COSINE: {
            const float denom = sqrtf(qq * xx);
            if (denom < 1e-12f) return 1.0f;
            const float c = fminf(fmaxf(dot / denom, -1.0f), 1.0f);
            return 1.0f - c;
        }
        case RINND_DIST_INNER_PRODUCT: return -dot;
        case RINND_DIST_ALT_DOT: return dot <= 0.0f ? FMAX : -log2f(dot);
        case RINND_DIST_DIRECT_COSINE: return 1.0f - fminf(fmaxf(dot, -1.0f), 1.0f);
        case RINND_DIST_ALT_COSINE:
            if (qq == 0.0f && xx == 0.0f) return 0.0f;
            if (qq == 0.0f || xx == 0.0f || dot <= 0.0f) return FMAX;
            return log2f(sqrtf(qq * xx) / dot);
        case RINND_DIST_ALT_INNER_PRODUCT: return dot <= 0.0f ? FMAX : 1.0f / dot;
        default: return __builtin_nanf("");
    }
}

// FP8 candidate distances on the matrix cores.  Both the rows and the query are e4m3 codes with one scale each
// (the query's codes sit in LDS); the dot products of a batch of up to 64 neighbours (lane t: neighbour x, valid
// ok) are 16 x 16 x 32 FP8 MFMAs with FP32 accumulation: 16 neighbours' code rows form the A operand, the query's
// codes the B operand (the same in all 16 columns), and lane t takes its neighbour's row of the result.  Every
// lane of the wave must call it.  Returns the dot product of the codes (scale it by both rows' scales).
typedef long fp8x8;
__device__ __forceinline__ float fp8_wave_dot(const uint8_t* __restrict__ q8, int64_t ld8, const uint8_t* qc8,
                                              int32_t x, bool ok, int lane) {
    const unsigned long long live = __ballot(ok);
    const int g = lane >> 4, m = lane & 15;
    float mine = 0.f;
    for (int tile = 0; tile < 4; ++tile) {
        if (!((live >> (16 * tile)) & 0xffffull)) continue;  // wave-uniform
        const int src = 16 * tile + m;
        const int32_t xr = __shfl(x, src);
        const bool okr = (live >> src) & 1ull;
        const uint8_t* row = q8 + (int64_t)(okr ? xr : 0) * ld8 + 8 * g;
        v4f acc = {0.f, 0.f, 0.f, 0.f};
        for (int64_t k0 = 0; k0 < ld8; k0 += 32) {
            const fp8x8 av = okr ? *reinterpret_cast<const fp8x8*>(row + k0) : (fp8x8)0;
            const fp8x8 bv = *reinterpret_cast<const fp8x8*>(qc8 + k0 + 8 * g);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(av, bv, acc, 0, 0, 0);
        }
        // acc[r] = dot(tile row 4 g + r, q) in every column: lane 16 tile + i takes row i from lane group i / 4
        const int i = lane - 16 * tile;
        const int from = ((i >> 2) & 3) * 16;
        const float v0 = __shfl(acc[0], from), v1 = __shfl(acc[1], from), v2 = __shfl(acc[2], from),
                    v3 = __shfl(acc[3], from);
        if (i >= 0 && i < 16) mine = (i & 3) == 0 ? v0 : (i & 3) == 1 ? v1 : (i & 3) == 2 ? v2 : v3;
    }
    return mine;
}
// The same dot product by one lane (the random fill of the seeds).
__device__ __forceinline__ float fp8_lane_dot(const uint8_t* __restrict__ c, const uint8_t* qc8, int64_t ld8) {
    float dot = 0.f;
    for (int64_t i = 0; i < ld8; i += 4) {
        const uint32_t wx = *reinterpret_cast<const uint32_t*>(c + i), wq = *reinterpret_cast<const uint32_t*>(qc8 + i);
        dot = fmaf(__builtin_amdgcn_cvt_f32_fp8(wx, 0), __builtin_amdgcn_cvt_f32_fp8(wq, 0), dot);
        dot = fmaf(__builtin_amdgcn_cvt_f32_fp8(wx, 1), __builtin_amdgcn_cvt_f32_fp8(wq, 1), dot);
        dot = fmaf(__builtin_amdgcn_cvt_f32_fp8(wx, 2), __builtin_amdgcn_cvt_f32_fp8(wq, 2), dot);
        dot = fmaf(__builtin_amdgcn_cvt_f32_fp8(wx, 3), __builtin_amdgcn_cvt_f32_fp8(wq, 3), dot);
    }
    return dot;
}

// int8 candidate distances (the quantized index's codes): exact integer dot products, so any order of the sums
// gives the reference's i32.  A batch of up to 64 neighbours goes through 16 x 16 x 32 int8 MFMAs with i32
// accumulation, laid out as fp8_wave_dot; every lane of the wave must call it.
typedef int v4i __attribute__((ext_vector_type(4)));
__device__ __forceinline__ int32_t i8_wave_dot(const int8_t* __restrict__ c8, int64_t ldi8, const int8_t* qc,
                                               int32_t x, bool ok, int lane) {
    const unsigned long long live = __ballot(ok);
    const int g = lane >> 4, m = lane & 15;
    int32_t mine = 0;
    for (int tile = 0; tile < 4; ++tile) {
        if (!((live >> (16 * tile)) & 0xffffull)) continue;  // wave-uniform
        const int src = 16 * tile + m;
        const int32_t xr = __shfl(x, src);
        const bool okr = (live >> src) & 1ull;
        const int8_t* row = c8 + (int64_t)(okr ? xr : 0) * ldi8 + 8 * g;
        v4i acc = {0, 0, 0, 0};
        for (int64_t k0 = 0; k0 < ldi8; k0 += 32) {
            const long av = okr ? *reinterpret_cast<const long*>(row + k0) : 0l;
            const long bv = *reinterpret_cast<const long*>(qc + k0 + 8 * g);
            acc = __builtin_amdgcn_mfma_i32_16x16x32_i8(av, bv, acc, 0, 0, 0);
        }
        const int i = lane - 16 * tile;
        const int from = ((i >> 2) & 3) * 16;
        const int32_t v0 = __shfl(acc[0], from), v1 = __shfl(acc[1], from), v2 = __shfl(acc[2], from),
                      v3 = __shfl(acc[3], from);
        if (i >= 0 && i < 16) mine = (i & 3) == 0 ? v0 : (i & 3) == 1 ? v1 : (i & 3) == 2 ? v2 : v3;
    }
    return mine;
}
// The same dot product by one lane, four codes per dot4.
__device__ __forceinline__ int32_t i8_lane_dot(const int8_t* __restrict__ c, const int8_t* qc, int64_t ldi8) {
    int32_t dot = 0;
    for (int64_t i = 0; i < ldi8; i += 4)
        dot = __builtin_amdgcn_sdot4(*reinterpret_cast<const int*>(c + i), *reinterpret_cast<const int*>(qc + i), dot,
                                     false);
    return dot;
}

// Sorts the first len entries of (d, id) by (d, id) ascending with a wave-wide bitonic network in LDS (padding to
// a power of two with (+inf, INT32_MAX)); the sorted array is a valid min-heap.  len <= capacity, the padded
// length must fit the arrays too.  Every lane of the wave calls it.
__device__ void wave_sort_pairs(float* d, int32_t* id, int len, int lane) {
    int P = 1;
    while (P < len) P <<= 1;
    for (int i = len + lane; i < P; i += 64) {
        d[i] = __builtin_inff();
        id[i] = 0x7fffffff;
    }
    wave_sync();
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = lane; i < P; i += 64) {
                const int l = i ^ j;
                if (l > i) {
                    const float di = d[i], dl = d[l];
                    const int32_t ii = id[i], il = id[l];
                    const bool gt = di > dl || (di == dl && ii > il);  // entry i after entry l
                    if (gt == ((i & k) == 0)) {                       // ascending halves where (i & k) == 0
                        d[i] = dl;
                        d[l] = di;
                        id[i] = il;
                        id[l] = ii;
                    }
                }
            }
            wave_sync();
        }
    }
}

// Candidate distances: FP32 rows in the reference's arithmetic, FP8 e4m3 codes, or the quantized index's int8
// codes.
enum : int { kRepF32 = 0, kRepFP8 = 1, kRepI8 = 2 };

// One instance per metric (METRIC >= 0; the switch folds away, so only that metric's accumulators take
// registers) and one for any metric (METRIC < 0).
template <int REP, int METRIC>
__global__ __launch_bounds__(64) void k_search(SearchArgs a) {
    constexpr bool Q8 = REP == kRepFP8;
    const int metric = METRIC >= 0 ? METRIC : a.metric;
    __shared__ WaveState S;
    extern __shared__ __align__(16) float sdyn[];
    float* const sq = sdyn;                 // a.qcap floats (0: the query is read from memory)
    float* const rh_d = sdyn + a.qcap;      // result heap, a.kcap entries
    int32_t* const rh_i = (int32_t*)(rh_d + a.kcap);
    uint8_t* const qc8 = (uint8_t*)(rh_i + a.kcap);  // FP8 / int8 search: the query's codes, a.c8cap bytes
    int8_t* const qi8 = (int8_t*)(qc8 + a.c8cap);     // int8 routing: the query's codes, a.i8cap bytes
    const int lane = threadIdx.x;
    const int slot = blockIdx.x;
    uint8_t* vis = a.visited + (int64_t)slot * a.n;
    float* gsd = a.seed_spill_d + (int64_t)slot * a.seed_cap;
    int32_t* gsi = a.seed_spill_i + (int64_t)slot * a.seed_cap;
    const int dim = a.dim;
    const bool q_in_lds = a.qcap > 0;
    for (int64_t qi = blockIdx.x; qi < a.nq; qi += gridDim.x) {
        const int64_t qid = a.q_qid ? a.q_qid[qi] : a.qid_base + qi;
        const float eps = a.q_eps ? a.q_eps[qi] : a.eps;
        const float* qg = a.Q + qi * a.ldq;
        if (q_in_lds)
            for (int i = lane; i < dim; i += 64) sq[i] = qg[i];
        const float* q = q_in_lds ? sq : qg;
        // FP8 search: the query's e4m3 codes (scale max|q| / 240) in LDS and the squared norm of its dequantized form
        float qsc = 0.f, qq = 0.f;
        if (Q8) {
            float mx = 0.f;
            for (int i = lane; i < dim; i += 64) mx = fmaxf(mx, fabsf(qg[i]));
            for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
            qsc = mx > 0.f ? mx / 240.0f : 1.0f;
            const float inv = 1.0f / qsc;
            for (int i4 = lane; i4 < (int)(a.ld8 >> 2); i4 += 64) {
                float v[4];
#pragma unroll
                for (int b = 0; b < 4; ++b) v[b] = 4 * i4 + b < dim ? qg[4 * i4 + b] * inv : 0.f;
                int word = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], 0, false);
                word = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], word, true);
                reinterpret_cast<uint32_t*>(qc8)[i4] = (uint32_t)word;
                const float f[4] = {qsc * __builtin_amdgcn_cvt_f32_fp8(word, 0), qsc * __builtin_amdgcn_cvt_f32_fp8(word, 1),
                                    qsc * __builtin_amdgcn_cvt_f32_fp8(word, 2), qsc * __builtin_amdgcn_cvt_f32_fp8(word, 3)};
#pragma unroll
                for (int b = 0; b < 4; ++b) qq = fmaf(f[b], f[b], qq);
            }
            for (int o = 32; o > 0; o >>= 1) qq += __shfl_xor(qq, o);
        }
        // int8 search: the query's codes under the index's encoding (quantize_query, crates/rinnd/src/lib.rs:325-358),
        // zero to ldi8, and its inverse norm (int8), the global dequantisation scale (sq8u) or 1 / its multiplier
        // (sq8p)
        float qinv = 0.f;
        if (REP == kRepI8) {
            float mul = a.i8gm;
            if (a.i8enc != 1) {
                float mx = 0.f;
                for (int i = lane; i < dim; i += 64) mx = fmaxf(mx, fabsf(qg[i]));
                for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
                mul = mx > 0.f ? 127.0f / mx : 0.f;
            }
            int8_t* const qc = (int8_t*)qc8;
            int32_t ns = 0;
            for (int i = lane; i < (int)a.ldi8; i += 64) {
                const int8_t c = i < dim ? (int8_t)fminf(127.f, fmaxf(-127.f, roundf(qg[i] * mul))) : (int8_t)0;
                qc[i] = c;
                ns += (int32_t)c * (int32_t)c;
            }
            for (int o = 32; o > 0; o >>= 1) ns += __shfl_xor(ns, o);
            qinv = a.i8enc == 0 ? (ns > 0 ? 1.0f / sqrtf((float)ns) : 0.f)
                 : a.i8enc == 1 ? a.i8gd
                                : (mul > 0.f ? 1.0f / mul : 0.f);
        }
        // int8 routing trees: the query's per-vector symmetric codes (crates/rinnd/src/lib.rs:320-347) and its scale
        // (tree_query_scale, lib.rs:303-313)
        float tqs = 0.f;
        if (a.hq) {
            float mx = 0.f;
            for (int i = lane; i < dim; i += 64) mx = fmaxf(mx, fabsf(q[i]));
            for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
            const float mul = mx > 0.f ? 127.0f / mx : 0.f;
            for (int i = lane; i < dim; i += 64) qi8[i] = (int8_t)fminf(127.f, fmaxf(-127.f, roundf(q[i] * mul)));
            tqs = mx > 0.f ? mx / 127.0f : 0.f;
        }
        __threadfence_block();
        wave_sync();
        // a search tree node's margin: FP32 hyperplane (flat_tree.rs:253-273), or the int8 one reconstructed
        // (flat_tree.rs:95-101: i8 dot as f32, times the node's scale, times the query's scale, plus the offset)
        auto tree_margin = [&](int t, int64_t node) -> float {  // lane 0 only
            if (a.hq) {
                const int8_t* hc = a.hq[t] + node * a.hdim;
                int32_t dot = 0;
                for (int i = 0; i < dim; ++i) dot += (int32_t)hc[i] * (int32_t)qi8[i];
                return (float)dot * a.hqs[t][node] * tqs + a.offsets[t][node];
            }
            return d_dot8(q, a.hyper[t] + node * a.hdim, dim) + a.offsets[t][node];
        };
        auto fp8_dist = [&](int32_t x, float code_dot) -> float {
            const float dot = code_dot * a.q8s[x] * qsc;
            const float xx = a.q8n[x];
            return distance_from_parts(metric, dot, qq, xx, fmaxf(qq + xx - 2.0f * dot, 0.0f));
        };
        // greedy.rs:381-387: 1 - clamp(i8 dot as f32 * query inverse norm * row inverse norm, -1, 1)
        auto i8_dist = [&](int32_t x, int32_t dot) -> float {
            const float sim = (float)dot * qinv * a.i8inv[x];
            return 1.0f - fminf(fmaxf(sim, -1.0f), 1.0f);
        };
        // candidate distance of one lane: FP32 in the reference's arithmetic, or from the FP8 / int8 codes
        auto cand_dist = [&](int32_t x) -> float {
            if (Q8) return fp8_dist(x, fp8_lane_dot(a.q8 + (int64_t)x * a.ld8, qc8, a.ld8));
            if (REP == kRepI8) return i8_dist(x, i8_lane_dot(a.i8 + (int64_t)x * a.ldi8, (const int8_t*)qc8, a.ldi8));
            return distance_exact(metric, q, a.X + (int64_t)x * a.ld, dim);
        };
        // distances of a batch of neighbours (every lane calls it; lane t: neighbour x, valid ok)
        auto batch_dist = [&](int32_t x, bool ok) -> float {
            if (Q8) {
                const float cd = fp8_wave_dot(a.q8, a.ld8, qc8, x, ok, lane);
                return ok ? fp8_dist(x, cd) : 0.f;
            }
            if (REP == kRepI8) {
                const int32_t cd = i8_wave_dot(a.i8, a.ldi8, (const int8_t*)qc8, x, ok, lane);
                return ok ? i8_dist(x, cd) : 0.f;
            }
            return ok ? distance_exact(metric, q, a.X + (int64_t)x * a.ld, dim) : 0.f;
        };
        // epoch (VisitedSet::clear, visited.rs:51-57)
        int ep = a.epochs[slot] + 1;
        if (ep > 255) {
            for (int64_t i = lane; i < a.n; i += 64) vis[i] = 0;
            ep = 1;
        }
        if (lane == 0) {
            a.epochs[slot] = ep;
            S.rh_len = 0;
            S.sh_len = 0;
            S.global_mode = 0;
        }
        __threadfence_block();
        wave_sync();
        Xoshiro rng = xoshiro_seed(a.seed + (a.seed_fixed ? 0ull : (uint64_t)qid));  // FastRng::new(seed + query_id)
        const int k = a.q_k ? a.q_k[qi] : a.k;

        // seed heap storage (LDS, or the per-slot global region after an overflow)
        auto seed_push = [&](float d, int32_t x) {  // lane 0 only
            if (!S.global_mode && S.sh_len >= kSeedLds) {
                for (int64_t t = 0; t < S.sh_len; ++t) {
                    gsd[t] = S.sh_d[t];
                    gsi[t] = S.sh_i[t];
                }
                S.global_mode = 1;
            }
            if (S.global_mode) {
                if (S.sh_len >= a.seed_cap) {
                    *a.status = 1;
                    return;
                }
                mh_push(gsd, gsi, S.sh_len, d, x);
            } else {
                mh_push(S.sh_d, S.sh_i, S.sh_len, d, x);
            }
        };

        // The seeds enter the frontier in bulk: a min-heap over unique (distance, id) keys pops in key order
        // whatever its layout, so the seeds are appended in parallel and the array is sorted once (a sorted array
        // is a heap); a frontier that would outgrow the LDS array is sorted at that point and takes the seeds
        // after it one by one (seed_push, which spills to memory).  The result heap takes every push in the
        // reference's order, its layout decides the order of ties in into_sorted.
        bool bulk = true;
        // ---- initialize_fp32 (greedy.rs:70-99) ----
        for (int t = 0; t < a.n_trees; ++t) {
            const int32_t* ch = a.children[t];
            if (lane == 0) {
                int nl = 0;
                if (a.leaf_budget == 1) {  // hard descent (flat_tree.rs:253-273)
                    int64_t node = 0;
                    while (ch[2 * node] > 0) {
                        const float margin = tree_margin(t, node);
                        const int side = fabsf(margin) < 1e-8f ? (int)rng.next_bool() : (int)(margin > 0.0f);
                        node = ch[2 * node + side];
                    }
                    S.leaves[0] = -ch[2 * node];
                    S.leaves[1] = -ch[2 * node + 1];
                    nl = 1;
                } else {  // bounded best-first spill (flat_tree.rs:275-308)
                    int nf = 1;
                    S.fr_p[0] = 0.f;
                    S.fr_n[0] = 0;
                    while (nl < a.leaf_budget && nf > 0) {
                        int best = 0;
                        for (int i = 1; i < nf; ++i) {  // min_by(total_cmp): first minimum
                            int32_t bi = __builtin_bit_cast(int32_t, S.fr_p[i]);
                            int32_t bb = __builtin_bit_cast(int32_t, S.fr_p[best]);
                            bi ^= (int32_t)(((uint32_t)(bi >> 31)) >> 1);
                            bb ^= (int32_t)(((uint32_t)(bb >> 31)) >> 1);
                            if (bi < bb) best = i;
                        }
                        const float prio = S.fr_p[best];
                        const int32_t node = S.fr_n[best];
                        for (int i = best; i + 1 < nf; ++i) {
                            S.fr_p[i] = S.fr_p[i + 1];
                            S.fr_n[i] = S.fr_n[i + 1];
                        }
                        --nf;
                        if (ch[2 * node] <= 0) {
                            S.leaves[2 * nl] = -ch[2 * node];
                            S.leaves[2 * nl + 1] = -ch[2 * node + 1];
                            ++nl;
                            continue;
                        }
                        const float margin = tree_margin(t, node);
                        const int side = fabsf(margin) < 1e-8f ? (int)rng.next_bool() : (int)(margin > 0.0f);
                        if (nf + 2 > kFrontier) {
                            *a.status = 2;
                            break;
                        }
                        S.fr_p[nf] = prio;
                        S.fr_n[nf] = ch[2 * node + side];
                        S.fr_p[nf + 1] = fmaxf(prio, fabsf(margin));
                        S.fr_n[nf + 1] = ch[2 * node + 1 - side];
                        nf += 2;
                    }
                }
                S.n_leaves = nl;
            }
            wave_sync();
            const int nl = S.n_leaves;
            const int32_t* tix = a.tindices[t];
            for (int l = 0; l < nl; ++l) {
                const int64_t s0 = S.leaves[2 * l], s1 = S.leaves[2 * l + 1];
                for (int64_t c0 = s0; c0 < s1; c0 += 64) {
                    const int64_t p = c0 + lane;
                    bool ok = false;
                    int32_t x = -1;
                    float d = 0.f;
                    if (p < s1) {
                        x = tix[p];
                        if (x >= 0) {
                            ok = vis[x] != ep;
                            vis[x] = (uint8_t)ep;
                        }
                    }
                    d = batch_dist(x, ok);
                    S.nb_i[lane] = x;
                    S.nb_d[lane] = d;
                    S.nb_ok[lane] = ok;
                    const uint64_t okm = __ballot(ok);
                    const int64_t len0 = S.sh_len;  // uniform: lane 0 last wrote it before a wave sync
                    bool appended = false;
                    if (bulk) {
                        if (len0 + __popcll(okm) <= kSeedLds) {
                            if (ok) {
                                const int64_t pos = len0 + __popcll(okm & ((1ull << lane) - 1ull));
                                S.sh_d[pos] = d;
                                S.sh_i[pos] = x;
                            }
                            appended = true;
                        } else {
                            __threadfence_block();
                            wave_sync();
                            wave_sort_pairs(S.sh_d, S.sh_i, (int)len0, lane);
                            bulk = false;
                        }
                    }
                    __threadfence_block();
                    wave_sync();
                    if (lane == 0) {
                        const int cnt = (int)min((int64_t)64, s1 - c0);
                        for (int j = 0; j < cnt; ++j)
                            if (S.nb_ok[j]) {
                                bounded_push(rh_d, rh_i, S.rh_len, k, S.nb_d[j], S.nb_i[j]);
                                if (!appended) seed_push(S.nb_d[j], S.nb_i[j]);
                            }
                        if (appended) S.sh_len = len0 + __popcll(okm);
                    }
                    wave_sync();
                }
            }
        }
        // ---- fill_random (greedy.rs:148-163) ----
        if (lane == 0) {
            const int64_t nfill = (int64_t)k > S.sh_len ? (int64_t)k - S.sh_len : 0;
            for (int64_t r = 0; r < nfill; ++r) {
                const int32_t x = (int32_t)rng.next_index((uint64_t)a.n);
                if (vis[x] != ep) {
                    vis[x] = (uint8_t)ep;
                    const float d = cand_dist(x);
                    bounded_push(rh_d, rh_i, S.rh_len, k, d, x);
                    if (bulk) {  // fills only up to k < kSeedLds entries: the array has room
                        S.sh_d[S.sh_len] = d;
                        S.sh_i[S.sh_len] = x;
                        ++S.sh_len;
                    } else {
                        seed_push(d, x);
                    }
                }
            }
            const float maxd = S.rh_len > 0 ? rh_d[0] : __builtin_inff();
            S.bound = maxd + eps * (maxd - a.min_distance);
        }
        __threadfence_block();
        wave_sync();
        if (bulk) wave_sort_pairs(S.sh_d, S.sh_i, (int)S.sh_len, lane);  // uniform
        __threadfence_block();
        wave_sync();
        // ---- traverse (greedy.rs:165-201) ----
        for (;;) {
            if (lane == 0) {
                S.done = 1;
                if (S.sh_len > 0) {
                    float dv;
                    int32_t v;
                    if (S.global_mode) mh_pop(gsd, gsi, S.sh_len, dv, v);
                    else mh_pop(S.sh_d, S.sh_i, S.sh_len, dv, v);
                    if (!(S.rh_len >= k && dv >= S.bound)) {
                        S.done = 0;
                        S.v = v;
                    }
                }
            }
            wave_sync();
            if (S.done) break;
            const int32_t v = S.v;
            const int32_t beg = a.indptr[v];
            const int deg = a.indptr[v + 1] - beg;
            for (int c0 = 0; c0 < deg; c0 += 64) {
                const int t = c0 + lane;
                bool ok = false;
                int32_t x = -1;
                float d = 0.f;
                if (t < deg) {
                    x = a.indices[beg + t];
                    if (x >= 0) {
                        ok = vis[x] != ep;
                        vis[x] = (uint8_t)ep;
                    }
                }
                d = batch_dist(x, ok);
                S.nb_i[lane] = x;
                S.nb_d[lane] = d;
                S.nb_ok[lane] = ok;
                __threadfence_block();
                wave_sync();
                if (lane == 0) {
                    const int cnt = min(64, deg - c0);
                    float bound = S.bound;
                    for (int j = 0; j < cnt; ++j) {
                        if (!S.nb_ok[j]) continue;
                        const float dj = S.nb_d[j];
                        if (dj < bound) {
                            const bool ins = bounded_push(rh_d, rh_i, S.rh_len, k, dj, S.nb_i[j]);
                            seed_push(dj, S.nb_i[j]);
                            if (ins) {
                                const float maxd = rh_d[0];
                                bound = maxd + eps * (maxd - a.min_distance);
                            }
                        }
                    }
                    S.bound = bound;
                }
                wave_sync();
            }
        }
        // ---- into_sorted + id remap + correction + padding (index.rs:182-199) ----
        if (lane == 0) {
            const int m = S.rh_len;
            for (int i = 1; i < m; ++i) {  // stable insertion sort by distance
                const float d = rh_d[i];
                const int32_t x = rh_i[i];
                int j = i - 1;
                while (j >= 0 && rh_d[j] > d) {
                    rh_d[j + 1] = rh_d[j];
                    rh_i[j + 1] = rh_i[j];
                    --j;
                }
                rh_d[j + 1] = d;
                rh_i[j + 1] = x;
            }
        }
        wave_sync();
        int m = S.rh_len;
        const int rerank = a.q_rerank ? a.q_rerank[qi] : a.rerank;
        if (rerank > 0) {
            // rerank_quantized_candidates (core/index.rs:103-124): the first `rerank` candidates get their
            // FP32 distance, then sort by (distance total order, internal id) and keep k
            const int rw = min(rerank, m);
            for (int j = lane; j < rw; j += 64)
                rh_d[j] = distance_exact(metric, q, a.X + (int64_t)rh_i[j] * a.ld, dim);
            __threadfence_block();
            wave_sync();
            if (lane == 0) {
                auto key = [](float d) {  // f32::total_cmp order
                    int32_t b = __builtin_bit_cast(int32_t, d);
                    return b ^ (int32_t)(((uint32_t)(b >> 31)) >> 1);
                };
                for (int i = 1; i < rw; ++i) {
                    const float d = rh_d[i];
                    const int32_t x = rh_i[i];
                    int j = i - 1;
                    while (j >= 0 && (key(rh_d[j]) > key(d) || (key(rh_d[j]) == key(d) && rh_i[j] > x))) {
                        rh_d[j + 1] = rh_d[j];
                        rh_i[j + 1] = rh_i[j];
                        --j;
                    }
                    rh_d[j + 1] = d;
                    rh_i[j + 1] = x;
                }
            }
            wave_sync();
            m = rw;
        }
        const int ko = a.q_kout ? a.q_kout[qi] : a.k_out;
        int32_t* const oip = a.q_oi ? a.q_oi[qi] : a.out_idx + qi * ko;
        float* const odp = a.q_od ? a.q_od[qi] : a.out_dist + qi * ko;
        for (int j = lane; j < ko; j += 64) {
            int32_t oi = -1;
            float od = __builtin_inff();
            if (j < m) {
                const int32_t x = rh_i[j];
                oi = x >= 0 ? (int32_t)a.vertex_order[x] : x;
                od = corrected(a.corr, rh_d[j]);
            }
            oip[j] = oi;
            odp[j] = od;
        }
        wave_sync();
    }
}

template <int REP>
static void launch_search_q(hipStream_t s, const SearchArgs& a, int slots, size_t lds) {
    switch (a.metric) {
        case RINND_DIST_SQEUCLIDEAN:
            hipLaunchKernelGGL((k_search<REP, RINND_DIST_SQEUCLIDEAN>), dim3(slots), dim3(64), lds, s, a);
            break;
        case RINND_DIST_ALT_DOT:
            hipLaunchKernelGGL((k_search<REP, RINND_DIST_ALT_DOT>), dim3(slots), dim3(64), lds, s, a);
            break;
        case RINND_DIST_DIRECT_COSINE:
            hipLaunchKernelGGL((k_search<REP, RINND_DIST_DIRECT_COSINE>), dim3(slots), dim3(64), lds, s, a);
            break;
        default: hipLaunchKernelGGL((k_search<REP, -1>), dim3(slots), dim3(64), lds, s, a);
    }
}
void launch_search(hipStream_t s, SearchArgs a, int slots) {
    if (a.i8 && (a.q8 || a.metric != RINND_DIST_DIRECT_COSINE))
        throw Error{RINND_ERR_INVALID_ARG, "int8 codes are searched alone, on a directly normalized cosine index"};
    a.qcap = a.dim <= kQLds ? (int)round_up(a.dim, 4) : 0;
    a.kcap = (int)round_up(std::max(a.k, 1), 4);
    a.c8cap = a.q8 ? (int)round_up(a.ld8, (int64_t)16) : a.i8 ? (int)round_up(a.ldi8, (int64_t)16) : 0;
    a.i8cap = a.hq ? (int)round_up((int64_t)a.dim, (int64_t)16) : 0;
    const size_t lds = ((size_t)a.qcap + 2 * (size_t)a.kcap) * 4 + (size_t)a.c8cap + (size_t)a.i8cap;
    if (a.q8) launch_search_q<kRepFP8>(s, a, slots, lds);
    else if (a.i8) hipLaunchKernelGGL((k_search<kRepI8, RINND_DIST_DIRECT_COSINE>), dim3(slots), dim3(64), lds, s, a);
    else launch_search_q<kRepF32>(s, a, slots, lds);
}

// One wave per node: max |h| over the plane by a wave reduction, the node's scale, then its codes (the padding
// columns stay zero).
__global__ __launch_bounds__(256) void k_quantize_tree(const float* __restrict__ hyper, int64_t n_nodes, int64_t hdim,
                                                       int dim, int8_t* __restrict__ codes, float* __restrict__ scales) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int64_t v = (int64_t)blockIdx.x * 4 + w; v < n_nodes; v += (int64_t)gridDim.x * 4) {
        const float* h = hyper + v * hdim;
        float mx = 0.f;
        for (int i = lane; i < dim; i += 64) mx = fmaxf(mx, fabsf(h[i]));
        for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
        const float sc = mx > 0.f ? mx / 127.0f : 0.f;
        if (lane == 0) scales[v] = sc;
        for (int64_t i = lane; i < hdim; i += 64)
            codes[v * hdim + i] =
                (int8_t)(i < dim && sc > 0.f ? fminf(127.f, fmaxf(-127.f, roundf(h[i] / sc))) : 0.f);
    }
}
void launch_quantize_tree(hipStream_t s, const float* hyper, int64_t n_nodes, int64_t hdim, int dim, int8_t* codes,
                          float* scales) {
    if (n_nodes <= 0) return;
    hipLaunchKernelGGL(k_quantize_tree, dim3(grid_for(n_nodes, 4, 4096)), dim3(256), 0, s, hyper, n_nodes, hdim, dim,
                       codes, scales);
    HIP_TRY(hipGetLastError());
}

// One wave per row: max |x| by a wave reduction, then the codes, two values per conversion.
__global__ __launch_bounds__(256) void k_fp8_encode(const float* __restrict__ X, int64_t ld, int dim, int64_t n,
                                                    uint8_t* __restrict__ codes, int64_t ld8, float* __restrict__ scales,
                                                    float* __restrict__ norms) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int64_t r = (int64_t)blockIdx.x * 4 + w; r < n; r += (int64_t)gridDim.x * 4) {
        const float* x = X + r * ld;
        float mx = 0.f;
        for (int i = lane; i < dim; i += 64) mx = fmaxf(mx, fabsf(x[i]));
        for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
        const float sc = mx > 0.f ? mx / 240.0f : 1.0f;  // 240: representable in both e4m3 variants
        const float inv = 1.0f / sc;
        if (lane == 0) scales[r] = sc;
        uint32_t* out = reinterpret_cast<uint32_t*>(codes + r * ld8);
        float xx = 0.f;
        for (int i4 = lane; i4 < (int)(ld8 >> 2); i4 += 64) {
            float v[4];
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int i = 4 * i4 + b;
                v[b] = i < dim ? x[i] * inv : 0.f;
            }
            int word = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], 0, false);
            word = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], word, true);
            out[i4] = (uint32_t)word;
            const float f[4] = {sc * __builtin_amdgcn_cvt_f32_fp8(word, 0), sc * __builtin_amdgcn_cvt_f32_fp8(word, 1),
                                sc * __builtin_amdgcn_cvt_f32_fp8(word, 2), sc * __builtin_amdgcn_cvt_f32_fp8(word, 3)};
#pragma unroll
            for (int b = 0; b < 4; ++b) xx = fmaf(f[b], f[b], xx);
        }
        for (int o = 32; o > 0; o >>= 1) xx += __shfl_xor(xx, o);
        if (lane == 0) norms[r] = xx;
    }
}
void launch_fp8_encode(hipStream_t s, const float* X, int64_t ld, int dim, int64_t n, uint8_t* codes, int64_t ld8,
                       float* scales, float* norms) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_fp8_encode, dim3((unsigned)std::min<int64_t>(ceil_div(n, (int64_t)4), 1 << 16)), dim3(256), 0,
                       s, X, ld, dim, n, codes, ld8, scales, norms);
}

// The quantized index's codes, one wave per row (QuantizedI8Index::new, crates/rinnd/src/lib.rs:166-224): the
// multiplier 127 / max|row| (or the global one, sq8u), codes round(x * multiplier) clamped to [-127, 127] and zero
// to ldi8, and per row 1 / sqrt(sum of squared codes) (int8), the global dequantisation scale (sq8u) or
// 1 / multiplier (sq8p).
__global__ __launch_bounds__(256) void k_i8_encode(const float* __restrict__ X, int64_t ld, int dim, int64_t n,
                                                   int enc, float gm, float gd, int8_t* __restrict__ codes,
                                                   int64_t ldi8, float* __restrict__ inv) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int64_t r = (int64_t)blockIdx.x * 4 + w; r < n; r += (int64_t)gridDim.x * 4) {
        const float* x = X + r * ld;
        float mul = gm;
        if (enc != 1) {
            float mx = 0.f;
            for (int i = lane; i < dim; i += 64) mx = fmaxf(mx, fabsf(x[i]));
            for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
            mul = mx > 0.f ? 127.0f / mx : 0.f;
        }
        int32_t ns = 0;
        for (int64_t i = lane; i < ldi8; i += 64) {
            const int8_t c = i < dim ? (int8_t)fminf(127.f, fmaxf(-127.f, roundf(x[i] * mul))) : (int8_t)0;
            codes[r * ldi8 + i] = c;
            ns += (int32_t)c * (int32_t)c;
        }
        for (int o = 32; o > 0; o >>= 1) ns += __shfl_xor(ns, o);
        if (lane == 0)
            inv[r] = enc == 0 ? (ns > 0 ? 1.0f / sqrtf((float)ns) : 0.f) : enc == 1 ? gd : (mul > 0.f ? 1.0f / mul : 0.f);
    }
}
void launch_i8_encode(hipStream_t s, const float* X, int64_t ld, int dim, int64_t n, int enc, float gm, float gd,
                      int8_t* codes, int64_t ldi8, float* inv) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_i8_encode, dim3((unsigned)std::min<int64_t>(ceil_div(n, (int64_t)4), 1 << 16)), dim3(256), 0,
                       s, X, ld, dim, n, enc, gm, gd, codes, ldi8, inv);
    HIP_TRY(hipGetLastError());
}

// max |x| over the n x dim block (rows of stride ld) into *out (which starts at zero): the bits of a non-negative
// float order as the float does, so an unsigned atomic max keeps the largest.
__global__ __launch_bounds__(256) void k_abs_max(const float* __restrict__ X, int64_t ld, int dim, int64_t n,
                                                 unsigned* out) {
    float mx = 0.f;
    const int64_t total = n * (int64_t)dim;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256)
        mx = fmaxf(mx, fabsf(X[(e / dim) * ld + e % dim]));
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(mx));
}
void launch_abs_max(hipStream_t s, const float* X, int64_t ld, int dim, int64_t n, unsigned* out) {
    if (n <= 0 || dim <= 0) return;
    const int64_t total = n * (int64_t)dim;
    hipLaunchKernelGGL(k_abs_max, dim3((unsigned)std::min<int64_t>(ceil_div(total, (int64_t)256), 4096)), dim3(256), 0,
                       s, X, ld, dim, n, out);
    HIP_TRY(hipGetLastError());
}

int search_max_k() { return kMaxK; }
int search_max_degree() { return kMaxDeg; }
int search_max_budget() { return 64; }

}  // namespace rinnd
