// rinnd_oracle.cpp - CPU restatement of lmcinnes/rinnd's NN-Descent build and greedy graph search.
//
// TEST INFRASTRUCTURE (the oracle): only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
// leg load this library, as the checker and as the timed CPU baseline.  The product (rinnd_amd/) never
// links or calls it.
//
// Build: oracle/build.sh (g++ -O3 -mavx2 -mfma -ffp-contract=off).  Each function cites the reference
// code it restates; "core" = crates/rinnd-core/src, "py" = crates/rinnd/src/lib.rs.
//
// Pinning status of each part (what checks it, see tests/test_oracle_*.py):
//   FastRng              - SplitMix64/xoshiro256++ published vectors and rand 0.8.5 conversion
//                          identities; the reference has no FastRng vector, so its stream is pinned
//                          only by the published algorithms (no Rust toolchain here).
//   distances            - reference KATs (euclidean.rs:187-234 ...) + exact-integer closed forms +
//                          an independent scalar re-derivation of the AVX2 lane/hsum order.
//   NeighborHeap         - reference KATs (neighbor_heap.rs:339-474) + brute-force replay.
//   candidate heap push  - reference KATs (candidates.rs:490-592).
//   RP tree              - reference invariants (builder.rs:367-430), flat_tree.rs:408-478 KATs,
//                          partition closed form; split choices pinned only through FastRng.
//   nn_descent pieces    - order argument (T- and block-independence) checked across T; invariants
//                          of nndescent/mod.rs:508-568; recall vs brute force.
//   search graph         - search_graph.rs:573-667 KATs (weighted union) + invariants.
//   BinaryHeap / search  - candidate_heap.rs:169-255 KATs, std BinaryHeap invariants, query-finds-self.
#include "rinnd_oracle.hpp"

#include <immintrin.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <queue>
#include <thread>

namespace oracle {

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

// Xoshiro256PlusPlus::seed_from_u64 = SplitMix64 stream -> 4 little-endian state words
// (rand_xoshiro 0.6.0 `seed_from_u64`, rand_core 0.6.4 `from_rng`; used at core/rng.rs:22-26).
FastRng::FastRng(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) {
        x += 0x9e3779b97f4a7c15ULL;
        uint64_t z = x;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        s[i] = z ^ (z >> 31);
    }
}

// xoshiro256++ next_u64 (core/rng.rs:30-32 via rand_xoshiro).
uint64_t FastRng::next_u64() {
    const uint64_t r = rotl64(s[0] + s[3], 23) + s[0];
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return r;
}

// rand 0.8.5 Standard f32: 24 high bits of next_u32 times 2^-24 (core/rng.rs:42-44).
float FastRng::next_float() { return (float)(next_u32() >> 8) * (1.0f / 16777216.0f); }

// rand 0.8.5 UniformInt<usize>::sample_single(0, n) -> sample_single_inclusive(0, n-1)
// (core/rng.rs:48-50): widening multiply with the "leading zeros" rejection zone.
uint64_t FastRng::next_index(uint64_t n) {
    const uint64_t range = n;  // (n-1) - 0 + 1
    if (range == 0) return next_u64();
    const uint64_t zone = (range << __builtin_clzll(range)) - 1;
    for (;;) {
        const uint64_t v = next_u64();
        const unsigned __int128 m = (unsigned __int128)v * range;
        const uint64_t lo = (uint64_t)m;
        if (lo <= zone) return (uint64_t)(m >> 64);
    }
}

// ---- distances ---------------------------------------------------------------------------------

// hsum256_ps_avx (core/distance/euclidean.rs:99-110): ((l0+l4)+(l1+l5)) + ((l2+l6)+(l3+l7)).
static inline float hsum_ps(__m256 v) {
    __m128 hi = _mm256_extractf128_ps(v, 1);
    __m128 lo = _mm256_castps256_ps128(v);
    __m128 sum128 = _mm_add_ps(hi, lo);
    __m128 shuf = _mm_movehdup_ps(sum128);
    __m128 sums = _mm_add_ps(sum128, shuf);
    __m128 shuf2 = _mm_movehl_ps(sums, sums);
    return _mm_cvtss_f32(_mm_add_ss(sums, shuf2));
}

// horizontal_sum_avx (core/distance/alternatives.rs:94-104): (s0+s2)+(s1+s3), s_i = l_i + l_{i+4}.
static inline float hsum_alt(__m256 v) {
    __m128 high = _mm256_extractf128_ps(v, 1);
    __m128 low = _mm256_castps256_ps128(v);
    __m128 sum = _mm_add_ps(low, high);
    __m128 high_pairs = _mm_movehl_ps(sum, sum);
    __m128 pairs = _mm_add_ps(sum, high_pairs);
    __m128 odd = _mm_shuffle_ps(pairs, pairs, 0x55);
    return _mm_cvtss_f32(_mm_add_ss(pairs, odd));
}

// l2_sqr_avx2 (core/distance/euclidean.rs:66-93).
static float l2_sqr_avx2(const float* a, const float* b, size_t n) {
    const size_t chunks = n / 8;
    __m256 sum = _mm256_setzero_ps();
    for (size_t i = 0; i < chunks; ++i) {
        __m256 aa = _mm256_loadu_ps(a + i * 8);
        __m256 bb = _mm256_loadu_ps(b + i * 8);
        __m256 diff = _mm256_sub_ps(aa, bb);
        sum = _mm256_fmadd_ps(diff, diff, sum);
    }
    float result = hsum_ps(sum);
    for (size_t i = chunks * 8; i < n; ++i) {
        float d = a[i] - b[i];
        result += d * d;
    }
    return result;
}

// cosine_avx2 (core/distance/cosine.rs:44-82).
static float cosine_avx2(const float* a, const float* b, size_t n) {
    const size_t chunks = n / 8;
    __m256 vdot = _mm256_setzero_ps(), vna = _mm256_setzero_ps(), vnb = _mm256_setzero_ps();
    for (size_t i = 0; i < chunks; ++i) {
        __m256 aa = _mm256_loadu_ps(a + i * 8);
        __m256 bb = _mm256_loadu_ps(b + i * 8);
        vdot = _mm256_fmadd_ps(aa, bb, vdot);
        vna = _mm256_fmadd_ps(aa, aa, vna);
        vnb = _mm256_fmadd_ps(bb, bb, vnb);
    }
    float dot = hsum_ps(vdot), na = hsum_ps(vna), nb = hsum_ps(vnb);
    for (size_t i = chunks * 8; i < n; ++i) {
        dot += a[i] * b[i];
        na += a[i] * a[i];
        nb += b[i] * b[i];
    }
    float denom = std::sqrt(na * nb);
    if (denom < 1e-12f) return 1.0f;
    float sim = dot / denom;
    sim = sim < -1.0f ? -1.0f : (sim > 1.0f ? 1.0f : sim);  // f32::clamp
    return 1.0f - sim;
}

// dot_avx2 (core/distance/inner_product.rs:64-83); also builder.rs:29-60 (same arithmetic).
static float dot_avx2(const float* a, const float* b, size_t n) {
    const size_t chunks = n / 8;
    __m256 v = _mm256_setzero_ps();
    for (size_t i = 0; i < chunks; ++i)
        v = _mm256_fmadd_ps(_mm256_loadu_ps(a + i * 8), _mm256_loadu_ps(b + i * 8), v);
    float dot = hsum_ps(v);
    for (size_t i = chunks * 8; i < n; ++i) dot += a[i] * b[i];
    return dot;
}

// normalized_dot_avx2 (core/distance/alternatives.rs:55-89): four 8-lane accumulators.
static float normalized_dot_avx2(const float* a, const float* b, size_t n) {
    __m256 s0 = _mm256_setzero_ps(), s1 = _mm256_setzero_ps(), s2 = _mm256_setzero_ps(),
           s3 = _mm256_setzero_ps();
    const size_t chunks = n / 32;
    for (size_t c = 0; c < chunks; ++c) {
        const size_t i = c * 32;
        s0 = _mm256_fmadd_ps(_mm256_loadu_ps(a + i), _mm256_loadu_ps(b + i), s0);
        s1 = _mm256_fmadd_ps(_mm256_loadu_ps(a + i + 8), _mm256_loadu_ps(b + i + 8), s1);
        s2 = _mm256_fmadd_ps(_mm256_loadu_ps(a + i + 16), _mm256_loadu_ps(b + i + 16), s2);
        s3 = _mm256_fmadd_ps(_mm256_loadu_ps(a + i + 24), _mm256_loadu_ps(b + i + 24), s3);
    }
    __m256 s01 = _mm256_add_ps(s0, s1), s23 = _mm256_add_ps(s2, s3);
    float dot = hsum_alt(_mm256_add_ps(s01, s23));
    for (size_t i = chunks * 32; i < n; ++i) dot += a[i] * b[i];
    return dot;
}

float tree_dot(const float* a, const float* b, size_t dim) { return dot_avx2(a, b, dim); }

float distance(int metric, const float* a, const float* b, size_t n) {
    const float FMAX = std::numeric_limits<float>::max();
    switch (metric) {
        case SQEUCLIDEAN: return l2_sqr_avx2(a, b, n);
        case COSINE: return cosine_avx2(a, b, n);
        case INNER_PRODUCT: return -dot_avx2(a, b, n);  // inner_product.rs:98-102
        case ALT_DOT: {  // alternatives.rs:39-51
            float dot = normalized_dot_avx2(a, b, n);
            return dot <= 0.0f ? FMAX : -std::log2(dot);
        }
        case DIRECT_COSINE: {  // alternatives.rs:133-149
            float s = normalized_dot_avx2(a, b, n);
            s = s < -1.0f ? -1.0f : (s > 1.0f ? 1.0f : s);
            return 1.0f - s;
        }
        case ALT_COSINE: {  // alternatives.rs:202-226 (scalar, sequential)
            float dot = 0.0f, na = 0.0f, nb = 0.0f;
            for (size_t i = 0; i < n; ++i) {
                dot += a[i] * b[i];
                na += a[i] * a[i];
                nb += b[i] * b[i];
            }
            if (na == 0.0f && nb == 0.0f) return 0.0f;
            if (na == 0.0f || nb == 0.0f) return FMAX;
            if (dot <= 0.0f) return FMAX;
            return std::log2(std::sqrt(na * nb) / dot);
        }
        case ALT_INNER_PRODUCT: {  // alternatives.rs:162-188 (scalar, 4-chunk)
            float dot = 0.0f;
            const size_t chunks = n / 4;
            for (size_t c = 0; c < chunks; ++c) {
                const size_t i = c * 4;
                dot += a[i] * b[i] + a[i + 1] * b[i + 1] + a[i + 2] * b[i + 2] + a[i + 3] * b[i + 3];
            }
            for (size_t i = chunks * 4; i < n; ++i) dot += a[i] * b[i];
            return dot <= 0.0f ? FMAX : 1.0f / dot;
        }
        default: return std::numeric_limits<float>::quiet_NaN();
    }
}

float correct(int corr, float d) {
    switch (corr) {
        case CORR_SQRT: return std::sqrt(d);
        case CORR_ALT_COSINE: return 1.0f - std::pow(2.0f, -d);  // alternatives.rs:276-278
        case CORR_ALT_IP:                                       // alternatives.rs:282-288
            return d >= std::numeric_limits<float>::max() ? 0.0f : -1.0f / d;
        default: return d;
    }
}

// ---- threads (Rayon `(0..T).into_par_iter()` with manual block ownership) ----------------------
static void par_for(int T, const std::function<void(int)>& f) {
    if (T <= 1) {
        f(0);
        return;
    }
    std::vector<std::thread> ts;
    ts.reserve(T);
    for (int t = 0; t < T; ++t) ts.emplace_back(f, t);
    for (auto& th : ts) th.join();
}

// ---- NeighborHeap (core/heap/neighbor_heap.rs) ------------------------------------------------
NeighborHeap::NeighborHeap(size_t n, size_t k_)
    : indices(n * k_, -1), distances(n * k_, std::numeric_limits<float>::infinity()),
      flags(n * k_, 0), n_points(n), k(k_) {}

// checked_flagged_push (neighbor_heap.rs:149-177).
bool NeighborHeap::checked_flagged_push(size_t p, int32_t nb, float d, bool is_new) {
    const size_t off = p * k;
    if (d >= distances[off]) return false;
    for (size_t i = 0; i < k; ++i)
        if (indices[off + i] == nb) return false;
    distances[off] = d;
    indices[off] = nb;
    flags[off] = is_new ? 1 : 0;
    sift_down(off, 0);
    return true;
}

// simple_push (neighbor_heap.rs:95-109).
bool NeighborHeap::simple_push(size_t p, int32_t nb, float d) {
    const size_t off = p * k;
    if (d >= distances[off]) return false;
    distances[off] = d;
    indices[off] = nb;
    flags[off] = 0;
    sift_down(off, 0);
    return true;
}

// unchecked_flagged_push (neighbor_heap.rs:184-197).
void NeighborHeap::unchecked_flagged_push(size_t p, int32_t nb, float d, bool is_new) {
    const size_t off = p * k;
    distances[off] = d;
    indices[off] = nb;
    flags[off] = is_new ? 1 : 0;
    sift_down(off, 0);
}

// sift_down (neighbor_heap.rs:207-243): hole method, larger child with ties to the left.
void NeighborHeap::sift_down(size_t off, size_t start) {
    const size_t end = k;
    const float vd = distances[off + start];
    const int32_t vi = indices[off + start];
    const uint8_t vf = flags[off + start];
    size_t pos = start, child = 2 * pos + 1;
    while (child < end) {
        size_t right = child + 1;
        if (right < end && distances[off + child] < distances[off + right]) child = right;
        if (vd >= distances[off + child]) break;
        distances[off + pos] = distances[off + child];
        indices[off + pos] = indices[off + child];
        flags[off + pos] = flags[off + child];
        pos = child;
        child = 2 * pos + 1;
    }
    distances[off + pos] = vd;
    indices[off + pos] = vi;
    flags[off + pos] = vf;
}

// sort_all / sort_row (neighbor_heap.rs:293-320): stable sort by distance.
void NeighborHeap::sort_all() {
    struct E { float d; int32_t i; uint8_t f; };
    std::vector<E> row(k);
    for (size_t p = 0; p < n_points; ++p) {
        const size_t off = p * k;
        for (size_t j = 0; j < k; ++j) row[j] = {distances[off + j], indices[off + j], flags[off + j]};
        std::stable_sort(row.begin(), row.end(), [](const E& a, const E& b) { return a.d < b.d; });
        for (size_t j = 0; j < k; ++j) {
            distances[off + j] = row[j].d;
            indices[off + j] = row[j].i;
            flags[off + j] = row[j].f;
        }
    }
}

// ---- RP tree (core/tree/builder.rs) -------------------------------------------------------------
namespace {
struct TreeBuilder {
    const float* data;
    size_t dim, leaf_size, max_depth;
    bool angular;
    std::vector<float> hyperplanes, offsets;
    std::vector<int32_t> children, leaf_indices;

    void push_leaf(const int32_t* idx, size_t m) {
        const int32_t start = (int32_t)leaf_indices.size();
        leaf_indices.insert(leaf_indices.end(), idx, idx + m);
        const int32_t end = (int32_t)leaf_indices.size();
        hyperplanes.resize(hyperplanes.size() + dim, 0.0f);
        offsets.push_back(0.0f);
        children.push_back(-start);
        children.push_back(-end);
    }

    // make_euclidean_split (builder.rs:249-279).
    float euclidean_split(const float* p1, const float* p2, std::vector<float>& h) {
        float norm_sq = 0.0f, dot_nm = 0.0f;
        for (size_t i = 0; i < dim; ++i) {
            float hv = p2[i] - p1[i];
            float m = (p1[i] + p2[i]) * 0.5f;
            h[i] = hv;
            norm_sq += hv * hv;
            dot_nm += hv * m;
        }
        float norm = std::sqrt(norm_sq);
        if (norm > 1e-8f) {
            float inv = 1.0f / norm;
            for (size_t i = 0; i < dim; ++i) h[i] *= inv;
            return -dot_nm * inv;
        }
        return 0.0f;
    }

    // make_angular_split (builder.rs:281-308).
    float angular_split(const float* p1, const float* p2, std::vector<float>& h) {
        float n1 = 0.0f, n2 = 0.0f;
        for (size_t i = 0; i < dim; ++i) n1 += p1[i] * p1[i];
        for (size_t i = 0; i < dim; ++i) n2 += p2[i] * p2[i];
        n1 = std::sqrt(n1);
        n2 = std::sqrt(n2);
        if (n1 < 1e-8f || n2 < 1e-8f) {
            std::fill(h.begin(), h.end(), 0.0f);
            return 0.0f;
        }
        for (size_t i = 0; i < dim; ++i) h[i] = p2[i] / n2 - p1[i] / n1;
        float nn = 0.0f;
        for (size_t i = 0; i < dim; ++i) nn += h[i] * h[i];
        nn = std::sqrt(nn);
        if (nn > 1e-8f)
            for (size_t i = 0; i < dim; ++i) h[i] /= nn;
        return 0.0f;
    }

    // partition_inplace (builder.rs:312-348): two-pointer swap partition.
    size_t partition(int32_t* idx, size_t m, const float* h, float off, FastRng& rng) {
        size_t left = 0, right = m;
        while (left < right) {
            const float* pt = data + (size_t)idx[left] * dim;
            float margin = dot_avx2(pt, h, dim) + off;
            bool goes_left = std::fabs(margin) < 1e-8f ? !rng.next_bool() : margin < 0.0f;
            if (goes_left) {
                ++left;
            } else {
                --right;
                std::swap(idx[left], idx[right]);
            }
        }
        return left;
    }

    // build_recursive (builder.rs:172-226) + make_split (228-247).
    int32_t build(int32_t* idx, size_t m, size_t depth, FastRng& rng) {
        const int32_t node = (int32_t)(children.size() / 2);
        if (m <= leaf_size || depth >= max_depth) {
            push_leaf(idx, m);
            return node;
        }
        size_t i1 = rng.next_index(m);
        size_t i2 = rng.next_index(m);
        while (i2 == i1 && m > 1) i2 = rng.next_index(m);
        const float* p1 = data + (size_t)idx[i1] * dim;
        const float* p2 = data + (size_t)idx[i2] * dim;
        std::vector<float> h(dim);
        float off = angular ? angular_split(p1, p2, h) : euclidean_split(p1, p2, h);
        size_t split = partition(idx, m, h.data(), off, rng);
        if (split == 0 || split == m) {
            push_leaf(idx, m);
            return node;
        }
        hyperplanes.insert(hyperplanes.end(), h.begin(), h.end());
        offsets.push_back(off);
        children.push_back(0);
        children.push_back(0);
        int32_t l = build(idx, split, depth + 1, rng);
        int32_t r = build(idx + split, m - split, depth + 1, rng);
        children[2 * node] = l;
        children[2 * node + 1] = r;
        return node;
    }
};
}  // namespace

// build_rp_tree (builder.rs:72-86).
FlatTree build_rp_tree(const float* data, size_t n, size_t dim, size_t leaf_size, FastRng& rng,
                       bool angular, size_t max_depth) {
    TreeBuilder b{data, dim, leaf_size, max_depth, angular, {}, {}, {}, {}};
    std::vector<int32_t> idx(n);
    for (size_t i = 0; i < n; ++i) idx[i] = (int32_t)i;
    b.build(idx.data(), n, 0, rng);
    FlatTree t;
    t.hyperplanes = std::move(b.hyperplanes);
    t.offsets = std::move(b.offsets);
    t.children = std::move(b.children);
    t.indices = std::move(b.leaf_indices);
    t.dim = dim;
    t.n_nodes = t.offsets.size();
    return t;
}

// build_rp_forest (builder.rs:89-118): one seed per tree from the main RNG, trees independent.
std::vector<FlatTree> build_rp_forest(const float* data, size_t n, size_t dim, size_t n_trees,
                                      size_t leaf_size, FastRng& rng, bool angular,
                                      size_t max_depth, int threads) {
    std::vector<uint64_t> seeds(n_trees);
    for (auto& s : seeds) s = rng.next_u64();
    std::vector<FlatTree> forest(n_trees);
    int T = std::max(1, std::min<int>(threads, (int)n_trees));
    par_for(T, [&](int t) {
        for (size_t i = t; i < n_trees; i += T) {
            FastRng tr(seeds[i]);
            forest[i] = build_rp_tree(data, n, dim, leaf_size, tr, angular, max_depth);
        }
    });
    return forest;
}

// get_all_leaves / collect_leaves (flat_tree.rs:339-356): DFS order.
std::vector<std::pair<size_t, size_t>> FlatTree::get_all_leaves() const {
    std::vector<std::pair<size_t, size_t>> out;
    if (children.empty()) return out;
    std::vector<size_t> stack{0};
    while (!stack.empty()) {
        size_t node = stack.back();
        stack.pop_back();
        if (children[2 * node] <= 0) {
            out.emplace_back((size_t)(-children[2 * node]), (size_t)(-children[2 * node + 1]));
        } else {
            stack.push_back((size_t)children[2 * node + 1]);
            stack.push_back((size_t)children[2 * node]);
        }
    }
    return out;
}

// search_leaves (flat_tree.rs:239-311).
size_t FlatTree::search_leaves(const float* point, FastRng& rng, size_t budget,
                               std::vector<std::pair<size_t, size_t>>& leaves) const {
    leaves.clear();
    if (children.empty() || budget == 0) return 0;
    size_t evals = 0;
    if (budget == 1) {
        size_t node = 0;
        while (children[2 * node] > 0) {
            float margin = dot_avx2(point, hyperplanes.data() + node * dim, dim) + offsets[node];
            ++evals;
            size_t side = std::fabs(margin) < 1e-8f ? (size_t)rng.next_bool() : (size_t)(margin > 0.0f);
            node = (size_t)children[2 * node + side];
        }
        leaves.emplace_back((size_t)(-children[2 * node]), (size_t)(-children[2 * node + 1]));
        return evals;
    }
    std::vector<std::pair<float, size_t>> frontier{{0.0f, 0}};
    auto total_lt = [](float a, float b) {  // f32::total_cmp(a, b) == Less
        int32_t ia, ib;
        std::memcpy(&ia, &a, 4);
        std::memcpy(&ib, &b, 4);
        ia ^= (int32_t)(((uint32_t)(ia >> 31)) >> 1);
        ib ^= (int32_t)(((uint32_t)(ib >> 31)) >> 1);
        return ia < ib;
    };
    while (leaves.size() < budget && !frontier.empty()) {
        size_t best = 0;  // min_by: first minimum
        for (size_t i = 1; i < frontier.size(); ++i)
            if (total_lt(frontier[i].first, frontier[best].first)) best = i;
        auto [prio, node] = frontier[best];
        frontier.erase(frontier.begin() + best);
        if (children[2 * node] <= 0) {
            leaves.emplace_back((size_t)(-children[2 * node]), (size_t)(-children[2 * node + 1]));
            continue;
        }
        float margin = dot_avx2(point, hyperplanes.data() + node * dim, dim) + offsets[node];
        ++evals;
        size_t side = std::fabs(margin) < 1e-8f ? (size_t)rng.next_bool() : (size_t)(margin > 0.0f);
        size_t primary = (size_t)children[2 * node + side];
        size_t alternate = (size_t)children[2 * node + 1 - side];
        frontier.emplace_back(prio, primary);
        frontier.emplace_back(std::max(prio, std::fabs(margin)), alternate);
    }
    return evals;
}

// ---- int8 routing trees (flat_tree.rs:35-183, quantized.rs:26-40, crates/rinnd/src/lib.rs:303-347) ------------
int32_t quantized_i8_dot(const int8_t* x, const int8_t* y, size_t n) {
    int32_t s = 0;  // exact in i32 whatever the summation order
    for (size_t i = 0; i < n; ++i) s += (int32_t)x[i] * (int32_t)y[i];
    return s;
}

// quantize_hyperplanes (flat_tree.rs:46-63): scale = max|h| / 127 (0 for a zero plane), codes round(h / scale)
// (half away from zero) clamped to [-127, 127]
QuantizedFlatTree QuantizedFlatTree::from_flat_tree(const FlatTree& t) {
    QuantizedFlatTree q;
    q.dim = t.dim;
    q.n_nodes = t.n_nodes;
    q.offsets = t.offsets;
    q.children = t.children;
    q.indices = t.indices;
    q.hyperplanes.resize(t.n_nodes * t.dim);
    q.scales.resize(t.n_nodes);
    for (size_t v = 0; v < t.n_nodes; ++v) {
        const float* h = t.hyperplanes.data() + v * t.dim;
        float max_abs = 0.0f;
        for (size_t i = 0; i < t.dim; ++i) max_abs = std::max(max_abs, std::fabs(h[i]));
        const float scale = max_abs > 0.0f ? max_abs / 127.0f : 0.0f;
        q.scales[v] = scale;
        for (size_t i = 0; i < t.dim; ++i) {
            float c = 0.0f;
            if (scale > 0.0f) c = std::min(127.0f, std::max(-127.0f, std::round(h[i] / scale)));
            q.hyperplanes[v * t.dim + i] = (int8_t)c;
        }
    }
    return q;
}

// margin (flat_tree.rs:95-101): i8 dot as f32, times the node's scale, times the query's scale, plus the offset
float QuantizedFlatTree::margin(size_t node, const int8_t* query, float query_scale) const {
    const float dot = (float)quantized_i8_dot(hyperplanes.data() + node * dim, query, dim);
    return dot * scales[node] * query_scale + offsets[node];
}

// search_leaves (flat_tree.rs:103-164): FlatTree::search_leaves with the reconstructed margin
size_t QuantizedFlatTree::search_leaves(const int8_t* query, float query_scale, FastRng& rng, size_t budget,
                                        std::vector<std::pair<size_t, size_t>>& leaves) const {
    leaves.clear();
    if (children.empty() || budget == 0) return 0;
    size_t evals = 0;
    if (budget == 1) {
        size_t node = 0;
        while (children[2 * node] > 0) {
            const float m = margin(node, query, query_scale);
            ++evals;
            const size_t side = std::fabs(m) < 1e-8f ? (size_t)rng.next_bool() : (size_t)(m > 0.0f);
            node = (size_t)children[2 * node + side];
        }
        leaves.emplace_back((size_t)(-children[2 * node]), (size_t)(-children[2 * node + 1]));
        return evals;
    }
    std::vector<std::pair<float, size_t>> frontier{{0.0f, 0}};
    auto total_lt = [](float a, float b) {  // f32::total_cmp(a, b) == Less
        int32_t ia, ib;
        std::memcpy(&ia, &a, 4);
        std::memcpy(&ib, &b, 4);
        ia ^= (int32_t)(((uint32_t)(ia >> 31)) >> 1);
        ib ^= (int32_t)(((uint32_t)(ib >> 31)) >> 1);
        return ia < ib;
    };
    while (leaves.size() < budget && !frontier.empty()) {
        size_t best = 0;
        for (size_t i = 1; i < frontier.size(); ++i)
            if (total_lt(frontier[i].first, frontier[best].first)) best = i;
        auto [prio, node] = frontier[best];
        frontier.erase(frontier.begin() + best);
        if (children[2 * node] <= 0) {
            leaves.emplace_back((size_t)(-children[2 * node]), (size_t)(-children[2 * node + 1]));
            continue;
        }
        const float m = margin(node, query, query_scale);
        ++evals;
        const size_t side = std::fabs(m) < 1e-8f ? (size_t)rng.next_bool() : (size_t)(m > 0.0f);
        frontier.emplace_back(prio, (size_t)children[2 * node + side]);
        frontier.emplace_back(std::max(prio, std::fabs(m)), (size_t)children[2 * node + 1 - side]);
    }
    return evals;
}

// quantize_query / tree_query_scale of the per-vector encodings (crates/rinnd/src/lib.rs:303-347)
float quantize_query_i8(const float* q, size_t dim, int8_t* codes) {
    float max_abs = 0.0f;
    for (size_t i = 0; i < dim; ++i) max_abs = std::max(max_abs, std::fabs(q[i]));
    const float multiplier = max_abs > 0.0f ? 127.0f / max_abs : 0.0f;
    for (size_t i = 0; i < dim; ++i)
        codes[i] = (int8_t)std::min(127.0f, std::max(-127.0f, std::round(q[i] * multiplier)));
    return max_abs > 0.0f ? max_abs / 127.0f : 0.0f;
}

// ---- NN-Descent -------------------------------------------------------------------------------

// initialize_from_leaves (core/nndescent/mod.rs:203-300).
void initialize_from_leaves(NeighborHeap& g, const std::vector<std::vector<int32_t>>& leaves,
                            const float* data, size_t dim, int metric, int threads) {
    const size_t n = g.n_points, n_leaves = leaves.size();
    const size_t T = (size_t)std::max(1, threads);
    const size_t block_size = std::max<size_t>(T * 64, 128);
    const size_t vbs = (n + T - 1) / T;
    struct U { int32_t p, q; float d; };
    std::vector<std::vector<U>> updates(T);
    for (size_t bs = 0; bs < n_leaves; bs += block_size) {
        const size_t be = std::min(bs + block_size, n_leaves);
        const size_t blen = be - bs;
        par_for((int)T, [&](int t) {
            auto& out = updates[t];
            out.clear();
            const size_t lpt = (blen + T - 1) / T;
            const size_t s = t * lpt, e = std::min(s + lpt, blen);
            for (size_t li = s; li < e; ++li) {
                const auto& leaf = leaves[bs + li];
                for (size_t i = 0; i < leaf.size(); ++i) {
                    const int32_t p = leaf[i];
                    if (p < 0) break;
                    const float* pp = data + (size_t)p * dim;
                    for (size_t j = i + 1; j < leaf.size(); ++j) {
                        const int32_t q = leaf[j];
                        if (q < 0) break;
                        float d = distance(metric, pp, data + (size_t)q * dim, dim);
                        float thr = std::fmax(g.max_distance(p), g.max_distance(q));
                        if (d < thr) out.push_back({p, q, d});
                    }
                }
            }
        });
        par_for((int)T, [&](int t) {
            const size_t vs = t * vbs, ve = std::min(vs + vbs, n);
            for (size_t src = 0; src < T; ++src)
                for (const U& u : updates[src]) {
                    if ((size_t)u.p >= vs && (size_t)u.p < ve) g.checked_flagged_push(u.p, u.q, u.d, true);
                    if ((size_t)u.q >= vs && (size_t)u.q < ve) g.checked_flagged_push(u.q, u.p, u.d, true);
                }
        });
    }
}

// checked_heap_push_flat (core/nndescent/candidates.rs:433-484).
static inline void cand_push(float* pri, int32_t* idx, size_t n, float priority, int32_t index) {
    if (priority >= pri[0]) return;
    for (size_t i = 0; i < n; ++i)
        if (idx[i] == index) return;
    pri[0] = priority;
    idx[0] = index;
    size_t pos = 0;
    for (;;) {
        size_t l = 2 * pos + 1, r = 2 * pos + 2, largest = pos;
        if (l < n && pri[l] > pri[largest]) largest = l;
        if (r < n && pri[r] > pri[largest]) largest = r;
        if (largest == pos) break;
        pri[pos] = pri[largest];
        idx[pos] = idx[largest];
        pri[largest] = priority;
        idx[largest] = index;
        pos = largest;
    }
}

// mark_old_flags (candidates.rs:354-417).
static void mark_old_flags(NeighborHeap& g, const std::vector<int32_t>& new_idx, size_t C, int T) {
    const size_t n = g.n_points, k = g.k;
    const int TT = (n < 256 || T <= 1) ? 1 : T;
    const size_t bs = (n + TT - 1) / TT;
    par_for(TT, [&](int t) {
        const size_t s = t * bs, e = std::min(s + bs, n);
        for (size_t i = s; i < e; ++i)
            for (size_t j = 0; j < k; ++j) {
                const int32_t nb = g.indices[i * k + j];
                if (nb < 0 || g.flags[i * k + j] == 0) continue;
                for (size_t c = 0; c < C; ++c)
                    if (new_idx[i * C + c] == nb) {
                        g.flags[i * k + j] = 0;
                        break;
                    }
            }
    });
}

// CandidateSets::build_from_graph (candidates.rs:51-65) with the sequential (68-144) and parallel
// (153-351) paths.
CandidateSets build_candidates(NeighborHeap& g, size_t C, FastRng& rng, int threads) {
    const size_t n = g.n_points, k = g.k;
    const size_t T = (size_t)std::max(1, threads);
    CandidateSets cs;
    cs.n_vertices = n;
    cs.max_candidates = C;
    cs.new_indices.assign(n * C, -1);
    cs.old_indices.assign(n * C, -1);
    const float INF = std::numeric_limits<float>::infinity();
    std::vector<float> new_pri(n * C, INF), old_pri(n * C, INF);
    auto push = [&](size_t v, bool is_new, float pri, int32_t x) {
        if (is_new) cand_push(&new_pri[v * C], &cs.new_indices[v * C], C, pri, x);
        else cand_push(&old_pri[v * C], &cs.old_indices[v * C], C, pri, x);
    };
    if (!(n >= 256 && T > 1)) {
        for (size_t i = 0; i < n; ++i)
            for (size_t j = 0; j < k; ++j) {
                const int32_t nb = g.indices[i * k + j];
                if (nb < 0) continue;
                const bool is_new = g.flags[i * k + j] != 0;
                const float pri = rng.next_float();
                push(i, is_new, pri, nb);
                const float rpri = rng.next_float();
                push((size_t)nb, is_new, rpri, (int32_t)i);
            }
    } else {
        const size_t bsz = (n + T - 1) / T;
        std::vector<uint64_t> seeds(T);
        for (auto& s : seeds) s = rng.next_u64();
        struct R { size_t dest; int32_t cand; float pri; bool is_new; };
        std::vector<std::vector<std::vector<R>>> buckets(T, std::vector<std::vector<R>>(T));
        par_for((int)T, [&](int t) {
            const size_t bs = t * bsz;
            if (bs >= n) return;
            const size_t be = std::min((t + 1) * bsz, n);
            FastRng lr(seeds[t]);
            for (size_t i = bs; i < be; ++i)
                for (size_t j = 0; j < k; ++j) {
                    const int32_t nb = g.indices[i * k + j];
                    if (nb < 0) continue;
                    const bool is_new = g.flags[i * k + j] != 0;
                    const float pri = lr.next_float();
                    push(i, is_new, pri, nb);
                    const size_t db = (size_t)nb / bsz;
                    if (db < T) {
                        if (db == (size_t)t) push((size_t)nb, is_new, pri, (int32_t)i);
                        else buckets[t][db].push_back({(size_t)nb, (int32_t)i, pri, is_new});
                    }
                }
        });
        par_for((int)T, [&](int t) {
            if ((size_t)t * bsz >= n) return;
            for (size_t src = 0; src < T; ++src) {
                if (src == (size_t)t) continue;
                for (const R& r : buckets[src][t]) push(r.dest, r.is_new, r.pri, r.cand);
            }
        });
    }
    mark_old_flags(g, cs.new_indices, C, (int)T);
    return cs;
}

// update_iteration (core/nndescent/mod.rs:321-491).
size_t update_iteration(NeighborHeap& g, const CandidateSets& c, const float* data, size_t dim,
                        int metric, int threads) {
    const size_t n = g.n_points, C = c.max_candidates;
    const size_t T = (size_t)std::max(1, threads);
    const size_t BLOCK = 16384;
    const size_t vbs = (n + T - 1) / T;
    struct U { int32_t p, q; float d; };
    std::vector<std::vector<std::vector<U>>> buckets(T, std::vector<std::vector<U>>(T));
    size_t total = 0;
    const size_t n_blocks = (n + BLOCK - 1) / BLOCK;
    std::vector<size_t> local(T);
    for (size_t b = 0; b < n_blocks; ++b) {
        const size_t bs = b * BLOCK, be = std::min((b + 1) * BLOCK, n), blen = be - bs;
        const size_t rpt = (blen + T - 1) / T;
        for (auto& tb : buckets)
            for (auto& x : tb) x.clear();
        par_for((int)T, [&](int t) {
            auto& lb = buckets[t];
            const size_t rs = bs + t * rpt, re = std::min(rs + rpt, be);
            for (size_t i = rs; i < re; ++i) {
                const int32_t* nc = &c.new_indices[i * C];
                const int32_t* oc = &c.old_indices[i * C];
                for (size_t j = 0; j < C; ++j) {
                    const int32_t p = nc[j];
                    if (p < 0) continue;
                    const float* dp = data + (size_t)p * dim;
                    const float tp = g.max_distance(p);
                    auto emit = [&](int32_t q, float d) {
                        const size_t pb = (size_t)p / vbs, qb = (size_t)q / vbs;
                        lb[std::min(pb, T - 1)].push_back({p, q, d});
                        if (pb != qb) lb[std::min(qb, T - 1)].push_back({p, q, d});
                    };
                    for (size_t kk = j + 1; kk < C; ++kk) {
                        const int32_t q = nc[kk];
                        if (q < 0) continue;
                        const float mt = std::fmax(tp, g.max_distance(q));
                        const float d = distance(metric, dp, data + (size_t)q * dim, dim);
                        if (d <= mt) emit(q, d);
                    }
                    for (size_t kk = 0; kk < C; ++kk) {
                        const int32_t q = oc[kk];
                        if (q < 0 || p == q) continue;
                        const float mt = std::fmax(tp, g.max_distance(q));
                        const float d = distance(metric, dp, data + (size_t)q * dim, dim);
                        if (d <= mt) emit(q, d);
                    }
                }
            }
        });
        par_for((int)T, [&](int t) {
            const size_t vs = t * vbs, ve = std::min(vs + vbs, n);
            size_t ch = 0;
            for (size_t src = 0; src < T; ++src)
                for (const U& u : buckets[src][t]) {
                    if ((size_t)u.p >= vs && (size_t)u.p < ve && g.checked_flagged_push(u.p, u.q, u.d, true)) ++ch;
                    if ((size_t)u.q >= vs && (size_t)u.q < ve && g.checked_flagged_push(u.q, u.p, u.d, true)) ++ch;
                }
            local[t] = ch;
        });
        for (size_t t = 0; t < T; ++t) total += local[t];
    }
    return total;
}

static double now_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch()).count();
}

// nn_descent (core/nndescent/mod.rs:79-197).
NeighborHeap nn_descent(const float* data, size_t n, size_t dim, int metric, const Params& p,
                        FastRng& rng, int threads, std::vector<FlatTree>* forest_out,
                        BuildTrace* trace) {
    const size_t C = std::min(std::min(p.max_candidates, (size_t)60), p.n_neighbors);
    double t0 = now_ms();
    std::vector<FlatTree> forest =
        build_rp_forest(data, n, dim, p.n_trees, p.leaf_size, rng, p.angular, p.max_depth, threads);
    double t1 = now_ms();
    NeighborHeap g(n, p.n_neighbors);
    std::vector<std::vector<int32_t>> leaves;  // rptree_leaf_array (builder.rs:123-135)
    for (const auto& t : forest)
        for (auto [s, e] : t.get_all_leaves())
            leaves.emplace_back(t.indices.begin() + s, t.indices.begin() + e);
    initialize_from_leaves(g, leaves, data, dim, metric, threads);
    double t2 = now_ms();
    BuildTrace tr;
    tr.ms_forest = t1 - t0;
    tr.ms_init = t2 - t1;
    for (size_t it = 0; it < p.n_iters; ++it) {
        double a = now_ms();
        CandidateSets cs = build_candidates(g, C, rng, threads);
        double b = now_ms();
        size_t ch = update_iteration(g, cs, data, dim, metric, threads);
        double c = now_ms();
        tr.ms_candidates += b - a;
        tr.ms_updates += c - b;
        tr.changes.push_back(ch);
        const size_t threshold = (size_t)(p.delta * (float)p.n_neighbors * (float)n);
        if (ch <= threshold) break;
    }
    if (forest_out) *forest_out = std::move(forest);
    if (trace) *trace = tr;
    return g;
}

// ---- search graph -------------------------------------------------------------------------------

// diversify_row (search_graph.rs:432-473) and diversify_sorted_csr (478-523).
static std::vector<std::pair<int32_t, float>> diversify(const int32_t* ri, const float* rd,
                                                        const std::vector<size_t>& order,
                                                        bool stop_at_neg, const float* data,
                                                        size_t dim, int metric, float prob,
                                                        size_t max_degree, FastRng& prng) {
    std::vector<std::pair<int32_t, float>> kept;
    for (size_t pos : order) {
        const int32_t j = ri[pos];
        if (j < 0) {
            if (stop_at_neg) break;
            continue;
        }
        const float dij = rd[pos];
        bool keep = true;
        for (auto [c, dic] : kept) {
            const float dcj = distance(metric, data + (size_t)c * dim, data + (size_t)j * dim, dim);
            if (dic > std::numeric_limits<float>::min() && dcj < dij) {
                if (prob >= 1.0f || prng.next_float() < prob) {
                    keep = false;
                    break;
                }
            }
        }
        if (keep) {
            kept.emplace_back(j, dij);
            if (kept.size() >= max_degree) break;
        }
    }
    return kept;
}

// weighted_union_pruned (search_graph.rs:368-426).
SearchGraph weighted_union_pruned(const std::vector<std::vector<std::pair<int32_t, float>>>& fwd,
                                  const std::vector<std::vector<std::pair<int32_t, float>>>& rev,
                                  size_t n, size_t max_degree) {
    SearchGraph g;
    g.n_vertices = n;
    g.indptr.push_back(0);
    std::vector<std::pair<int32_t, float>> row, dd;
    for (size_t v = 0; v < n; ++v) {
        row.clear();
        auto ok = [&](const std::pair<int32_t, float>& e) {
            return e.first >= 0 && e.first != (int32_t)v && std::isfinite(e.second);
        };
        for (auto& e : fwd[v]) if (ok(e)) row.push_back(e);
        for (auto& e : rev[v]) if (ok(e)) row.push_back(e);
        std::stable_sort(row.begin(), row.end(), [](auto& a, auto& b) { return a.first < b.first; });
        dd.clear();
        for (auto& e : row) {
            if (!dd.empty() && dd.back().first == e.first) {
                dd.back().second = std::fmax(dd.back().second, e.second);
                continue;
            }
            dd.push_back(e);
        }
        auto total_key = [](float f) {
            int32_t i;
            std::memcpy(&i, &f, 4);
            return i ^ (int32_t)(((uint32_t)(i >> 31)) >> 1);
        };
        std::sort(dd.begin(), dd.end(), [&](auto& a, auto& b) {
            int32_t ka = total_key(a.second), kb = total_key(b.second);
            if (ka != kb) return ka < kb;
            return a.first < b.first;
        });
        if (dd.size() > max_degree) dd.resize(max_degree);
        std::sort(dd.begin(), dd.end(), [](auto& a, auto& b) { return a.first < b.first; });
        for (auto& e : dd) g.indices.push_back(e.first);
        g.indptr.push_back((int32_t)g.indices.size());
    }
    return g;
}

// from_dense_diversified (search_graph.rs:263-360).  Reverse rows are ordered with a stable sort:
// identical to Rust 1.75's sort_unstable_by for rows of <= 20 entries (insertion sort) and for any
// row without exactly equal distances; pdqsort tie order on longer rows is not reproduced.
SearchGraph from_dense_diversified(const int32_t* idx, const float* dist, const float* data,
                                   size_t n, size_t k, size_t dim, int metric, float prob,
                                   float mult, int threads, float* min_distance_out) {
    const size_t max_degree = (size_t)std::round(mult * (float)k);
    std::vector<std::vector<std::pair<int32_t, float>>> fwd(n), rev(n);
    const int T = std::max(1, threads);
    std::vector<size_t> ident(k);
    for (size_t i = 0; i < k; ++i) ident[i] = i;
    par_for(T, [&](int t) {
        FastRng prng(0x5eed0000ULL + (uint64_t)t);
        for (size_t i = t; i < n; i += T)
            fwd[i] = diversify(idx + i * k, dist + i * k, ident, true, data, dim, metric, prob,
                               max_degree, prng);
    });
    // forward CSR + transpose_csr (search_graph.rs:298-313, 526-567)
    std::vector<int32_t> deg(n, 0);
    for (size_t v = 0; v < n; ++v)
        for (auto& e : fwd[v]) deg[e.first]++;
    std::vector<int64_t> tptr(n + 1, 0);
    for (size_t v = 0; v < n; ++v) tptr[v + 1] = tptr[v] + deg[v];
    std::vector<int32_t> tidx(tptr[n]);
    std::vector<float> tdat(tptr[n]);
    std::vector<int64_t> cur(tptr.begin(), tptr.end() - 1);
    for (size_t v = 0; v < n; ++v)
        for (auto& e : fwd[v]) {
            int64_t pos = cur[e.first]++;
            tidx[pos] = (int32_t)v;
            tdat[pos] = e.second;
        }
    par_for(T, [&](int t) {
        FastRng prng(0x5eed1000ULL + (uint64_t)t);
        std::vector<size_t> order;
        for (size_t i = t; i < n; i += T) {
            const int64_t s = tptr[i], e = tptr[i + 1];
            if (s == e) continue;
            order.resize(e - s);
            for (int64_t q = 0; q < e - s; ++q) order[q] = q;
            const float* rd = tdat.data() + s;
            std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return rd[a] < rd[b]; });
            rev[i] = diversify(tidx.data() + s, rd, order, false, data, dim, metric, prob,
                               max_degree, prng);
        }
    });
    SearchGraph g = weighted_union_pruned(fwd, rev, n, max_degree);
    float md = std::numeric_limits<float>::infinity();
    for (size_t v = 0; v < n; ++v)
        for (auto& e : fwd[v]) {
            float d = e.second == 0.0f ? std::numeric_limits<float>::epsilon() : e.second;
            if (std::isfinite(d)) md = std::fmin(md, d);
        }
    if (min_distance_out) *min_distance_out = md;
    return g;
}

// SearchGraph::reorder (search_graph.rs:199-229).
void SearchGraph::reorder(const std::vector<size_t>& order) {
    const size_t n = n_vertices;
    std::vector<size_t> inv(n);
    for (size_t np = 0; np < n; ++np) inv[order[np]] = np;
    std::vector<int32_t> ni, nptr{0};
    for (size_t nv = 0; nv < n; ++nv) {
        const size_t ov = order[nv];
        for (int32_t q = indptr[ov]; q < indptr[ov + 1]; ++q)
            if (indices[q] >= 0) ni.push_back((int32_t)inv[indices[q]]);
        nptr.push_back((int32_t)ni.size());
    }
    indptr = std::move(nptr);
    indices = std::move(ni);
}

// ---- BinaryHeap (Rust std) and BoundedHeap ------------------------------------------------------
static inline int ord_cmp(const std::pair<float, int32_t>& a, const std::pair<float, int32_t>& b) {
    // (OrderedFloat<f32>, i32) lexicographic; no NaN distances reach the heap.
    if (a.first < b.first) return -1;
    if (a.first > b.first) return 1;
    return a.second < b.second ? -1 : (a.second > b.second ? 1 : 0);
}

void BoundedHeap::std_push(std::pair<float, int32_t> item) {
    data.push_back(item);
    // sift_up(0, old_len) with a hole
    size_t pos = data.size() - 1;
    auto elem = data[pos];
    while (pos > 0) {
        size_t parent = (pos - 1) / 2;
        if (ord_cmp(elem, data[parent]) <= 0) break;
        data[pos] = data[parent];
        pos = parent;
    }
    data[pos] = elem;
}

std::pair<float, int32_t> BoundedHeap::std_pop() {
    auto item = data.back();
    data.pop_back();
    if (!data.empty()) {
        std::swap(item, data[0]);
        // sift_down_to_bottom(0) then sift_up(0, pos)
        const size_t end = data.size();
        size_t pos = 0;
        auto elem = data[0];
        size_t child = 1;
        while (end >= 2 && child <= end - 2) {
            child += (ord_cmp(data[child], data[child + 1]) <= 0) ? 1 : 0;
            data[pos] = data[child];
            pos = child;
            child = 2 * pos + 1;
        }
        if (child == end - 1) {
            data[pos] = data[child];
            pos = child;
        }
        while (pos > 0) {
            size_t parent = (pos - 1) / 2;
            if (ord_cmp(elem, data[parent]) <= 0) break;
            data[pos] = data[parent];
            pos = parent;
        }
        data[pos] = elem;
    }
    return item;
}

// BoundedHeap::push (candidate_heap.rs:97-113).
bool BoundedHeap::push(float d, int32_t idx) {
    if (data.size() < k) {
        std_push({d, idx});
        return true;
    }
    if (!data.empty() && d < data[0].first) {
        std_pop();
        std_push({d, idx});
        return true;
    }
    return false;
}

float BoundedHeap::max_distance() const {
    return data.empty() ? std::numeric_limits<float>::infinity() : data[0].first;
}

// into_sorted (candidate_heap.rs:140-148): stable sort of the internal layout by distance only.
void BoundedHeap::into_sorted(std::vector<int32_t>& idx, std::vector<float>& dist) const {
    auto v = data;
    std::stable_sort(v.begin(), v.end(), [](auto& a, auto& b) { return a.first < b.first; });
    idx.clear();
    dist.clear();
    for (auto& e : v) {
        idx.push_back(e.second);
        dist.push_back(e.first);
    }
}

// ---- index build + query ------------------------------------------------------------------------

// build_with_distance (core/index.rs:601-733).
Index build_index(const float* data, size_t n, size_t dim, int metric, int corr,
                  const IndexParams& ip, int threads) {
    Params p;
    p.n_neighbors = ip.n_neighbors;
    p.n_trees = ip.n_trees;
    p.leaf_size = ip.leaf_size > 0 ? (size_t)ip.leaf_size
                                    : std::max<size_t>(60, std::min<size_t>(256, 5 * ip.n_neighbors));
    p.max_candidates = ip.max_candidates > 0 ? (size_t)ip.max_candidates
                                             : std::min<size_t>(ip.n_neighbors, 60);
    p.n_iters = ip.n_iters > 0 ? (size_t)ip.n_iters
                               : std::max<size_t>(5, (size_t)std::ceil(std::log2((double)n)));
    p.delta = ip.delta;
    p.angular = ip.angular;
    p.max_depth = 200;
    FastRng rng(ip.random_seed);
    Index ix;
    std::vector<FlatTree> forest;
    NeighborHeap g = nn_descent(data, n, dim, metric, p, rng, threads, &forest, &ix.trace);
    g.sort_all();
    ix.search_graph = from_dense_diversified(g.indices.data(), g.distances.data(), data, n,
                                             ip.n_neighbors, dim, metric, ip.diversify_prob,
                                             ip.pruning_degree_multiplier, threads, &ix.min_distance);
    ix.neighbor_distances = g.distances;
    if (corr != CORR_NONE)
        for (auto& d : ix.neighbor_distances) d = correct(corr, d);
    ix.neighbor_indices = g.indices;
    const size_t nst = std::min(ip.n_search_trees, forest.size());
    for (size_t t = 0; t < nst; ++t) ix.search_trees.push_back(std::move(forest[t]));
    if (!ix.search_trees.empty()) {
        for (int32_t v : ix.search_trees[0].indices) ix.vertex_order.push_back((size_t)v);
    } else {
        for (size_t i = 0; i < n; ++i) ix.vertex_order.push_back(i);
    }
    std::vector<size_t> old_to_new(n);
    for (size_t i = 0; i < n; ++i) old_to_new[ix.vertex_order[i]] = i;
    ix.data.resize(n * dim);
    for (size_t i = 0; i < n; ++i)
        std::memcpy(&ix.data[i * dim], data + ix.vertex_order[i] * dim, dim * sizeof(float));
    ix.search_graph.reorder(ix.vertex_order);
    for (auto& t : ix.search_trees)
        for (auto& v : t.indices) v = (int32_t)old_to_new[v];
    ix.n_points = n;
    ix.dim = dim;
    ix.n_neighbors = ip.n_neighbors;
    ix.metric = metric;
    ix.corr = corr;
    ix.search_tree_leaf_budget = ip.search_tree_leaf_budget;
    ix.rng_seed = ip.random_seed;
    return ix;
}

// The body of search (greedy.rs:203-241) after the tree descent: the seed ids are the leaf points in
// the order initialize_fp32 (greedy.rs:70-99) visits them; then fill_random (greedy.rs:148-163) and
// traverse (greedy.rs:165-201); the result in into_sorted order (candidate_heap.rs:140-148).
void greedy_from_seeds(const float* q, const float* data, size_t dim, int metric, size_t n_vertices,
                       const int32_t* indptr, const int32_t* indices, const std::vector<int32_t>& seed_ids,
                       size_t k, float eps, float min_distance, FastRng& rng, std::vector<int32_t>& ri,
                       std::vector<float>& rd) {
    greedy_core([&](int32_t i) { return distance(metric, q, data + (size_t)i * dim, dim); }, n_vertices, indptr,
                indices, seed_ids, k, eps, min_distance, rng, ri, rd);
}

// The search of greedy.rs:70-241 with any candidate distance (greedy.rs:101-146, 365-408 take the int8 one).
void greedy_core(const std::function<float(int32_t)>& dist_to, size_t n_vertices, const int32_t* indptr,
                 const int32_t* indices, const std::vector<int32_t>& seed_ids, size_t k, float eps,
                 float min_distance, FastRng& rng, std::vector<int32_t>& ri, std::vector<float>& rd) {
    std::vector<uint8_t> visited(n_vertices, 0);
    BoundedHeap result(k);
    using E = std::pair<float, int32_t>;
    auto gt = [](const E& a, const E& b) { return ord_cmp(a, b) > 0; };
    std::priority_queue<E, std::vector<E>, decltype(gt)> seeds(gt);  // min-heap on (d, idx)
    auto check_and_mark = [&](int32_t i) {
        bool was = visited[i] != 0;
        visited[i] = 1;
        return was;
    };
    // initialize_fp32 (greedy.rs:70-99)
    for (int32_t idx : seed_ids)
        if (idx >= 0 && !check_and_mark(idx)) {
            float d = dist_to(idx);
            result.push(d, idx);
            seeds.push({d, idx});
        }
    // fill_random (greedy.rs:148-163)
    const size_t nfill = k > seeds.size() ? k - seeds.size() : 0;
    for (size_t r = 0; r < nfill; ++r) {
        int32_t idx = (int32_t)rng.next_index(n_vertices);
        if (!check_and_mark(idx)) {
            float d = dist_to(idx);
            result.push(d, idx);
            seeds.push({d, idx});
        }
    }
    // traverse (greedy.rs:165-201)
    float maxd = result.max_distance();
    float bound = maxd + eps * (maxd - min_distance);
    while (!seeds.empty()) {
        E top = seeds.top();
        seeds.pop();
        if (result.full() && top.first >= bound) break;
        const int32_t v = top.second;
        for (int32_t z = indptr[v]; z < indptr[v + 1]; ++z) {
            const int32_t nb = indices[z];
            if (nb < 0 || check_and_mark(nb)) continue;
            float d = dist_to(nb);
            if (d < bound) {
                bool ins = result.push(d, nb);
                seeds.push({d, nb});
                if (ins) {
                    maxd = result.max_distance();
                    bound = maxd + eps * (maxd - min_distance);
                }
            }
        }
    }
    result.into_sorted(ri, rd);
}

// The seed ids of a query: every tree's leaves, in order (initialize_fp32, greedy.rs:70-99), descending the int8
// routing trees with the query's codes and scale when there are any (initialize_quantized, greedy.rs:101-146).
// Distances draw nothing from rng, so the descents may all run before the first distance.
void Index::seed_ids(const float* q, const int8_t* tree_codes, float tree_scale, FastRng& rng,
                     std::vector<int32_t>& ids) const {
    std::vector<std::pair<size_t, size_t>> leaves;
    if (!quantized_trees.empty()) {
        for (const auto& t : quantized_trees) {
            t.search_leaves(tree_codes, tree_scale, rng, search_tree_leaf_budget, leaves);
            for (auto [s, e] : leaves)
                for (size_t z = s; z < e; ++z) ids.push_back(t.indices[z]);
        }
    } else {
        for (const auto& t : search_trees) {
            t.search_leaves(q, rng, search_tree_leaf_budget, leaves);
            for (auto [s, e] : leaves)
                for (size_t z = s; z < e; ++z) ids.push_back(t.indices[z]);
        }
    }
}

// process_query_with_workspace (core/index.rs:157-202) around search (greedy.rs:203-241).
void Index::query_one(const float* q, size_t qid, size_t k, float eps, int32_t* out_idx,
                      float* out_dist) const {
    FastRng rng(rng_seed + (uint64_t)qid);
    std::vector<int32_t> seed_ids;
    std::vector<int8_t> codes(dim);
    const float qs = quantized_trees.empty() ? 0.f : quantize_query_i8(q, dim, codes.data());
    this->seed_ids(q, codes.data(), qs, rng, seed_ids);
    std::vector<int32_t> ri;
    std::vector<float> rd;
    greedy_from_seeds(q, data.data(), dim, metric, search_graph.n_vertices, search_graph.indptr.data(),
                      search_graph.indices.data(), seed_ids, k, eps, min_distance, rng, ri, rd);
    for (size_t j = 0; j < k; ++j) {
        if (j < ri.size()) {
            out_idx[j] = ri[j] >= 0 ? (int32_t)vertex_order[ri[j]] : ri[j];
            out_dist[j] = correct(corr, rd[j]);
        } else {
            out_idx[j] = -1;
            out_dist[j] = std::numeric_limits<float>::infinity();
        }
    }
}

// QuantizedI8Index::new (crates/rinnd/src/lib.rs:158-224): per-vector codes round(x * 127 / max|row|), or the
// global multiplier 127 / max|data| for sq8u, clamped to [-127, 127]; per row the inverse norm of the codes
// (int8), the global dequantisation scale 1 / multiplier (sq8u) or the row's 1 / multiplier (sq8p).
I8Codes encode_i8(const float* data, size_t n, size_t dim, int encoding) {
    I8Codes c;
    c.encoding = encoding;
    c.dim = dim;
    c.codes.resize(n * dim);
    c.inv_norms.resize(n);
    float gmax = 0.f;
    for (size_t i = 0; i < n * dim; ++i) gmax = std::max(gmax, std::fabs(data[i]));
    c.global_multiplier = gmax > 0.f ? 127.0f / gmax : 0.f;
    c.global_dequant_scale = c.global_multiplier > 0.f ? 1.0f / c.global_multiplier : 0.f;
    for (size_t r = 0; r < n; ++r) {
        const float* row = data + r * dim;
        float mult = c.global_multiplier;
        if (encoding != I8_GLOBAL_SYMMETRIC) {
            float mx = 0.f;
            for (size_t j = 0; j < dim; ++j) mx = std::max(mx, std::fabs(row[j]));
            mult = mx > 0.f ? 127.0f / mx : 0.f;
        }
        int8_t* code = c.codes.data() + r * dim;
        int32_t norm_sq = 0;
        for (size_t j = 0; j < dim; ++j) {
            code[j] = (int8_t)std::min(127.0f, std::max(-127.0f, std::round(row[j] * mult)));
            norm_sq += (int32_t)code[j] * (int32_t)code[j];
        }
        if (encoding == I8_PER_VECTOR_COSINE)
            c.inv_norms[r] = norm_sq > 0 ? 1.0f / std::sqrt((float)norm_sq) : 0.f;
        else if (encoding == I8_PER_VECTOR_SYMMETRIC)
            c.inv_norms[r] = mult > 0.f ? 1.0f / mult : 0.f;
        else
            c.inv_norms[r] = c.global_dequant_scale;
    }
    return c;
}

// QuantizedI8Index::quantize_query (lib.rs:325-358): the query's codes under the index's encoding and its inverse
// norm (int8) or dequantisation scale (sq8u: the global one, sq8p: 1 / its multiplier).
float quantize_query_codes(const I8Codes& c, const float* q, int8_t* codes) {
    float mult = c.global_multiplier;
    if (c.encoding != I8_GLOBAL_SYMMETRIC) {
        float mx = 0.f;
        for (size_t j = 0; j < c.dim; ++j) mx = std::max(mx, std::fabs(q[j]));
        mult = mx > 0.f ? 127.0f / mx : 0.f;
    }
    int32_t norm_sq = 0;
    for (size_t j = 0; j < c.dim; ++j) {
        codes[j] = (int8_t)std::min(127.0f, std::max(-127.0f, std::round(q[j] * mult)));
        norm_sq += (int32_t)codes[j] * (int32_t)codes[j];
    }
    if (c.encoding == I8_PER_VECTOR_COSINE) return norm_sq > 0 ? 1.0f / std::sqrt((float)norm_sq) : 0.f;
    if (c.encoding == I8_GLOBAL_SYMMETRIC) return c.global_dequant_scale;
    return mult > 0.f ? 1.0f / mult : 0.f;
}

// greedy.rs:381-387: 1 - clamp(i8 dot as f32 * query inverse norm * row inverse norm, -1, 1).
float i8_cosine_distance(const int8_t* q, float q_inv_norm, const int8_t* x, float x_inv_norm, size_t dim) {
    const float similarity = (float)quantized_i8_dot(q, x, dim) * q_inv_norm * x_inv_norm;
    return 1.0f - std::min(1.0f, std::max(-1.0f, similarity));
}

// query_one_quantized_i8_with_workspace (core/index.rs:283-330): the search over the codes with candidate width
// resolve_quantized_candidate_width (index.rs:55-77), then rerank_quantized_candidates (index.rs:103-124) of the
// first rerank_width candidates in FP32 when should_rerank_quantized (index.rs:80-88) says so.  The trees
// descend with the FP32 query, or with the query's codes (and its tree scale max|q| / 127, lib.rs:303-313)
// when the index routes by int8 trees.
void Index::query_one_quantized(const I8Codes& c, const float* q, size_t qid, size_t k, size_t candidate_width,
                                size_t rerank_width, float eps, int32_t* out_idx, float* out_dist) const {
    const size_t search_k = candidate_width == 0 ? std::max(k, rerank_width) : candidate_width;
    const bool rerank = candidate_width == 0 ? rerank_width > k : rerank_width > 0;
    std::vector<int8_t> qc(dim);
    const float qinv = quantize_query_codes(c, q, qc.data());
    float tree_scale = 0.f;
    if (c.encoding == I8_GLOBAL_SYMMETRIC) {
        tree_scale = c.global_dequant_scale;
    } else {
        float mx = 0.f;
        for (size_t j = 0; j < dim; ++j) mx = std::max(mx, std::fabs(q[j]));
        tree_scale = mx > 0.f ? mx / 127.0f : 0.f;
    }
    (void)qid;
    FastRng rng(rng_seed);  // index.rs:298: every quantized query starts from the index's seed itself
    std::vector<int32_t> seeds;
    this->seed_ids(q, qc.data(), tree_scale, rng, seeds);
    std::vector<int32_t> ri;
    std::vector<float> rd;
    greedy_core([&](int32_t i) {
                    return i8_cosine_distance(qc.data(), qinv, c.codes.data() + (size_t)i * dim, c.inv_norms[i], dim);
                },
                search_graph.n_vertices, search_graph.indptr.data(), search_graph.indices.data(), seeds, search_k,
                eps, min_distance, rng, ri, rd);
    if (rerank) {
        std::vector<std::pair<float, int32_t>> rr;
        for (size_t j = 0; j < ri.size() && rr.size() < rerank_width; ++j)
            if (ri[j] >= 0) rr.push_back({distance(metric, q, data.data() + (size_t)ri[j] * dim, dim), ri[j]});
        auto key = [](float d) {  // f32::total_cmp
            int32_t b;
            std::memcpy(&b, &d, 4);
            return b ^ (int32_t)(((uint32_t)(b >> 31)) >> 1);
        };
        std::sort(rr.begin(), rr.end(), [&](const auto& a, const auto& b) {
            return key(a.first) != key(b.first) ? key(a.first) < key(b.first) : a.second < b.second;
        });
        if (rr.size() > k) rr.resize(k);
        ri.clear();
        rd.clear();
        for (auto& [d, i] : rr) {
            ri.push_back(i);
            rd.push_back(d);
        }
    }
    for (size_t j = 0; j < k; ++j) {
        if (j < ri.size()) {
            out_idx[j] = ri[j] >= 0 ? (int32_t)vertex_order[ri[j]] : ri[j];
            out_dist[j] = rd[j];
        } else {
            out_idx[j] = -1;
            out_dist[j] = std::numeric_limits<float>::infinity();
        }
    }
}

// NNDescentIndex::query (core/index.rs:406-439): query ids 0..nq, results independent of threads.
void Index::query(const float* queries, size_t nq, size_t k, float eps, int32_t* out_idx,
                  float* out_dist, int threads) const {
    const int T = std::max(1, std::min<int>(threads, (int)std::max<size_t>(nq, 1)));
    par_for(T, [&](int t) {
        for (size_t i = t; i < nq; i += T)
            query_one(queries + i * dim, i, k, eps, out_idx + i * k, out_dist + i * k);
    });
}

}  // namespace oracle

// =================================================================================================
// extern "C" surface used by the tests (ctypes) and by bench.py's CPU-baseline leg.
// =================================================================================================
using namespace oracle;

extern "C" {

void orc_rng_seed(uint64_t seed, uint64_t st[4]) {
    FastRng r(seed);
    std::memcpy(st, r.s, 32);
}
static FastRng rng_from(const uint64_t st[4]) {
    FastRng r(0);
    std::memcpy(r.s, st, 32);
    return r;
}
void orc_rng_u64(uint64_t st[4], int64_t count, uint64_t* out) {
    FastRng r = rng_from(st);
    for (int64_t i = 0; i < count; ++i) out[i] = r.next_u64();
    std::memcpy(st, r.s, 32);
}
void orc_rng_float(uint64_t st[4], int64_t count, float* out) {
    FastRng r = rng_from(st);
    for (int64_t i = 0; i < count; ++i) out[i] = r.next_float();
    std::memcpy(st, r.s, 32);
}
void orc_rng_bool(uint64_t st[4], int64_t count, uint8_t* out) {
    FastRng r = rng_from(st);
    for (int64_t i = 0; i < count; ++i) out[i] = r.next_bool() ? 1 : 0;
    std::memcpy(st, r.s, 32);
}
void orc_rng_index(uint64_t st[4], uint64_t n, int64_t count, uint64_t* out) {
    FastRng r = rng_from(st);
    for (int64_t i = 0; i < count; ++i) out[i] = r.next_index(n);
    std::memcpy(st, r.s, 32);
}

float orc_distance(int metric, const float* a, const float* b, int64_t dim) {
    return distance(metric, a, b, (size_t)dim);
}
void orc_pair_distances(int metric, const float* data, int64_t dim, const int32_t* a,
                        const int32_t* b, int64_t n_pairs, float* out) {
    for (int64_t p = 0; p < n_pairs; ++p)
        out[p] = distance(metric, data + (size_t)a[p] * dim, data + (size_t)b[p] * dim, (size_t)dim);
}
float orc_correct(int corr, float d) { return correct(corr, d); }

// Single-row heap operations for the reference's heap unit tests.
int orc_heap_checked_flagged_push(int32_t* idx, float* dist, uint8_t* flags, int64_t k, int32_t nb,
                                  float d, int is_new) {
    NeighborHeap h(1, (size_t)k);
    std::memcpy(h.indices.data(), idx, k * 4);
    std::memcpy(h.distances.data(), dist, k * 4);
    std::memcpy(h.flags.data(), flags, k);
    bool r = h.checked_flagged_push(0, nb, d, is_new != 0);
    std::memcpy(idx, h.indices.data(), k * 4);
    std::memcpy(dist, h.distances.data(), k * 4);
    std::memcpy(flags, h.flags.data(), k);
    return r ? 1 : 0;
}
int orc_heap_simple_push(int32_t* idx, float* dist, uint8_t* flags, int64_t k, int32_t nb, float d) {
    NeighborHeap h(1, (size_t)k);
    std::memcpy(h.indices.data(), idx, k * 4);
    std::memcpy(h.distances.data(), dist, k * 4);
    std::memcpy(h.flags.data(), flags, k);
    bool r = h.simple_push(0, nb, d);
    std::memcpy(idx, h.indices.data(), k * 4);
    std::memcpy(dist, h.distances.data(), k * 4);
    std::memcpy(flags, h.flags.data(), k);
    return r ? 1 : 0;
}
void orc_heap_unchecked_flagged_push(int32_t* idx, float* dist, uint8_t* flags, int64_t n,
                                     int64_t k, int64_t p, int32_t nb, float d, int is_new) {
    NeighborHeap h((size_t)n, (size_t)k);
    std::memcpy(h.indices.data(), idx, n * k * 4);
    std::memcpy(h.distances.data(), dist, n * k * 4);
    std::memcpy(h.flags.data(), flags, n * k);
    h.unchecked_flagged_push((size_t)p, nb, d, is_new != 0);
    std::memcpy(idx, h.indices.data(), n * k * 4);
    std::memcpy(dist, h.distances.data(), n * k * 4);
    std::memcpy(flags, h.flags.data(), n * k);
}
void orc_cand_push(float* pri, int32_t* idx, int64_t n, float priority, int32_t index) {
    cand_push(pri, idx, (size_t)n, priority, index);
}
void orc_sort_all(int32_t* idx, float* dist, uint8_t* flags, int64_t n, int64_t k) {
    NeighborHeap h((size_t)n, (size_t)k);
    std::memcpy(h.indices.data(), idx, n * k * 4);
    std::memcpy(h.distances.data(), dist, n * k * 4);
    std::memcpy(h.flags.data(), flags, n * k);
    h.sort_all();
    std::memcpy(idx, h.indices.data(), n * k * 4);
    std::memcpy(dist, h.distances.data(), n * k * 4);
    std::memcpy(flags, h.flags.data(), n * k);
}

// ---- forest handle ----
struct OrcForest { std::vector<FlatTree> trees; };

void* orc_build_forest(const float* data, int64_t n, int64_t dim, int64_t n_trees,
                       int64_t leaf_size, uint64_t st[4], int angular, int64_t max_depth,
                       int threads) {
    FastRng r = rng_from(st);
    auto* f = new OrcForest;
    f->trees = build_rp_forest(data, n, dim, n_trees, leaf_size, r, angular != 0, max_depth, threads);
    std::memcpy(st, r.s, 32);
    return f;
}
void* orc_build_tree(const float* data, int64_t n, int64_t dim, int64_t leaf_size, uint64_t st[4],
                     int angular, int64_t max_depth) {
    FastRng r = rng_from(st);
    auto* f = new OrcForest;
    f->trees.push_back(build_rp_tree(data, n, dim, leaf_size, r, angular != 0, max_depth));
    std::memcpy(st, r.s, 32);
    return f;
}
void* orc_forest_from_arrays(int64_t dim, int64_t n_nodes, const float* hp, const float* off,
                             const int32_t* ch, int64_t n_idx, const int32_t* idx) {
    auto* f = new OrcForest;
    FlatTree t;
    t.dim = dim;
    t.n_nodes = n_nodes;
    t.hyperplanes.assign(hp, hp + n_nodes * dim);
    t.offsets.assign(off, off + n_nodes);
    t.children.assign(ch, ch + 2 * n_nodes);
    t.indices.assign(idx, idx + n_idx);
    f->trees.push_back(std::move(t));
    return f;
}
void orc_forest_free(void* h) { delete (OrcForest*)h; }
int64_t orc_forest_n_trees(void* h) { return (int64_t)((OrcForest*)h)->trees.size(); }
int64_t orc_forest_n_nodes(void* h, int64_t t) { return (int64_t)((OrcForest*)h)->trees[t].n_nodes; }
int64_t orc_forest_n_indices(void* h, int64_t t) {
    return (int64_t)((OrcForest*)h)->trees[t].indices.size();
}
void orc_forest_copy(void* h, int64_t t, float* hp, float* off, int32_t* ch, int32_t* idx) {
    const FlatTree& tr = ((OrcForest*)h)->trees[t];
    if (hp) std::memcpy(hp, tr.hyperplanes.data(), tr.hyperplanes.size() * 4);
    if (off) std::memcpy(off, tr.offsets.data(), tr.offsets.size() * 4);
    if (ch) std::memcpy(ch, tr.children.data(), tr.children.size() * 4);
    if (idx) std::memcpy(idx, tr.indices.data(), tr.indices.size() * 4);
}
int64_t orc_forest_leaves(void* h, int64_t t, int64_t* out_pairs, int64_t cap) {
    auto lv = ((OrcForest*)h)->trees[t].get_all_leaves();
    for (size_t i = 0; i < lv.size() && (int64_t)i < cap; ++i) {
        out_pairs[2 * i] = (int64_t)lv[i].first;
        out_pairs[2 * i + 1] = (int64_t)lv[i].second;
    }
    return (int64_t)lv.size();
}
int64_t orc_tree_search_leaves(void* h, int64_t t, const float* point, uint64_t st[4],
                               int64_t budget, int64_t* out_pairs, int64_t cap, int64_t* n_evals) {
    FastRng r = rng_from(st);
    std::vector<std::pair<size_t, size_t>> lv;
    size_t ev = ((OrcForest*)h)->trees[t].search_leaves(point, r, (size_t)budget, lv);
    std::memcpy(st, r.s, 32);
    for (size_t i = 0; i < lv.size() && (int64_t)i < cap; ++i) {
        out_pairs[2 * i] = (int64_t)lv[i].first;
        out_pairs[2 * i + 1] = (int64_t)lv[i].second;
    }
    if (n_evals) *n_evals = (int64_t)ev;
    return (int64_t)lv.size();
}

// ---- NN-Descent pieces on caller arrays ----
static NeighborHeap heap_in(const int32_t* idx, const float* dist, const uint8_t* flags, int64_t n,
                            int64_t k) {
    NeighborHeap h((size_t)n, (size_t)k);
    std::memcpy(h.indices.data(), idx, n * k * 4);
    std::memcpy(h.distances.data(), dist, n * k * 4);
    std::memcpy(h.flags.data(), flags, n * k);
    return h;
}
static void heap_out(const NeighborHeap& h, int32_t* idx, float* dist, uint8_t* flags) {
    std::memcpy(idx, h.indices.data(), h.indices.size() * 4);
    std::memcpy(dist, h.distances.data(), h.distances.size() * 4);
    std::memcpy(flags, h.flags.data(), h.flags.size());
}

void orc_init_from_leaves(int32_t* idx, float* dist, uint8_t* flags, int64_t n, int64_t k,
                          void* forest, const float* data, int64_t dim, int metric, int threads) {
    NeighborHeap h = heap_in(idx, dist, flags, n, k);
    std::vector<std::vector<int32_t>> leaves;
    for (const auto& t : ((OrcForest*)forest)->trees)
        for (auto [s, e] : t.get_all_leaves())
            leaves.emplace_back(t.indices.begin() + s, t.indices.begin() + e);
    initialize_from_leaves(h, leaves, data, (size_t)dim, metric, threads);
    heap_out(h, idx, dist, flags);
}

void orc_build_candidates(int32_t* idx, float* dist, uint8_t* flags, int64_t n, int64_t k,
                          int64_t C, uint64_t st[4], int threads, int32_t* out_new,
                          int32_t* out_old) {
    NeighborHeap h = heap_in(idx, dist, flags, n, k);
    FastRng r = rng_from(st);
    CandidateSets cs = build_candidates(h, (size_t)C, r, threads);
    std::memcpy(st, r.s, 32);
    heap_out(h, idx, dist, flags);
    std::memcpy(out_new, cs.new_indices.data(), n * C * 4);
    std::memcpy(out_old, cs.old_indices.data(), n * C * 4);
}

int64_t orc_update_iteration(int32_t* idx, float* dist, uint8_t* flags, int64_t n, int64_t k,
                             const int32_t* new_c, const int32_t* old_c, int64_t C,
                             const float* data, int64_t dim, int metric, int threads) {
    NeighborHeap h = heap_in(idx, dist, flags, n, k);
    CandidateSets cs;
    cs.n_vertices = n;
    cs.max_candidates = C;
    cs.new_indices.assign(new_c, new_c + n * C);
    cs.old_indices.assign(old_c, old_c + n * C);
    size_t ch = update_iteration(h, cs, data, (size_t)dim, metric, threads);
    heap_out(h, idx, dist, flags);
    return (int64_t)ch;
}

// nn_descent on caller arrays; returns the forest handle (free with orc_forest_free).
void* orc_nn_descent(const float* data, int64_t n, int64_t dim, int metric, int64_t k,
                     int64_t n_trees, int64_t leaf_size, int64_t max_candidates, int64_t n_iters,
                     float delta, int angular, int64_t max_depth, uint64_t st[4], int threads,
                     int32_t* out_idx, float* out_dist, uint8_t* out_flags, int64_t* out_changes,
                     int64_t* out_iters, double* out_ms4) {
    Params p;
    p.n_neighbors = k;
    p.n_trees = n_trees;
    p.leaf_size = leaf_size;
    p.max_candidates = max_candidates;
    p.n_iters = n_iters;
    p.delta = delta;
    p.angular = angular != 0;
    p.max_depth = max_depth;
    FastRng r = rng_from(st);
    auto* f = new OrcForest;
    BuildTrace tr;
    NeighborHeap h = nn_descent(data, n, dim, metric, p, r, threads, &f->trees, &tr);
    std::memcpy(st, r.s, 32);
    heap_out(h, out_idx, out_dist, out_flags);
    for (size_t i = 0; i < tr.changes.size() && i < 64; ++i) out_changes[i] = (int64_t)tr.changes[i];
    *out_iters = (int64_t)tr.changes.size();
    if (out_ms4) {
        out_ms4[0] = tr.ms_forest;
        out_ms4[1] = tr.ms_init;
        out_ms4[2] = tr.ms_candidates;
        out_ms4[3] = tr.ms_updates;
    }
    return f;
}

int64_t orc_search_graph(const int32_t* idx, const float* dist, const float* data, int64_t n,
                         int64_t k, int64_t dim, int metric, float prob, float mult, int threads,
                         int32_t* out_indptr, int32_t* out_indices, int64_t cap,
                         float* out_min_distance) {
    SearchGraph g = from_dense_diversified(idx, dist, data, n, k, dim, metric, prob, mult, threads,
                                           out_min_distance);
    std::memcpy(out_indptr, g.indptr.data(), (n + 1) * 4);
    const int64_t nnz = (int64_t)g.indices.size();
    std::memcpy(out_indices, g.indices.data(), std::min(nnz, cap) * 4);
    return nnz;
}

// weighted_union_pruned on CSR-encoded (idx, dist) rows.
int64_t orc_weighted_union(int64_t n, int64_t max_degree, const int64_t* fptr, const int32_t* fidx,
                           const float* fdst, const int64_t* rptr, const int32_t* ridx,
                           const float* rdst, int32_t* out_indptr, int32_t* out_indices,
                           int64_t cap) {
    std::vector<std::vector<std::pair<int32_t, float>>> fw(n), rv(n);
    for (int64_t v = 0; v < n; ++v) {
        for (int64_t q = fptr[v]; q < fptr[v + 1]; ++q) fw[v].emplace_back(fidx[q], fdst[q]);
        for (int64_t q = rptr[v]; q < rptr[v + 1]; ++q) rv[v].emplace_back(ridx[q], rdst[q]);
    }
    SearchGraph g = weighted_union_pruned(fw, rv, n, max_degree);
    std::memcpy(out_indptr, g.indptr.data(), (n + 1) * 4);
    const int64_t nnz = (int64_t)g.indices.size();
    std::memcpy(out_indices, g.indices.data(), std::min(nnz, cap) * 4);
    return nnz;
}

// BoundedHeap replay: push (d, idx) pairs with capacity k; returns the into_sorted length.
int64_t orc_bounded_heap(int64_t k, const float* d, const int32_t* ix, int64_t n, uint8_t* inserted,
                         int32_t* out_idx, float* out_dist) {
    BoundedHeap h((size_t)k);
    for (int64_t i = 0; i < n; ++i) {
        bool r = h.push(d[i], ix[i]);
        if (inserted) inserted[i] = r ? 1 : 0;
    }
    std::vector<int32_t> oi;
    std::vector<float> od;
    h.into_sorted(oi, od);
    std::memcpy(out_idx, oi.data(), oi.size() * 4);
    std::memcpy(out_dist, od.data(), od.size() * 4);
    return (int64_t)oi.size();
}

// ---- index handle ----
void* orc_index_build(const float* data, int64_t n, int64_t dim, int metric, int corr, int64_t k,
                      int64_t n_trees, int64_t n_search_trees, int64_t leaf_budget,
                      int64_t leaf_size, int64_t max_candidates, int64_t n_iters, float delta,
                      float diversify_prob, float mult, int angular, uint64_t seed, int threads) {
    IndexParams ip;
    ip.n_neighbors = k;
    ip.n_trees = n_trees;
    ip.n_search_trees = n_search_trees;
    ip.search_tree_leaf_budget = leaf_budget < 1 ? 1 : leaf_budget;
    ip.leaf_size = leaf_size;
    ip.max_candidates = max_candidates;
    ip.n_iters = n_iters;
    ip.delta = delta;
    ip.diversify_prob = diversify_prob;
    ip.pruning_degree_multiplier = mult;
    ip.angular = angular != 0;
    ip.random_seed = seed;
    return new Index(build_index(data, n, dim, metric, corr, ip, threads));
}
void orc_index_free(void* h) { delete (Index*)h; }
// tree_quantization = "int8": the search descends int8 copies of the search trees from now on
void orc_index_quantize_trees(void* h) {
    Index& ix = *(Index*)h;
    ix.quantized_trees.clear();
    for (const auto& t : ix.search_trees) ix.quantized_trees.push_back(QuantizedFlatTree::from_flat_tree(t));
}
// one node's int8 plane, its scale and the reconstructed margin of an int8 query (tests of the formulas)
float orc_quantized_margin(const float* hyperplane, float offset, int64_t dim, const int8_t* query, float query_scale,
                           int8_t* codes_out, float* scale_out) {
    FlatTree t;
    t.dim = (size_t)dim;
    t.n_nodes = 1;
    t.hyperplanes.assign(hyperplane, hyperplane + dim);
    t.offsets = {offset};
    t.children = {0, 0};
    const QuantizedFlatTree q = QuantizedFlatTree::from_flat_tree(t);
    std::memcpy(codes_out, q.hyperplanes.data(), (size_t)dim);
    *scale_out = q.scales[0];
    return q.margin(0, query, query_scale);
}
float orc_quantize_query_i8(const float* q, int64_t dim, int8_t* codes) { return quantize_query_i8(q, (size_t)dim, codes); }
// The quantized index's codes (QuantizedI8Index::new) of the index's (reordered) rows; freed by orc_i8_free.
void* orc_index_encode_i8(void* h, int encoding) {
    const Index& ix = *(const Index*)h;
    return new I8Codes(encode_i8(ix.data.data(), ix.n_points, ix.dim, encoding));
}
void orc_i8_free(void* c) { delete (I8Codes*)c; }
// codes (n x dim), inverse norms (n) and the global multiplier / dequantisation scale
void orc_i8_codes(void* c, int8_t* codes, float* inv_norms, float* global2) {
    const I8Codes& q = *(const I8Codes*)c;
    std::memcpy(codes, q.codes.data(), q.codes.size());
    std::memcpy(inv_norms, q.inv_norms.data(), q.inv_norms.size() * 4);
    global2[0] = q.global_multiplier;
    global2[1] = q.global_dequant_scale;
}
// standalone encoding of rows (no index) and of a query under it
void* orc_encode_i8_rows(const float* data, int64_t n, int64_t dim, int encoding) {
    return new I8Codes(encode_i8(data, (size_t)n, (size_t)dim, encoding));
}
float orc_quantize_query_codes(void* c, const float* q, int8_t* codes) {
    return quantize_query_codes(*(const I8Codes*)c, q, codes);
}
float orc_i8_cosine_distance(const int8_t* q, float qinv, const int8_t* x, float xinv, int64_t dim) {
    return i8_cosine_distance(q, qinv, x, xinv, (size_t)dim);
}
void orc_index_query_i8(void* h, void* c, const float* queries, int64_t nq, int64_t k, int64_t candidate_width,
                        int64_t rerank_width, float eps, int threads, int32_t* oi, float* od) {
    const Index& ix = *(const Index*)h;
    const I8Codes& codes = *(const I8Codes*)c;
    const int T = std::max(1, std::min<int>(threads, (int)std::max<int64_t>(nq, 1)));
    par_for(T, [&](int t) {
        for (int64_t i = t; i < nq; i += T)
            ix.query_one_quantized(codes, queries + i * ix.dim, (size_t)i, (size_t)k, (size_t)candidate_width,
                                   (size_t)rerank_width, eps, oi + i * k, od + i * k);
    });
}
// leaves of one quantized tree (t-th of a forest) for an int8 query
int64_t orc_tree_search_leaves_q(void* h, int64_t t, const int8_t* query, float query_scale, uint64_t st[4],
                                 int64_t budget, int64_t* out) {
    FastRng r;
    std::memcpy(r.s, st, 32);
    const QuantizedFlatTree q = QuantizedFlatTree::from_flat_tree(((OrcForest*)h)->trees[t]);
    std::vector<std::pair<size_t, size_t>> lv;
    q.search_leaves(query, query_scale, r, (size_t)budget, lv);
    std::memcpy(st, r.s, 32);
    for (size_t i = 0; i < lv.size(); ++i) {
        out[2 * i] = (int64_t)lv[i].first;
        out[2 * i + 1] = (int64_t)lv[i].second;
    }
    return (int64_t)lv.size();
}
void orc_index_query(void* h, const float* q, int64_t nq, int64_t k, float eps, int threads,
                     int32_t* out_idx, float* out_dist) {
    ((Index*)h)->query(q, nq, k, eps, out_idx, out_dist, threads);
}
void orc_index_query_one(void* h, const float* q, int64_t k, float eps, int32_t* out_idx,
                         float* out_dist) {
    ((Index*)h)->query_one(q, 0, k, eps, out_idx, out_dist);
}
void orc_index_neighbor_graph(void* h, int32_t* idx, float* dist) {
    Index* ix = (Index*)h;
    std::memcpy(idx, ix->neighbor_indices.data(), ix->neighbor_indices.size() * 4);
    std::memcpy(dist, ix->neighbor_distances.data(), ix->neighbor_distances.size() * 4);
}
int64_t orc_index_search_graph_nnz(void* h) { return (int64_t)((Index*)h)->search_graph.indices.size(); }
float orc_index_min_distance(void* h) { return ((Index*)h)->min_distance; }
// internal (reordered) CSR
void orc_index_search_graph(void* h, int32_t* indptr, int32_t* indices) {
    Index* ix = (Index*)h;
    std::memcpy(indptr, ix->search_graph.indptr.data(), ix->search_graph.indptr.size() * 4);
    std::memcpy(indices, ix->search_graph.indices.data(), ix->search_graph.indices.size() * 4);
}
void orc_index_vertex_order(void* h, int64_t* out) {
    Index* ix = (Index*)h;
    for (size_t i = 0; i < ix->vertex_order.size(); ++i) out[i] = (int64_t)ix->vertex_order[i];
}
int64_t orc_index_n_trees(void* h) { return (int64_t)((Index*)h)->search_trees.size(); }
void* orc_index_forest(void* h) {  // copy of the retained search trees (free with orc_forest_free)
    auto* f = new OrcForest;
    f->trees = ((Index*)h)->search_trees;
    return f;
}
void orc_index_trace(void* h, int64_t* changes, int64_t* n_iters, double* ms4) {
    const BuildTrace& tr = ((Index*)h)->trace;
    for (size_t i = 0; i < tr.changes.size() && i < 64; ++i) changes[i] = (int64_t)tr.changes[i];
    *n_iters = (int64_t)tr.changes.size();
    ms4[0] = tr.ms_forest;
    ms4[1] = tr.ms_init;
    ms4[2] = tr.ms_candidates;
    ms4[3] = tr.ms_updates;
}

// diversify_row (stop_at_neg = 1) / diversify_sorted_csr (stop_at_neg = 0) on explicit inputs; prob >= 1
// (the reference's default, no random pruning).  Returns the retained count.
int64_t orc_diversify(const int32_t* ri, const float* rd, const int64_t* order, int64_t n_order, int stop_at_neg,
                      const float* data, int64_t dim, int metric, int64_t max_degree, int32_t* out_idx,
                      float* out_dist) {
    std::vector<size_t> ord(order, order + n_order);
    FastRng prng(0);
    auto kept = diversify(ri, rd, ord, stop_at_neg != 0, data, (size_t)dim, metric, 1.0f, (size_t)max_degree, prng);
    for (size_t i = 0; i < kept.size(); ++i) {
        out_idx[i] = kept[i].first;
        out_dist[i] = kept[i].second;
    }
    return (int64_t)kept.size();
}

// The greedy search from explicit seed ids over a caller-given CSR graph (greedy_from_seeds); st is the
// query's FastRng state, advanced by fill_random.  Writes k results padded with (-1, +inf), returns the count.
int64_t orc_greedy(const float* q, const float* data, int64_t n, int64_t dim, int metric, const int32_t* indptr,
                   const int32_t* indices, const int32_t* seed_ids, int64_t n_seeds, int64_t k, float eps,
                   float min_distance, uint64_t st[4], int32_t* out_idx, float* out_dist) {
    FastRng rng = rng_from(st);
    std::vector<int32_t> seeds(seed_ids, seed_ids + n_seeds), ri;
    std::vector<float> rd;
    greedy_from_seeds(q, data, (size_t)dim, metric, (size_t)n, indptr, indices, seeds, (size_t)k, eps,
                      min_distance, rng, ri, rd);
    std::memcpy(st, rng.s, 32);
    for (int64_t j = 0; j < k; ++j) {
        out_idx[j] = j < (int64_t)ri.size() ? ri[j] : -1;
        out_dist[j] = j < (int64_t)ri.size() ? rd[j] : std::numeric_limits<float>::infinity();
    }
    return (int64_t)ri.size();
}

}  // extern "C"
