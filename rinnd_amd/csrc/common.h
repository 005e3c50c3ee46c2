// common.h - shared host/device definitions for the gfx950 NN-Descent library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/rinnd_hip.h"

#define RINND_WAVE 64

namespace rinnd {

// ---- error plumbing ------------------------------------------------------------------------------
struct Error {
    rinnd_status code;
    std::string msg;
};

#define HIP_TRY(expr)                                                                           \
    do {                                                                                        \
        hipError_t _e = (expr);                                                                 \
        if (_e != hipSuccess)                                                                   \
            throw ::rinnd::Error{RINND_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)}; \
    } while (0)

#define RINND_CHECK_ARG(cond, text)                                              \
    do {                                                                         \
        if (!(cond)) throw ::rinnd::Error{RINND_ERR_INVALID_ARG, std::string(text)}; \
    } while (0)

// ---- small helpers -----------------------------------------------------------------------------
__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
__host__ __device__ inline int64_t round_up(int64_t a, int64_t b) { return ceil_div(a, b) * b; }
// workgroups for `work` items of `block` each: at least 1, at most cap (the kernels loop over the rest)
inline unsigned grid_for(int64_t work, int block, int64_t cap = 1 << 20) {
    int64_t g = ceil_div(work, block);
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

// Wave-local LDS ordering: wait for this wave's LDS traffic, then re-converge the wave.
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

__device__ inline int lane_id() { return threadIdx.x & (RINND_WAVE - 1); }

// Moves the k-entry neighbour-heap rows of a wave's 64 lanes between global memory (row r's slot j at
// r * k + j) and LDS columns (slot j of column c at j * S + c).  Each row is handled by 16 lanes, so
// every global access covers consecutive slots of four rows instead of one slot of 64 scattered rows;
// the four rows of an access are 16 columns apart, which keeps the LDS side conflict-free for odd S.
// row: this lane's global row, or -1 for none; col0: the wave's first column.  Every lane of the wave
// calls it; LOAD ends, and the store starts, with a wave sync.
template <bool LOAD>
__device__ __forceinline__ void wave_heap_rows(int32_t row, int k, float* dist, int32_t* idx, uint8_t* flags,
                                               float* hd, int32_t* hi, uint8_t* hf, int S, int col0) {
    const int lane = threadIdx.x & 63, sub = lane & 15, g = lane >> 4;
    if (!LOAD) wave_sync();
    if (k <= 16) {
#pragma unroll 4
        for (int q = 0; q < 16; ++q) {
            const int rr = q + 16 * g;
            const int32_t rid = __shfl(row, rr);
            if (rid >= 0 && sub < k) {
                const int64_t gi = (int64_t)rid * k + sub;
                const int li = sub * S + col0 + rr;
                if (LOAD) {
                    hd[li] = dist[gi];
                    hi[li] = idx[gi];
                    hf[li] = flags[gi];
                } else {
                    dist[gi] = hd[li];
                    idx[gi] = hi[li];
                    flags[gi] = hf[li];
                }
            }
        }
    } else {
        for (int q = 0; q < 16; ++q) {
            const int rr = q + 16 * g;
            const int32_t rid = __shfl(row, rr);
            if (rid < 0) continue;
            for (int j = sub; j < k; j += 16) {
                const int64_t gi = (int64_t)rid * k + j;
                const int li = j * S + col0 + rr;
                if (LOAD) {
                    hd[li] = dist[gi];
                    hi[li] = idx[gi];
                    hf[li] = flags[gi];
                } else {
                    dist[gi] = hd[li];
                    idx[gi] = hi[li];
                    flags[gi] = hf[li];
                }
            }
        }
    }
    if (LOAD) wave_sync();
}

// The same for one array of rows with k entries (g: global rows, l: LDS columns).
template <bool LOAD, typename T>
__device__ __forceinline__ void wave_rows(int32_t row, int k, T* g, T* l, int S, int col0) {
    const int lane = threadIdx.x & 63, sub = lane & 15, grp = lane >> 4;
    if (!LOAD) wave_sync();
#pragma unroll 4
    for (int q = 0; q < 16; ++q) {
        const int rr = q + 16 * grp;
        const int32_t rid = __shfl(row, rr);
        if (rid < 0) continue;
        for (int j = sub; j < k; j += 16) {
            if (LOAD) l[j * S + col0 + rr] = g[(int64_t)rid * k + j];
            else g[(int64_t)rid * k + j] = l[j * S + col0 + rr];
        }
    }
    if (LOAD) wave_sync();
}

// A neighbour-heap push (neighbor_heap.rs:207-243) on an LDS column once the caller has shown that it is
// accepted: d is below the root (kept in a register by
// the caller) and x is not in the heap.  Returns the new root distance; evicted receives the id of the old
// root, which leaves the heap.
__device__ __forceinline__ float heap_push_col_new(float* hd, int32_t* hi, uint8_t* hf, int B, int c, int k, int32_t x,
                                                   float d, int32_t& evicted) {
    evicted = hi[c];
    int pos = 0;
    float top = d;
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
        if (pos == 0) top = dc;
        hd[pos * B + c] = dc;
        hi[pos * B + c] = right ? ir : il;
        hf[pos * B + c] = right ? fr : fl;
        pos = right ? r : l;
    }
    hd[pos * B + c] = d;
    hi[pos * B + c] = x;
    hf[pos * B + c] = 1;
    return top;
}

// A workgroup barrier that orders LDS only: unlike __syncthreads(), whose workgroup fence also waits for every
// outstanding global access, it leaves global loads in flight (prefetches issued before it keep going).  For
// kernels whose waves communicate through LDS alone.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// operand and accumulator vectors of the 16x16x16 bf16 MFMA
typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));  // A/B operand of a 16x16x32 bf16 MFMA

// Order-preserving float -> u32 map; -0.0 is folded onto +0.0 so that ties compare equal like
// Rust's partial_cmp (neighbor_heap.rs:308-310).
__host__ __device__ inline uint32_t float_order_key(float f) {
    if (f == 0.0f) f = 0.0f;
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// ---- reference-order distance kernels (lane-serial) --------------------------------------------
// Each function reproduces the operation order of the reference's AVX2 kernel: lane l of the
// 8-wide accumulator owns elements l, l+8, ...; fused multiply-adds where the reference uses
// _mm256_fmadd_ps; the horizontal sums as written in the reference; scalar tails without FMA.
// The library is compiled with -ffp-contract=off so `a*b + c` stays two roundings.

// hsum256_ps_avx (core/distance/euclidean.rs:99-110): ((l0+l4)+(l1+l5)) + ((l2+l6)+(l3+l7))
__device__ __forceinline__ float hsum8(const float* v) {
    return ((v[0] + v[4]) + (v[1] + v[5])) + ((v[2] + v[6]) + (v[3] + v[7]));
}
// horizontal_sum_avx (core/distance/alternatives.rs:94-104): ((l0+l4)+(l2+l6)) + ((l1+l5)+(l3+l7))
__device__ __forceinline__ float hsum8_alt(const float* v) {
    return ((v[0] + v[4]) + (v[2] + v[6])) + ((v[1] + v[5]) + (v[3] + v[7]));
}

// l2_sqr_avx2 (core/distance/euclidean.rs:66-93)
__device__ __forceinline__ float d_sqeuclidean(const float* __restrict__ a,
                                               const float* __restrict__ b, int d) {
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const int chunks = d >> 3;
#pragma unroll 4
    for (int c = 0; c < chunks; ++c) {
        const float4 a0 = *reinterpret_cast<const float4*>(a + c * 8);
        const float4 a1 = *reinterpret_cast<const float4*>(a + c * 8 + 4);
        const float4 b0 = *reinterpret_cast<const float4*>(b + c * 8);
        const float4 b1 = *reinterpret_cast<const float4*>(b + c * 8 + 4);
        float t;
        t = a0.x - b0.x; acc[0] = __builtin_fmaf(t, t, acc[0]);
        t = a0.y - b0.y; acc[1] = __builtin_fmaf(t, t, acc[1]);
        t = a0.z - b0.z; acc[2] = __builtin_fmaf(t, t, acc[2]);
        t = a0.w - b0.w; acc[3] = __builtin_fmaf(t, t, acc[3]);
        t = a1.x - b1.x; acc[4] = __builtin_fmaf(t, t, acc[4]);
        t = a1.y - b1.y; acc[5] = __builtin_fmaf(t, t, acc[5]);
        t = a1.z - b1.z; acc[6] = __builtin_fmaf(t, t, acc[6]);
        t = a1.w - b1.w; acc[7] = __builtin_fmaf(t, t, acc[7]);
    }
    float r = hsum8(acc);
    for (int i = chunks * 8; i < d; ++i) {
        const float t = a[i] - b[i];
        r += t * t;
    }
    return r;
}

// dot_avx2 (core/distance/inner_product.rs:64-83) == dot_product_avx2 (core/tree/builder.rs:29-60)
__device__ __forceinline__ float d_dot8(const float* __restrict__ a, const float* __restrict__ b,
                                        int d) {
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const int chunks = d >> 3;
#pragma unroll 4
    for (int c = 0; c < chunks; ++c) {
        const float4 a0 = *reinterpret_cast<const float4*>(a + c * 8);
        const float4 a1 = *reinterpret_cast<const float4*>(a + c * 8 + 4);
        const float4 b0 = *reinterpret_cast<const float4*>(b + c * 8);
        const float4 b1 = *reinterpret_cast<const float4*>(b + c * 8 + 4);
        acc[0] = __builtin_fmaf(a0.x, b0.x, acc[0]);
        acc[1] = __builtin_fmaf(a0.y, b0.y, acc[1]);
        acc[2] = __builtin_fmaf(a0.z, b0.z, acc[2]);
        acc[3] = __builtin_fmaf(a0.w, b0.w, acc[3]);
        acc[4] = __builtin_fmaf(a1.x, b1.x, acc[4]);
        acc[5] = __builtin_fmaf(a1.y, b1.y, acc[5]);
        acc[6] = __builtin_fmaf(a1.z, b1.z, acc[6]);
        acc[7] = __builtin_fmaf(a1.w, b1.w, acc[7]);
    }
    float r = hsum8(acc);
    for (int i = chunks * 8; i < d; ++i) r += a[i] * b[i];
    return r;
}

// cosine_avx2 (core/distance/cosine.rs:44-82)
__device__ __forceinline__ float d_cosine(const float* __restrict__ a, const float* __restrict__ b,
                                          int d) {
    float vd[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float va[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float vb[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const int chunks = d >> 3;
#pragma unroll 2
    for (int c = 0; c < chunks; ++c) {
#pragma unroll
        for (int l = 0; l < 8; ++l) {
            const float x = a[c * 8 + l], y = b[c * 8 + l];
            vd[l] = __builtin_fmaf(x, y, vd[l]);
            va[l] = __builtin_fmaf(x, x, va[l]);
            vb[l] = __builtin_fmaf(y, y, vb[l]);
        }
    }
    float dot = hsum8(vd), na = hsum8(va), nb = hsum8(vb);
    for (int i = chunks * 8; i < d; ++i) {
        dot += a[i] * b[i];
        na += a[i] * a[i];
        nb += b[i] * b[i];
    }
    const float denom = sqrtf(na * nb);
    if (denom < 1e-12f) return 1.0f;
    float s = dot / denom;
    s = s < -1.0f ? -1.0f : (s > 1.0f ? 1.0f : s);
    return 1.0f - s;
}

// normalized_dot_avx2 (core/distance/alternatives.rs:55-89): four 8-lane accumulators over 32-chunks
__device__ __forceinline__ float d_normdot(const float* __restrict__ a, const float* __restrict__ b,
                                           int d) {
    float s[4][8];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int l = 0; l < 8; ++l) s[q][l] = 0.f;
    const int chunks = d / 32;
#pragma unroll 2
    for (int c = 0; c < chunks; ++c) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int l = 0; l < 8; ++l)
                s[q][l] = __builtin_fmaf(a[c * 32 + q * 8 + l], b[c * 32 + q * 8 + l], s[q][l]);
    }
    float v[8];
#pragma unroll
    for (int l = 0; l < 8; ++l) v[l] = (s[0][l] + s[1][l]) + (s[2][l] + s[3][l]);
    float dot = hsum8_alt(v);
    for (int i = chunks * 32; i < d; ++i) dot += a[i] * b[i];
    return dot;
}

__device__ __forceinline__ float distance_exact(int metric, const float* __restrict__ a,
                                                const float* __restrict__ b, int d) {
    const float FMAX = 3.40282347e+38f;
    switch (metric) {
        case RINND_DIST_SQEUCLIDEAN: return d_sqeuclidean(a, b, d);
        case RINND_DIST_COSINE: return d_cosine(a, b, d);
        case RINND_DIST_INNER_PRODUCT: return -d_dot8(a, b, d);  // inner_product.rs:98-102
        case RINND_DIST_ALT_DOT: {                                // alternatives.rs:39-51
            const float dot = d_normdot(a, b, d);
            return dot <= 0.0f ? FMAX : -log2f(dot);
        }
        case RINND_DIST_DIRECT_COSINE: {  // alternatives.rs:133-149
            float s = d_normdot(a, b, d);
            s = s < -1.0f ? -1.0f : (s > 1.0f ? 1.0f : s);
            return 1.0f - s;
        }
        case RINND_DIST_ALT_COSINE: {  // alternatives.rs:202-226 (scalar, sequential)
            float dot = 0.f, na = 0.f, nb = 0.f;
            for (int i = 0; i < d; ++i) {
                dot += a[i] * b[i];
                na += a[i] * a[i];
                nb += b[i] * b[i];
            }
            if (na == 0.0f && nb == 0.0f) return 0.0f;
            if (na == 0.0f || nb == 0.0f) return FMAX;
            if (dot <= 0.0f) return FMAX;
            return log2f(sqrtf(na * nb) / dot);
        }
        case RINND_DIST_ALT_INNER_PRODUCT: {  // alternatives.rs:162-188 (scalar, 4-chunk)
            float dot = 0.f;
            const int chunks = d / 4;
            for (int c = 0; c < chunks; ++c) {
                const int i = c * 4;
                dot += a[i] * b[i] + a[i + 1] * b[i + 1] + a[i + 2] * b[i + 2] + a[i + 3] * b[i + 3];
            }
            for (int i = chunks * 4; i < d; ++i) dot += a[i] * b[i];
            return dot <= 0.0f ? FMAX : 1.0f / dot;
        }
        default: return __builtin_nanf("");
    }
}

// ---- FastRng (core/rng.rs:15-80) on host and device ---------------------------------------------
struct Xoshiro {
    uint64_t s[4];
    __host__ __device__ static inline uint64_t rotl(uint64_t x, int k) {
        return (x << k) | (x >> (64 - k));
    }
    __host__ __device__ inline uint64_t next_u64() {
        const uint64_t r = rotl(s[0] + s[3], 23) + s[0];
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    __host__ __device__ inline uint32_t next_u32() { return (uint32_t)(next_u64() >> 32); }
    __host__ __device__ inline float next_float() {
        return (float)(next_u32() >> 8) * (1.0f / 16777216.0f);
    }
    __host__ __device__ inline bool next_bool() { return (int32_t)next_u32() < 0; }
    // rand 0.8.5 gen_range(0..n) for usize
    __host__ __device__ inline uint64_t next_index(uint64_t n) {
        const uint64_t range = n;
        if (range == 0) return next_u64();
        const uint64_t zone = (range << __builtin_clzll(range)) - 1;
        for (;;) {
            const uint64_t v = next_u64();
#if defined(__HIP_DEVICE_COMPILE__)
            const uint64_t lo = v * range;
            const uint64_t hi = __umul64hi(v, range);
#else
            const unsigned __int128 m = (unsigned __int128)v * range;
            const uint64_t lo = (uint64_t)m;
            const uint64_t hi = (uint64_t)(m >> 64);
#endif
            if (lo <= zone) return hi;
        }
    }
};

// Xoshiro256PlusPlus::seed_from_u64 (SplitMix64 expansion).
__host__ __device__ inline Xoshiro xoshiro_seed(uint64_t seed) {
    Xoshiro r;
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) {
        x += 0x9e3779b97f4a7c15ULL;
        uint64_t z = x;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        r.s[i] = z ^ (z >> 31);
    }
    return r;
}

// Counter-based 64-bit mixer for RINND_RNG_COUNTER (SplitMix64 finaliser).
__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// Jump-ahead: the xoshiro256 state transition is linear over GF(2)^256.  jump_tables holds
// M^(2^j) for j = 0..63 as 256 rows x 4 words; row r of the product gives output bit r.
constexpr int kJumpLevels = 64;
__host__ __device__ inline void gf2_apply(const uint64_t* __restrict__ mat, uint64_t s[4]) {
    uint64_t out[4] = {0, 0, 0, 0};
    for (int r = 0; r < 256; ++r) {
        const uint64_t* row = mat + r * 4;
        const uint64_t v = (row[0] & s[0]) ^ (row[1] & s[1]) ^ (row[2] & s[2]) ^ (row[3] & s[3]);
        const uint64_t bit = (uint64_t)(__builtin_popcountll(v) & 1);
        out[r >> 6] |= bit << (r & 63);
    }
    s[0] = out[0];
    s[1] = out[1];
    s[2] = out[2];
    s[3] = out[3];
}
__host__ __device__ inline void xoshiro_jump(const uint64_t* __restrict__ tables, uint64_t s[4],
                                             uint64_t steps) {
    for (int j = 0; j < kJumpLevels && steps; ++j, steps >>= 1)
        if (steps & 1) gf2_apply(tables + (size_t)j * 1024, s);
}

// Host-side table (built once per process).
const uint64_t* host_jump_tables();

}  // namespace rinnd
