// leaf_init.hip - initialize_from_leaves (core/nndescent/mod.rs:203-300), one pass per tree.
//
// Within one tree every point lies in exactly one leaf.  The reference generates a leaf's pairs
// (i < j, row-major over leaf positions) and applies each to both endpoints in that order, leaves in
// forest order, so the pushes member a of a leaf receives from its tree are: every other member o in
// increasing leaf position, with d = distance(leaf[min(a,o)], leaf[max(a,o)]).  Its threshold filter
// (d < max of both heap roots taken at block start) only drops pushes that the live heap rejects
// anyway (roots only shrink), so one thread per member applies its own pushes directly to its heap:
// no update list, no grouping, the same heaps bit for bit.
//
// One workgroup per leaf.  The leaf's rows are staged in LDS (padded stride, conflict-free float4
// reads of each thread's own row; the other row is a broadcast) when the leaf fits; larger leaves
// use a second launch that reads rows from global memory.  Heaps sit in LDS column-major
// (slot j of thread c at j * B + c).
#include "internal.h"

namespace rinnd {

constexpr int kInitLdsBudget = 96 * 1024;

// checked_flagged_push (neighbor_heap.rs:149-177) on a column of an LDS heap block.
__device__ __forceinline__ void heap_push_col(float* hd, int32_t* hi, uint8_t* hf, int B, int c, int k, int32_t x,
                                              float d) {
    if (d >= hd[c]) return;
    bool dup = false;  // no early exit: the slot loads issue together
#pragma unroll 8
    for (int j = 0; j < k; ++j) dup |= hi[j * B + c] == x;
    if (dup) return;
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
}

// A tree offers each partner id to a row at most once (leaf members are distinct), so an offered id can
// already be in the row's heap only if it was there when the tree's pass began and has not been evicted
// since.  REG (k <= 16): those ids are kept in registers (an evicted id is cleared), which replaces the
// heap's duplicate scan by register compares.
template <bool REG>
__device__ __forceinline__ void leaf_offer(float* hd, int32_t* hi, uint8_t* hf, int S, int c, int k, int32_t x, float d,
                                           float& root, int32_t (&sid)[16]) {
    if (d >= root) return;
    if (REG) {
        bool dup = false;
#pragma unroll
        for (int j = 0; j < 16; ++j) dup |= sid[j] == x;
        if (dup) return;
        int32_t ev;
        root = heap_push_col_new(hd, hi, hf, S, c, k, x, d, ev);
#pragma unroll
        for (int j = 0; j < 16; ++j) sid[j] = sid[j] == ev ? -1 : sid[j];
    } else {
        heap_push_col(hd, hi, hf, S, c, k, x, d);
        root = hd[c];
    }
}

template <bool STAGED>
__global__ __launch_bounds__(256) void k_leaf_init(int metric, const float* __restrict__ X, int64_t ld, int dim,
                                                   const int32_t* __restrict__ tidx, const int64_t* __restrict__ leaves,
                                                   const int64_t* __restrict__ n_leaves_dev, int stage_cap,
                                                   int64_t ldl, int32_t* idx, float* dist, uint8_t* flags, int k) {
    extern __shared__ __align__(16) float lsm[];
    const int B = blockDim.x, tid = threadIdx.x;
    float* hd = lsm;
    int32_t* hi = (int32_t*)(hd + B * k);
    uint8_t* hf = (uint8_t*)(hi + B * k);
    float* rows = lsm + round_up((int64_t)B * k * 2 + (B * k + 3) / 4, 4);
    const int64_t n_leaves = *n_leaves_dev;
    for (int64_t leaf = blockIdx.x; leaf < n_leaves; leaf += gridDim.x) {
        const int64_t s = leaves[2 * leaf];
        const int m = (int)(leaves[2 * leaf + 1] - s);
        if (STAGED != (m <= stage_cap)) continue;  // uniform across the block
        if (STAGED) {
            const int vec = (int)(ld >> 2);
            for (int t = tid; t < m * vec; t += B) {
                const int c = t / vec, e4 = t - c * vec;
                *reinterpret_cast<float4*>(rows + c * ldl + e4 * 4) =
                    *reinterpret_cast<const float4*>(X + (int64_t)tidx[s + c] * ld + e4 * 4);
            }
            __syncthreads();
        }
        for (int a0 = 0; a0 < m; a0 += B) {
            const int a = a0 + tid;
            const bool act = a < m;
            const int32_t p = act ? tidx[s + a] : 0;
            if (act)
                for (int j = 0; j < k; ++j) {
                    hd[j * B + tid] = dist[(int64_t)p * k + j];
                    hi[j * B + tid] = idx[(int64_t)p * k + j];
                    hf[j * B + tid] = flags[(int64_t)p * k + j];
                }
            const float* ra = STAGED ? rows + (int64_t)a * ldl : X + (int64_t)p * ld;
            for (int o = 0; o < m; ++o) {
                const int32_t q = tidx[s + o];
                if (!act || o == a) continue;
                const float* ro = STAGED ? rows + (int64_t)o * ldl : X + (int64_t)q * ld;
                // every metric here is bitwise symmetric (products commute, the lane sums see the
                // same terms), so distance(leaf[min], leaf[max]) == distance(ra, ro)
                const float d = distance_exact(metric, ra, ro, dim);
                heap_push_col(hd, hi, hf, B, tid, k, q, d);
            }
            if (act)
                for (int j = 0; j < k; ++j) {
                    dist[(int64_t)p * k + j] = hd[j * B + tid];
                    idx[(int64_t)p * k + j] = hi[j * B + tid];
                    flags[(int64_t)p * k + j] = hf[j * B + tid];
                }
        }
        __syncthreads();  // rows are restaged by the next leaf
    }
}

// Two-pass form for leaves whose rows fit in LDS.  k_leaf_dist evaluates every pair of a leaf once
// (all threads busy) into D, an n x Dl4 matrix whose row p holds the distances from tree position p
// to the members of its leaf; k_leaf_push gives each position one lane that applies its pushes in
// leaf order with its heap in an LDS column and the root in a register, so rejected pushes cost
// one compare.
// Stages rows tidx[s..s+m) of X (stride ld) into LDS rows of stride ldl, rows m..mp-1 zero.  The member ids
// go to rid first, so the row loads (eight float4 in flight per thread) do not each wait on an id load.
// Ends with a barrier.
__device__ __forceinline__ void stage_leaf_rows(const float* __restrict__ X, int64_t ld, const int32_t* __restrict__ tidx,
                                                int64_t s, int m, int mp, float* rsm, int64_t ldl, int32_t* rid) {
    const int B = blockDim.x, t = threadIdx.x;
    for (int i = t; i < mp; i += B) rid[i] = i < m ? tidx[s + i] : -1;
    __syncthreads();
    const int vec = (int)(ld >> 2);
    const int total = mp * vec;
    for (int i0 = 0; i0 < total; i0 += 8 * B) {
        float4 v[8];
        bool pad[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = min(i0 + u * B + t, total - 1);
            const int c = i / vec, e4 = i - c * vec;
            const int32_t id = rid[c];
            pad[u] = id < 0;
            v[u] = *reinterpret_cast<const float4*>(X + (int64_t)(id < 0 ? 0 : id) * ld + e4 * 4);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = i0 + u * B + t;
            if (i < total) {
                const int c = i / vec, e4 = i - c * vec;
                *reinterpret_cast<float4*>(rsm + c * ldl + e4 * 4) = pad[u] ? make_float4(0.f, 0.f, 0.f, 0.f) : v[u];
            }
        }
    }
    __syncthreads();
}

__device__ __forceinline__ void tri_pair(int q, int m, int& a, int& o) {
    // q-th pair (a < o) in row-major order: counted back from the last row, r = k(k+1)/2 + rest
    const int r = m * (m - 1) / 2 - 1 - q;
    int k = (int)((sqrtf(8.0f * (float)r + 1.0f) - 1.0f) * 0.5f);
    while ((k + 1) * (k + 2) / 2 <= r) ++k;
    while (k * (k + 1) / 2 > r) --k;
    a = m - 2 - k;
    o = m - 1 - (r - k * (k + 1) / 2);
}

__global__ __launch_bounds__(256) void k_leaf_dist(int metric, const float* __restrict__ X, int64_t ld, int dim,
                                                   const int32_t* __restrict__ tidx,
                                                   const int64_t* __restrict__ leaves,
                                                   const int64_t* __restrict__ n_leaves_dev, int mcap, int64_t ldl,
                                                   float* __restrict__ D, int64_t Dl4, unsigned long long* pairs) {
    extern __shared__ __align__(16) float rsm[];
    const int B = blockDim.x, t = threadIdx.x;
    const int64_t n_leaves = *n_leaves_dev;
    unsigned long long np = 0;
    for (int64_t leaf = blockIdx.x; leaf < n_leaves; leaf += gridDim.x) {
        const int64_t s = leaves[2 * leaf];
        const int m = (int)(leaves[2 * leaf + 1] - s);
        if (m > mcap || m < 2) continue;  // uniform; larger leaves go to the untiled kernel
        np += (unsigned long long)m * (m - 1);
        stage_leaf_rows(X, ld, tidx, s, m, m, rsm, ldl, (int32_t*)(rsm + (int64_t)round_up((int64_t)mcap, (int64_t)4) * ldl));
        // every metric here is bitwise symmetric, so one evaluation serves both members
        const int T = m * (m - 1) / 2;
        for (int q = t; q < T; q += B) {
            int a, o;
            tri_pair(q, m, a, o);
            const float d = distance_exact(metric, rsm + a * ldl, rsm + o * ldl, dim);
            D[(s + a) * Dl4 + o] = d;
            D[(s + o) * Dl4 + a] = d;
        }
        __syncthreads();
    }
    if (t == 0 && np) atomicAdd(pairs, np);
}

// Squared Euclidean with 4 x 4 register tiles: a thread evaluates rows 4A..4A+3 against rows
// 4O..4O+3 (A <= O), loading each 8-float chunk of the 8 rows once for 16 pairs.  Each pair keeps the
// reference's arithmetic: 8 lane accumulators (fma of the difference with itself) over the chunks in
// order, the hsum of l2_sqr_avx2, then the scalar tail without fma.
__global__ __launch_bounds__(256) void k_leaf_dist_sq(const float* __restrict__ X, int64_t ld, int dim,
                                                      const int32_t* __restrict__ tidx,
                                                      const int64_t* __restrict__ leaves,
                                                      const int64_t* __restrict__ n_leaves_dev, int mcap, int64_t ldl,
                                                      float* __restrict__ D, int64_t Dl4, unsigned long long* pairs) {
    extern __shared__ __align__(16) float rsm[];
    const int B = blockDim.x, t = threadIdx.x;
    const int64_t n_leaves = *n_leaves_dev;
    const int chunks = dim >> 3;
    unsigned long long np = 0;
    for (int64_t leaf = blockIdx.x; leaf < n_leaves; leaf += gridDim.x) {
        const int64_t s = leaves[2 * leaf];
        const int m = (int)(leaves[2 * leaf + 1] - s);
        if (m > mcap || m < 2) continue;  // uniform; larger leaves go to the untiled kernel
        np += (unsigned long long)m * (m - 1);
        const int mp = round_up(m, 4);  // rows m..mp-1 are zero padding for the edge tiles
        stage_leaf_rows(X, ld, tidx, s, m, mp, rsm, ldl, (int32_t*)(rsm + (int64_t)round_up((int64_t)mcap, (int64_t)4) * ldl));
        const int nb = mp >> 2;
        const int T = nb * (nb + 1) / 2;
        for (int q = t; q < T; q += B) {
            int A, O;
            tri_pair(q, nb + 1, A, O);  // pairs A < O' over nb + 1 blocks -> A <= O = O' - 1
            O -= 1;
            const float* ra = rsm + (4 * A) * ldl;
            const float* ro = rsm + (4 * O) * ldl;
            float acc[4][4][8];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int l = 0; l < 8; ++l) acc[i][j][l] = 0.f;
            for (int c = 0; c < chunks; ++c) {
                float4 xa[4][2], xo[4][2];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    xa[i][0] = *reinterpret_cast<const float4*>(ra + i * ldl + c * 8);
                    xa[i][1] = *reinterpret_cast<const float4*>(ra + i * ldl + c * 8 + 4);
                    xo[i][0] = *reinterpret_cast<const float4*>(ro + i * ldl + c * 8);
                    xo[i][1] = *reinterpret_cast<const float4*>(ro + i * ldl + c * 8 + 4);
                }
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        float d;
                        d = xa[i][0].x - xo[j][0].x; acc[i][j][0] = __builtin_fmaf(d, d, acc[i][j][0]);
                        d = xa[i][0].y - xo[j][0].y; acc[i][j][1] = __builtin_fmaf(d, d, acc[i][j][1]);
                        d = xa[i][0].z - xo[j][0].z; acc[i][j][2] = __builtin_fmaf(d, d, acc[i][j][2]);
                        d = xa[i][0].w - xo[j][0].w; acc[i][j][3] = __builtin_fmaf(d, d, acc[i][j][3]);
                        d = xa[i][1].x - xo[j][1].x; acc[i][j][4] = __builtin_fmaf(d, d, acc[i][j][4]);
                        d = xa[i][1].y - xo[j][1].y; acc[i][j][5] = __builtin_fmaf(d, d, acc[i][j][5]);
                        d = xa[i][1].z - xo[j][1].z; acc[i][j][6] = __builtin_fmaf(d, d, acc[i][j][6]);
                        d = xa[i][1].w - xo[j][1].w; acc[i][j][7] = __builtin_fmaf(d, d, acc[i][j][7]);
                    }
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int a = 4 * A + i, o = 4 * O + j;
                    if (a < o && o < m) {
                        float r = hsum8(acc[i][j]);
                        for (int e = chunks * 8; e < dim; ++e) {
                            const float df = ra[i * ldl + e] - ro[j * ldl + e];
                            r += df * df;
                        }
                        D[(s + a) * Dl4 + o] = r;
                        D[(s + o) * Dl4 + a] = r;
                    }
                }
        }
        __syncthreads();
    }
    if (t == 0 && np) atomicAdd(pairs, np);
}

// One lane per tree position: its heap in an LDS column, its root in a register (a rejected push costs
// one compare).  A wave's 64 positions are consecutive, so their rows of D form one contiguous block:
// the wave streams it through a private LDS tile in COLS-column chunks, COLS/4 lanes per row segment
// (coalesced, each byte read once), with the next chunk loaded into registers while the current one is
// replayed.  Waves never wait for each other.  16-column tiles leave room for a third workgroup per CU,
// which hides the sift latency of the many accepted pushes of the first trees; 32-column tiles halve the
// chunk overhead once most pushes are rejected.
template <int COLS, bool REG>
__global__ __launch_bounds__(256) void k_leaf_push(const int32_t* __restrict__ tidx, const int64_t* __restrict__ leaves,
                                                   const int32_t* __restrict__ mark, const int64_t* __restrict__ rank,
                                                   int64_t n, int mcap, const float* __restrict__ D, int64_t Dl4,
                                                   int32_t* idx, float* dist, uint8_t* flags, int k) {
    extern __shared__ __align__(16) float psm[];
    const int B = blockDim.x, c = threadIdx.x;
    const int S = B + 1;  // heap column stride (odd: the row staging is conflict-free)
    const int lane = c & 63, w = c >> 6;
    float* hd = psm;
    int32_t* hi = (int32_t*)(hd + S * k);
    uint8_t* hf = (uint8_t*)(hi + S * k);
    constexpr int kPushCols = COLS, kPushTL = COLS + 1;  // padded tile row: lane-strided reads conflict-free
    float* tile = psm + round_up((int64_t)S * k * 2 + (S * k + 3) / 4, 4) + w * 64 * kPushTL;
    constexpr int LPR = kPushCols / 4;  // loader lanes per row (one float4 each)
    constexpr int RPI = 64 / LPR;       // rows per load instruction
    const int rl = lane / LPR, q4 = (lane % LPR) * 4;  // loader role: rows rl + RPI i, columns q4..q4+3
    for (int64_t base = ((int64_t)blockIdx.x * B) + w * 64; base < n; base += (int64_t)gridDim.x * B) {
        const int64_t pos = base + lane;
        int m = 0, a = 0;
        int64_t s = 0;
        if (pos < n) {
            const int64_t l = rank[pos] + mark[pos] - 1;  // leaf holding this position
            s = leaves[2 * l];
            m = (int)(leaves[2 * l + 1] - s);
            if (m > mcap || m < 2) m = 0;  // not a two-pass leaf
            a = (int)(pos - s);
        }
        const int32_t hrow = m ? tidx[pos] : -1;
        wave_heap_rows<true>(hrow, k, dist, idx, flags, hd, hi, hf, S, w * 64);
        float root = m ? hd[c] : 0.f;
        int32_t sid[16];  // REG: the row's ids at the start of the pass that are still in the heap
#pragma unroll
        for (int j = 0; j < 16; ++j) sid[j] = REG && m && j < k ? hi[j * S + c] : -1;
        int mw = m;  // the wave's longest row
        for (int o = 32; o > 0; o >>= 1) mw = max(mw, __shfl_xor(mw, o));
        const int64_t rows = min((int64_t)64, n - base);
        float4 nx[LPR];
        auto load = [&](int c0) {
#pragma unroll
            for (int i = 0; i < LPR; ++i) {
                const int r = rl + RPI * i;
                nx[i] = (r < rows && c0 + q4 < Dl4)
                            ? *reinterpret_cast<const float4*>(D + (base + r) * Dl4 + c0 + q4)
                            : make_float4(0.f, 0.f, 0.f, 0.f);
            }
        };
        if (mw) load(0);
        for (int c0 = 0; c0 < mw; c0 += kPushCols) {
            wave_sync();  // every lane is done with the previous chunk
#pragma unroll
            for (int i = 0; i < LPR; ++i) {
                float* t = tile + (rl + RPI * i) * kPushTL + q4;
                t[0] = nx[i].x;
                t[1] = nx[i].y;
                t[2] = nx[i].z;
                t[3] = nx[i].w;
            }
            wave_sync();
            if (c0 + kPushCols < mw) load(c0 + kPushCols);  // in flight while this chunk is replayed
            if (m) {
                // survivors of this chunk against the current root (a superset of the accepted ones:
                // the root only falls), then pushed in leaf order
                uint32_t msk = 0;
#pragma unroll
                for (int j = 0; j < kPushCols; ++j) {
                    const int o = c0 + j;
                    if (o < m && o != a && tile[lane * kPushTL + j] < root) msk |= 1u << j;
                }
                // the next survivor's id is loaded while the current one is pushed
                int j = msk ? __ffs((int)msk) - 1 : -1;
                int32_t xj = j >= 0 ? tidx[s + c0 + j] : 0;
                while (j >= 0) {
                    msk &= msk - 1;
                    const int jn = msk ? __ffs((int)msk) - 1 : -1;
                    const int32_t xn = jn >= 0 ? tidx[s + c0 + jn] : 0;
                    const float d = tile[lane * kPushTL + j];
                    leaf_offer<REG>(hd, hi, hf, S, c, k, xj, d, root, sid);
                    j = jn;
                    xj = xn;
                }
            }
        }
        wave_heap_rows<false>(hrow, k, dist, idx, flags, hd, hi, hf, S, w * 64);
    }
}

// The pushes of a group of survivor trees (k_leaf_surv_mfma, point-major): one lane per point, its heap in an
// LDS column (64 consecutive rows staged per wave, coalesced), the root in a register.  For each tree of the
// group in order: the point's leaf (info); for a leaf that k_leaf_surv_mfma evaluated, its survivors in leaf
// order (mask bits, distance at the leaf column of its D row, partner id from the tree's indices), the next
// survivor's distance and partner id in flight while the current one is pushed; for a leaf above mcap, every
// other member in leaf order with the distance in the reference arithmetic (all metrics here are bitwise
// symmetric).  A point's pushes come only from its own leaves, so replaying its trees one after another is the
// reference's order for its heap.
template <bool REG>
__global__ __launch_bounds__(256) void k_leaf_push_group(SurvGroup grp, const float* __restrict__ X, int64_t ld,
                                                         int dim, int64_t n, int mcap, int64_t Dl4, int32_t* idx,
                                                         float* dist, uint8_t* flags, int k) {
    extern __shared__ __align__(16) float qsm[];
    const int B = blockDim.x, c = threadIdx.x;
    const int S = B + 1;  // heap column stride (odd: the row staging is conflict-free)
    const int lane = c & 63, w = c >> 6;
    float* hd = qsm;
    int32_t* hi = (int32_t*)(hd + S * k);
    uint8_t* hf = (uint8_t*)(hi + S * k);
    for (int64_t base = ((int64_t)blockIdx.x * B) + w * 64; base < n; base += (int64_t)gridDim.x * B) {
        const int64_t p = base + lane;
        const int32_t hrow = p < n ? (int32_t)p : -1;
        wave_heap_rows<true>(hrow, k, dist, idx, flags, hd, hi, hf, S, w * 64);
        if (p < n) {
            float root = hd[c];
            int32_t sid[16];  // REG: the row's ids at the start of the tree's pass that are still in the heap
            for (int g = 0; g < grp.G; ++g) {
                const int2 inf = grp.info[(int64_t)g * n + p];
                const int64_t s = inf.x;
                const int m = inf.y;
                if (m < 2) continue;
                const int32_t* tidx = grp.tidx[g];
#pragma unroll
                for (int j = 0; j < 16; ++j) sid[j] = REG && j < k ? hi[j * S + c] : -1;
                if (m <= mcap) {
                    const uint32_t* mw = grp.Msk + ((int64_t)g * n + p) * 8;
                    const uint4 m0 = *reinterpret_cast<const uint4*>(mw);
                    const uint4 m1 = *reinterpret_cast<const uint4*>(mw + 4);
                    const uint32_t words[8] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w};
                    const float* Dr = grp.D + ((int64_t)g * n + p) * Dl4;  // survivor o's distance at column o
#pragma unroll
                    for (int wi = 0; wi < 8; ++wi) {
                        uint32_t bits = words[wi];
                        int j = bits ? __ffs((int)bits) - 1 : -1;
                        int32_t xj = j >= 0 ? tidx[s + 32 * wi + j] : 0;
                        float dj = j >= 0 ? Dr[32 * wi + j] : 0.f;
                        while (j >= 0) {  // the next survivor's id and distance in flight
                            bits &= bits - 1;
                            const int jn = bits ? __ffs((int)bits) - 1 : -1;
                            const int32_t xn = jn >= 0 ? tidx[s + 32 * wi + jn] : 0;
                            const float dnx = jn >= 0 ? Dr[32 * wi + jn] : 0.f;
                            leaf_offer<REG>(hd, hi, hf, S, c, k, xj, dj, root, sid);
                            j = jn;
                            xj = xn;
                            dj = dnx;
                        }
                    }
                } else {
                    const float* xp = X + p * ld;
                    for (int o = 0; o < m; ++o) {
                        const int32_t q = tidx[s + o];
                        if (q == (int32_t)p) continue;
                        const float d = distance_exact(RINND_DIST_SQEUCLIDEAN, xp, X + (int64_t)q * ld, dim);
                        leaf_offer<REG>(hd, hi, hf, S, c, k, q, d, root, sid);
                    }
                }
            }
        }
        wave_heap_rows<false>(hrow, k, dist, idx, flags, hd, hi, hf, S, w * 64);
    }
}

// The same 4 x 4 register tiles for leaves whose rows do not fit in LDS together (wide rows): the leaf's rows
// are staged KC dimensions at a time for all members, and every tile's 8 lane accumulators stay in registers
// across the chunks, so each pair still sees the reference's arithmetic (chunks in order, then the hsum and the
// scalar tail).  512 threads; a leaf with more tiles than threads takes several passes over its rows.
template <int KC>
__global__ __launch_bounds__(512) void k_leaf_dist_sq_kc(const float* __restrict__ X, int64_t ld, int dim,
                                                         const int32_t* __restrict__ tidx,
                                                         const int64_t* __restrict__ leaves,
                                                         const int64_t* __restrict__ n_leaves_dev, int mcap,
                                                         float* __restrict__ D, int64_t Dl4,
                                                         unsigned long long* pairs) {
    constexpr int LDK = KC + 4;  // staged row stride (floats): 16-byte aligned, rows start on different banks
    extern __shared__ __align__(16) float csm[];
    const int B = blockDim.x, t = threadIdx.x;
    int32_t* rid = (int32_t*)(csm + (int64_t)round_up((int64_t)mcap, (int64_t)4) * LDK);
    const int64_t n_leaves = *n_leaves_dev;
    const int chunks = dim >> 3;
    const int dim8 = chunks * 8;
    unsigned long long np = 0;
    for (int64_t leaf = blockIdx.x; leaf < n_leaves; leaf += gridDim.x) {
        const int64_t s = leaves[2 * leaf];
        const int m = (int)(leaves[2 * leaf + 1] - s);
        if (m > mcap || m < 2) continue;  // uniform; larger leaves go to the untiled kernel
        np += (unsigned long long)m * (m - 1);
        const int mp = round_up(m, 4);  // rows m..mp-1 are zero padding for the edge tiles
        __syncthreads();                // the previous leaf is done with rid
        for (int i = t; i < mp; i += B) rid[i] = i < m ? tidx[s + i] : -1;
        const int nb = mp >> 2;
        const int T = nb * (nb + 1) / 2;
        for (int q0 = 0; q0 < T; q0 += B) {  // uniform
            const int q = q0 + t;
            const bool act = q < T;
            int A = 0, O = 0;
            if (act) {
                tri_pair(q, nb + 1, A, O);  // pairs A < O' over nb + 1 blocks -> A <= O = O' - 1
                O -= 1;
            }
            float acc[4][4][8];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int l = 0; l < 8; ++l) acc[i][j][l] = 0.f;
            for (int c0 = 0; c0 < dim8; c0 += KC) {
                const int kc = min(KC, dim8 - c0);  // a multiple of 8
                const int v4 = kc >> 2;
                __syncthreads();  // the last chunk is consumed (and rid is written)
                for (int i = t; i < mp * v4; i += B) {
                    const int r = i / v4, e = i - r * v4;
                    const int32_t id = rid[r];
                    *reinterpret_cast<float4*>(csm + r * LDK + 4 * e) =
                        id >= 0 ? *reinterpret_cast<const float4*>(X + (int64_t)id * ld + c0 + 4 * e)
                                : make_float4(0.f, 0.f, 0.f, 0.f);
                }
                __syncthreads();
                if (act) {
                    const float* ra = csm + (4 * A) * LDK;
                    const float* ro = csm + (4 * O) * LDK;
                    for (int c = 0; c < kc; c += 8) {
                        float4 xa[4][2], xo[4][2];
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            xa[i][0] = *reinterpret_cast<const float4*>(ra + i * LDK + c);
                            xa[i][1] = *reinterpret_cast<const float4*>(ra + i * LDK + c + 4);
                            xo[i][0] = *reinterpret_cast<const float4*>(ro + i * LDK + c);
                            xo[i][1] = *reinterpret_cast<const float4*>(ro + i * LDK + c + 4);
                        }
#pragma unroll
                        for (int i = 0; i < 4; ++i)
#pragma unroll
                            for (int j = 0; j < 4; ++j) {
                                float d;
                                d = xa[i][0].x - xo[j][0].x; acc[i][j][0] = __builtin_fmaf(d, d, acc[i][j][0]);
                                d = xa[i][0].y - xo[j][0].y; acc[i][j][1] = __builtin_fmaf(d, d, acc[i][j][1]);
                                d = xa[i][0].z - xo[j][0].z; acc[i][j][2] = __builtin_fmaf(d, d, acc[i][j][2]);
                                d = xa[i][0].w - xo[j][0].w; acc[i][j][3] = __builtin_fmaf(d, d, acc[i][j][3]);
                                d = xa[i][1].x - xo[j][1].x; acc[i][j][4] = __builtin_fmaf(d, d, acc[i][j][4]);
                                d = xa[i][1].y - xo[j][1].y; acc[i][j][5] = __builtin_fmaf(d, d, acc[i][j][5]);
                                d = xa[i][1].z - xo[j][1].z; acc[i][j][6] = __builtin_fmaf(d, d, acc[i][j][6]);
                                d = xa[i][1].w - xo[j][1].w; acc[i][j][7] = __builtin_fmaf(d, d, acc[i][j][7]);
                            }
                    }
                }
            }
            if (act) {
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int a = 4 * A + i, o = 4 * O + j;
                        if (a < o && o < m) {
                            float r = hsum8(acc[i][j]);
                            const float* xa = X + (int64_t)rid[a] * ld;
                            const float* xo = X + (int64_t)rid[o] * ld;
                            for (int e = dim8; e < dim; ++e) {  // scalar tail, no fma
                                const float df = xa[e] - xo[e];
                                r += df * df;
                            }
                            D[(s + a) * Dl4 + o] = r;
                            D[(s + o) * Dl4 + a] = r;
                        }
                    }
            }
        }
    }
    if (t == 0 && np) atomicAdd(pairs, np);
}

struct TilePlan {
    int mcap = 0;
    int64_t Dl4 = 0;
    size_t rows_lds = 0;
    bool chunked = false;  // rows staged kLeafKC dimensions at a time (squared Euclidean only)
};
constexpr int kLeafKC = 64;
static TilePlan tile_plan(int64_t ld, int64_t leaf_size, int metric = RINND_DIST_SQEUCLIDEAN) {
    TilePlan P;
    const int mcap = (int)std::min<int64_t>(std::max<int64_t>(leaf_size, 2), 256);
    const size_t rows = (size_t)round_up((int64_t)mcap, (int64_t)4) * (ld + 4 + 1) * 4;  // rows + member ids
    if (rows > 150 * 1024) {
        if (metric != RINND_DIST_SQEUCLIDEAN) return P;
        P.chunked = true;
        P.rows_lds = (size_t)round_up((int64_t)mcap, (int64_t)4) * (kLeafKC + 4 + 1) * 4;
    } else {
        P.rows_lds = rows;
    }
    P.mcap = mcap;
    P.Dl4 = round_up((int64_t)mcap, (int64_t)4);
    return P;
}

static int init_block(int64_t leaf_size) {
    return (int)std::min<int64_t>(256, round_up(std::max<int64_t>(leaf_size, 1), (int64_t)64));
}
static size_t init_heap_bytes(int B, int k) {
    return (size_t)round_up((int64_t)B * k * 2 + (B * k + 3) / 4, 4) * 4;
}

// Lanes per workgroup of the push kernels: one LDS heap column per lane (plus one for the odd stride) in 64 KiB.
static int push_block(int k) {
    int PB = 256;
    while (PB > 64 && init_heap_bytes(PB + 1, k) > 64 * 1024) PB -= 64;
    return PB;
}

bool leaf_init_supported(int k, int64_t leaf_size) {
    return init_heap_bytes(init_block(leaf_size), k) <= (size_t)kInitLdsBudget;
}
int64_t leaf_init_scratch_floats(int64_t n, int64_t ld, int64_t leaf_size) {
    const TilePlan P = tile_plan(ld, leaf_size);  // the widest plan (squared Euclidean)
    return P.mcap ? std::max<int64_t>(n, 1) * P.Dl4 : 0;
}

template <int COLS>
static void launch_leaf_push(hipStream_t s, unsigned grid, int B, size_t lds, const int32_t* tree_indices,
                             const int64_t* leaves, const LeafInitWs& ws, int64_t n, const TilePlan& P, int32_t* idx,
                             float* dist, uint8_t* flags, int k) {
    if (k <= 16)
        hipLaunchKernelGGL((k_leaf_push<COLS, true>), dim3(grid), dim3(B), lds, s, tree_indices, leaves, ws.mark, ws.rank,
                           n, P.mcap, ws.D, P.Dl4, idx, dist, flags, k);
    else
        hipLaunchKernelGGL((k_leaf_push<COLS, false>), dim3(grid), dim3(B), lds, s, tree_indices, leaves, ws.mark,
                           ws.rank, n, P.mcap, ws.D, P.Dl4, idx, dist, flags, k);
}

void launch_leaf_init(hipStream_t s, int metric, const float* X, int64_t ld, int dim, const int32_t* tree_indices,
                      const int64_t* leaves, const int64_t* n_leaves_dev, int64_t max_leaves, int64_t leaf_size,
                      int32_t* idx, float* dist, uint8_t* flags, int k, const LeafInitWs& ws,
                      const KTimer& dist_timer, const KTimer& push_timer, bool early) {
    if (max_leaves <= 0) return;
    const int B = init_block(leaf_size);
    const size_t heap = init_heap_bytes(B, k);
    const int64_t ldl = ld + 4;
    const int64_t n = max_leaves;  // positions in the tree == points
    const unsigned grid = (unsigned)std::min<int64_t>(max_leaves, 8192);  // persistent over the leaves
    const TilePlan P = tile_plan(ld, leaf_size, metric);
    int untiled_cap = 0;  // leaves with m <= untiled_cap are done by the two-pass kernels
    const int PB = push_block(k);
    if (P.mcap > 0 && ws.D && init_heap_bytes(PB + 1, k) <= 64 * 1024) {
        const bool mfma = ws.exact_bf16 && metric == RINND_DIST_SQEUCLIDEAN &&
                          leaf_dist_mfma_lds(ws.exact_bf16->ldb, P.mcap) <= 150 * 1024;
        dist_timer.begin(s);
        if (mfma)
            launch_leaf_dist_mfma(s, *ws.exact_bf16, tree_indices, leaves, n_leaves_dev, max_leaves, P.mcap, ws.D,
                                  P.Dl4, ws.pairs);
        else if (P.chunked)
            hipLaunchKernelGGL(k_leaf_dist_sq_kc<kLeafKC>, dim3(std::min<unsigned>(grid, 4096)), dim3(512), P.rows_lds,
                               s, X, ld, dim, tree_indices, leaves, n_leaves_dev, P.mcap, ws.D, P.Dl4, ws.pairs);
        else if (metric == RINND_DIST_SQEUCLIDEAN)
            hipLaunchKernelGGL(k_leaf_dist_sq, dim3(std::min<unsigned>(grid, 4096)), dim3(256), P.rows_lds, s, X, ld,
                               dim, tree_indices, leaves, n_leaves_dev, P.mcap, ldl, ws.D, P.Dl4, ws.pairs);
        else
            hipLaunchKernelGGL(k_leaf_dist, dim3(std::min<unsigned>(grid, 4096)), dim3(256), P.rows_lds, s, metric, X,
                               ld, dim, tree_indices, leaves, n_leaves_dev, P.mcap, ldl, ws.D, P.Dl4, ws.pairs);
        dist_timer.stop(s);
        push_timer.begin(s);
        const unsigned pgrid = (unsigned)std::min<int64_t>(ceil_div(n, (int64_t)PB), 1 << 16);
        if (early)
            launch_leaf_push<16>(s, pgrid, PB, init_heap_bytes(PB + 1, k) + (size_t)PB * 17 * 4, tree_indices, leaves,
                                 ws, n, P, idx, dist, flags, k);
        else
            launch_leaf_push<32>(s, pgrid, PB, init_heap_bytes(PB + 1, k) + (size_t)PB * 33 * 4, tree_indices, leaves,
                                 ws, n, P, idx, dist, flags, k);
        push_timer.stop(s);
        // the parts of the algorithmic bytes that do not depend on the leaf sizes; the caller adds the
        // per-pair parts (4 B of D written per ordered pair; 4 B of D and 4 B of partner id read) once it
        // has read the pair counter
        // dist: each member's row (bf16 + norm on MFMA) and id once
        dist_timer.collect((double)n * ((mfma ? ws.exact_bf16->ldb * 2.0 + 4.0 : dim * 4.0) + 4.0));
        // push: every member's heap read and written, its leaf lookup
        push_timer.collect((double)n * k * 9.0 * 2.0 + n * 40.0);
        untiled_cap = P.mcap;
    } else {
        const int64_t cap_rows = ((int64_t)kInitLdsBudget - (int64_t)heap) / (ldl * 4);
        untiled_cap = (int)std::max<int64_t>(0, std::min<int64_t>(cap_rows, std::max<int64_t>(leaf_size, 1)));
        if (untiled_cap > 0)
            hipLaunchKernelGGL(k_leaf_init<true>, dim3(grid), dim3(B), heap + (size_t)untiled_cap * ldl * 4, s, metric,
                               X, ld, dim, tree_indices, leaves, n_leaves_dev, untiled_cap, ldl, idx, dist, flags, k);
    }
    hipLaunchKernelGGL(k_leaf_init<false>, dim3(std::min<unsigned>(grid, 4096)), dim3(B), heap, s, metric, X, ld,
                       dim, tree_indices, leaves, n_leaves_dev, untiled_cap, ldl, idx, dist, flags, k);
    HIP_TRY(hipGetLastError());
}

double surv_group_budget() {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        (void)hipGetLastError();
        total_b = (size_t)64 << 30;
    }
    return std::min((double)total_b / 8.0, 40.0 * (double)(1ull << 30));
}

int64_t leaf_surv_dl4(int metric, const Bf16Rows* R, int64_t ld, int64_t leaf_size, int k) {
    if (!R || !R->exact || metric != RINND_DIST_SQEUCLIDEAN) return 0;
    const TilePlan P = tile_plan(ld, leaf_size);
    if (P.mcap <= 0 || P.chunked || leaf_surv_mfma_lds(R->ldb, P.mcap) > 150 * 1024) return 0;
    if (init_heap_bytes(push_block(k) + 1, k) > 64 * 1024) return 0;
    return P.Dl4;
}

void launch_leaf_surv_tree(hipStream_t s, const Bf16Rows& R, const int32_t* tree_indices, const int64_t* leaves,
                           const int64_t* n_leaves_dev, int64_t n, int64_t leaf_size, const float* hdist, int k,
                           const SurvGroup& grp, int g, unsigned long long* surv, const KTimer& timer) {
    if (n <= 0) return;
    const int mcap = (int)std::min<int64_t>(std::max<int64_t>(leaf_size, 2), 256);
    const int64_t Dl4 = round_up((int64_t)mcap, (int64_t)4);
    timer.begin(s);
    launch_leaf_surv_mfma(s, R, tree_indices, leaves, n_leaves_dev, n, mcap, hdist, k, grp.D + (int64_t)g * n * Dl4,
                          Dl4, grp.Msk + (int64_t)g * n * 8, grp.info + (int64_t)g * n, surv);
    timer.stop(s);
    // per point: bf16 row, norm, root and id read; mask and leaf info written (survivors added by the caller)
    timer.collect((double)n * (R.ldb * 2.0 + 12.0 + 32.0 + 8.0));
}

void launch_leaf_push_group(hipStream_t s, const SurvGroup& grp, const float* X, int64_t ld, int dim, int64_t n,
                            int64_t leaf_size, int32_t* idx, float* dist, uint8_t* flags, int k, const KTimer& timer) {
    if (n <= 0 || grp.G <= 0) return;
    const int mcap = (int)std::min<int64_t>(std::max<int64_t>(leaf_size, 2), 256);
    const int64_t Dl4 = round_up((int64_t)mcap, (int64_t)4);
    const int PB = push_block(k);
    const unsigned pgrid = (unsigned)std::min<int64_t>(ceil_div(n, (int64_t)PB), 1 << 16);
    timer.begin(s);
    if (k <= 16)
        hipLaunchKernelGGL(k_leaf_push_group<true>, dim3(pgrid), dim3(PB), init_heap_bytes(PB + 1, k), s, grp, X, ld,
                           dim, n, mcap, Dl4, idx, dist, flags, k);
    else
        hipLaunchKernelGGL(k_leaf_push_group<false>, dim3(pgrid), dim3(PB), init_heap_bytes(PB + 1, k), s, grp, X, ld,
                           dim, n, mcap, Dl4, idx, dist, flags, k);
    HIP_TRY(hipGetLastError());
    timer.stop(s);
    // per point: heap read and written once; per tree its leaf info and mask (survivors added by the caller)
    timer.collect((double)n * k * 9.0 * 2.0 + (double)n * grp.G * (8.0 + 32.0));
}

}  // namespace rinnd
