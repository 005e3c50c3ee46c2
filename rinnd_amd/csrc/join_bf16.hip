// join_bf16.hip - local join with BF16 rows on MFMA and FP32 re-evaluation inside a proven margin.
//
// For row i the join needs d(p, q) for p in new[i] (C) and q in new[i] ++ old[i] (2C)
// (update_iteration, core/nndescent/mod.rs:383-447).  One wavefront per row forms the Gram matrix
// G = A B^T of the bf16-rounded candidate rows on v_mfma_f32_16x16x16bf16 (FP32 accumulation) and
// the screened distance  dt = |a|^2 + |b|^2 - 2 G,  where |a|^2 are FP32 norms of the rounded rows.
//
// Error bound.  Rows are centred by the column means mu first (distances are translation invariant; the
// norms shrink): c = fl(x - mu), x^ = bf16(c).  The preparation pass stores, per row, an upper bound e_x of
// |x^ - (x - mu)| (the rounding x^ - c is exact in FP32, its squared norm is summed with relative error g_d,
// and the centring adds at most 1.01 u24 |c|).  With a = x - y, a^ = x^ - y^ and e = e_x + e_y, the triangle
// inequality gives |a^ - a| <= e, so |a^|^2 <= (|a| + e)^2.  The screened value dt = (n_x + n_y) - 2 G, from the
// FP32 norms n of the rounded rows and the MFMA Gram G with FP32 accumulation, is within
// F = (2 g_d + 4 u24) (n_x + n_y) (1 + 2 g_d) of |a^|^2, and the reference's FP32 sum D_ref is within g_r |a|^2
// of |a|^2 (u24 = 2^-24, g_d = d u24 / (1 - d u24), g_r = 1.01 (d + 8) u24).  A pair can pass the reference
// filter D_ref < thr only if |a|^2 < thr / (1 - g_r) =: Db, and then
//   dt < (sqrt(Db) + e)^2 + F = Db + 2 e sqrt(Db) + e^2 + F.
// Every pair with dt <= Db + 1.05 (2 e sqrt(Db) + e^2 + F) is therefore re-evaluated with the reference-order
// FP32 distance and filtered exactly; no other pair can be kept, so the emitted updates - and the final heaps -
// equal the FP32 path's.  The pairs to re-evaluate are queued in LDS by the screen and evaluated afterwards
// with every lane busy.
//
// Exact mode.  If every coordinate is an integer with |x| <= 256 (exact in bf16) and either every
// |x|^2 <= 2^22, or no coordinate is negative and every |x|^2 <= 2^23, then all products, partial sums,
// norms, dt and every partial sum of the reference's (x - y)^2 accumulation are integers of at most
// 2^24: dt equals the true distance, which also equals the reference's FP32 result.  No
// re-evaluation is needed.
#include <cstring>

#include "internal.h"

namespace rinnd {

__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

// column sums (FP64) for centring
__global__ void k_col_sums(const float* __restrict__ X, int64_t ld, int64_t n, int dim, double* sums) {
    for (int c = threadIdx.x; c < dim; c += blockDim.x) {
        double s = 0.0;
        for (int64_t r = blockIdx.x; r < n; r += gridDim.x) s += X[r * ld + c];
        atomicAdd(&sums[c], s);
    }
}
__global__ void k_col_mean(const double* sums, int64_t n, int dim, float* mu) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < dim; c += gridDim.x * blockDim.x)
        mu[c] = (float)(sums[c] / (double)n);
}

// One pass over the rows for the BF16 join: the exactness test (flag bit 0: a value that is not an integer
// in [-256, 256]; bit 1: a negative value; the largest squared row norm in FP64) and the uncentred bf16 rows with the FP32 norms of the rounded rows, which are the rows exact mode uses
// (other data are converted again after centring).  16 lanes per row, 8 consecutive dims per lane.  Exact
// data have integer squared norms below 2^24, so the FP32 norms do not depend on the summation order.
template <bool VEC>
__global__ __launch_bounds__(256) void k_check_to_bf16(const float* __restrict__ X, int64_t ld, int64_t n, int dim,
                                                       uint16_t* __restrict__ Xb, int64_t ldb,
                                                       float* __restrict__ nrm, int* flag,
                                                       unsigned long long* maxnorm) {
    const int lane = threadIdx.x & 63, gl = lane & 15;
    const int64_t wid = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    double mx = 0.0;
    int bad = 0;
    for (int64_t r0 = wid * 4; r0 < n; r0 += nw * 4) {  // wave-uniform: four rows per wave
        const int64_t r = r0 + (lane >> 4);
        const bool valid = r < n;
        double s = 0.0;
        float sf = 0.f;
        for (int64_t c0 = 8 * gl; c0 < ldb; c0 += 128) {
            float v[8];
            if (VEC && valid && c0 + 8 <= dim) {
                const float4 p = *reinterpret_cast<const float4*>(X + r * ld + c0);
                const float4 q = *reinterpret_cast<const float4*>(X + r * ld + c0 + 4);
                v[0] = p.x, v[1] = p.y, v[2] = p.z, v[3] = p.w, v[4] = q.x, v[5] = q.y, v[6] = q.z, v[7] = q.w;
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) v[e] = valid && c0 + e < dim ? X[r * ld + c0 + e] : 0.f;  // pads are 0
            }
            uint32_t w[4];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                bad |= !(v[e] == rintf(v[e]) && fabsf(v[e]) <= 256.0f) ? 1 : 0;
                bad |= v[e] < 0.0f ? 2 : 0;
                s += (double)v[e] * (double)v[e];
                const uint16_t h = f2bf(v[e]);
                const float f = bf2f(h);
                sf = __builtin_fmaf(f, f, sf);
                if (e & 1) w[e >> 1] |= (uint32_t)h << 16;
                else w[e >> 1] = h;
            }
            if (valid) *reinterpret_cast<uint4*>(Xb + r * ldb + c0) = make_uint4(w[0], w[1], w[2], w[3]);
        }
        for (int o = 8; o > 0; o >>= 1) {
            s += __shfl_xor(s, o, 16);
            sf += __shfl_xor(sf, o, 16);
        }
        if (valid) {
            mx = fmax(mx, s);
            if (gl == 0) nrm[r] = sf;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        bad |= __shfl_xor(bad, o);
        mx = fmax(mx, __shfl_xor(mx, o));
    }
    // one pair of atomics per workgroup: atomics on one address serialise in the L2
    __shared__ double wmx[4];
    __shared__ int wbad[4];
    if (lane == 0) {
        wmx[threadIdx.x >> 6] = mx;
        wbad[threadIdx.x >> 6] = bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) {
            mx = fmax(mx, wmx[i]);
            bad |= wbad[i];
        }
        if (bad) atomicOr(flag, bad);
        atomicMax(maxnorm, (unsigned long long)__double_as_longlong(mx));
    }
}

// bf16 rows (optionally centred) and FP32 norms of the rounded rows; one wave per row.  err (if not null) gets an
// upper bound of |x^ - (x - mu)|: the rounding x^ - c of each centred coordinate c is exact in FP32, the sum of
// their squares carries a relative error of at most g_d, and the centring c = fl(x - mu) adds at most
// 1.01 u24 |c| <= 1.01 u24 (|x^| + |x^ - c|).
__global__ void k_to_bf16(const float* __restrict__ X, int64_t ld, int64_t n, int dim,
                          const float* __restrict__ mu, uint16_t* __restrict__ Xb, int64_t ldb,
                          float* __restrict__ nrm, float* __restrict__ err, float g_d) {
    const int lane = threadIdx.x & 63;
    const int64_t wid = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t r = wid; r < n; r += nw) {
        float s = 0.f, se = 0.f;
        for (int c = lane; c < ldb; c += 64) {
            uint16_t h = 0;
            if (c < dim) {
                const float v = mu ? X[r * ld + c] - mu[c] : X[r * ld + c];
                h = f2bf(v);
                const float f = bf2f(h);
                s = __builtin_fmaf(f, f, s);
                const float dv = f - v;  // exact
                se = __builtin_fmaf(dv, dv, se);
            }
            Xb[r * ldb + c] = h;
        }
        for (int o = 32; o > 0; o >>= 1) {
            s += __shfl_xor(s, o);
            se += __shfl_xor(se, o);
        }
        if (lane == 0) {
            nrm[r] = s;
            if (err) {
                const float eb = sqrtf(se * (1.0f + g_d)) * (1.0f + 2.4e-7f);
                const float en = sqrtf(s * (1.0f + g_d)) * (1.0f + 2.4e-7f);
                err[r] = (eb + 1.01f * 5.9604645e-8f * (en + eb)) * (1.0f + 4.8e-7f);
            }
        }
    }
}

struct JoinBF16Args {
    const float* X;  // FP32 rows (reference-order re-evaluation)
    int64_t ld;
    int dim;
    const uint16_t* Xb;  // bf16 rows
    int64_t ldb;
    const float* nrm;
    const float* err;  // per-row rounding error bound (non-exact rows)
    const int32_t* new_idx;
    const int32_t* old_idx;
    int C;
    int64_t row_begin, row_end;
    const float* hd;
    int k;
    int exact;
    float g_d, g_r;
    JoinOut out;
    unsigned long long* rescored;  // kKeptSlots counters
};

constexpr int kJoinQueue = 1024;  // re-evaluation queue entries per wave (flushed when 256 more might not fit)
template <int TA, int TB, bool EXACT>
__global__ __launch_bounds__(256) void k_join_bf16(JoinBF16Args a) {
    constexpr int W = 4;
    __shared__ int32_t ids[W][TB * 16];
    __shared__ float roots[W][TB * 16];
    __shared__ float nr[W][TB * 16];
    __shared__ float er[W][EXACT ? 1 : TB * 16];
    __shared__ uint16_t rq[W][EXACT ? 1 : kJoinQueue];  // queued pair slots j * 2C + kk
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const unsigned long long lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    const int C = a.C, C2 = 2 * C;
    const float u24 = 1.0f / 16777216.0f;
    const int m = lane & 15, kq = (lane >> 4) * 4;
    unsigned long long resc = 0, kept = 0;
    for (int64_t i = a.row_begin + (int64_t)blockIdx.x * W + w; i < a.row_end; i += (int64_t)gridDim.x * W) {
        for (int c = lane; c < TB * 16; c += 64) {
            int32_t id = -1;
            if (c < C) id = a.new_idx[i * C + c];
            else if (c < C2) id = a.old_idx[i * C + (c - C)];
            ids[w][c] = id;
            roots[w][c] = id >= 0 ? a.hd[(int64_t)id * a.k] : 0.f;
            nr[w][c] = id >= 0 ? a.nrm[id] : 0.f;
            if (!EXACT) er[w][c] = id >= 0 ? a.err[id] : 0.f;
        }
        wave_sync();
        const int64_t r0 = i - a.row_begin;
        JRec* U = a.out.U + a.out.rowoff[r0];
        unsigned run = 0, recs = 0;
        // operand rows: absent candidates read row 0 and are zeroed after the load, so every load of a batch
        // is unconditional and KB k-steps of fragments are in flight together.  The new candidates (the A
        // operand) are the first candidates of the B operand, so A's fragments are B's first TA tiles, zeroed
        // past the new candidates.
        const uint16_t* brow[TB];
        bool aval[TA], bval[TB];
#pragma unroll
        for (int t = 0; t < TA; ++t) aval[t] = 16 * t + m < C && ids[w][16 * t + m] >= 0;
#pragma unroll
        for (int t = 0; t < TB; ++t) {
            const int32_t id = ids[w][16 * t + m];
            bval[t] = id >= 0;
            brow[t] = a.Xb + (int64_t)(id >= 0 ? id : 0) * a.ldb + kq;
        }
        v4f acc[TA][TB];
#pragma unroll
        for (int x = 0; x < TA; ++x)
#pragma unroll
            for (int y = 0; y < TB; ++y) acc[x][y] = v4f{0.f, 0.f, 0.f, 0.f};
        constexpr int KB = 2;
        const v4s z4 = {0, 0, 0, 0};
        for (int64_t k0 = 0; k0 < a.ldb; k0 += 16 * KB) {
            v4s B[KB][TB];
#pragma unroll
            for (int kb = 0; kb < KB; ++kb) {
                const int64_t ko = min(k0 + 16 * kb, a.ldb - 16);
#pragma unroll
                for (int t = 0; t < TB; ++t) B[kb][t] = *reinterpret_cast<const v4s*>(brow[t] + ko);
            }
#pragma unroll
            for (int kb = 0; kb < KB; ++kb) {
                const bool live = k0 + 16 * kb < a.ldb;  // uniform: steps past ldb add zeros
#pragma unroll
                for (int x = 0; x < TA; ++x)
#pragma unroll
                    for (int y = 0; y < TB; ++y)
                        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(
                            aval[x] && live ? B[kb][x] : z4, bval[y] ? B[kb][y] : z4, acc[x][y], 0, 0, 0);
            }
        }
        // kept pairs go to the row's own region in any order (the slot carries the key)
        auto emit = [&](bool keep, bool ep, bool eq, int slot, float d) {
            const unsigned long long bal = __ballot(keep);
            if (keep) {
                U[run + __popcll(bal & lt)] = JRec{(uint32_t)slot | (ep ? kEmitP : 0u) | (eq ? kEmitQ : 0u), d};
                recs += (unsigned)ep + (unsigned)eq;
            }
            run += (unsigned)__popcll(bal);
        };
        // non-exact rows: the pairs inside the margin are queued, then re-evaluated with every lane busy
        unsigned nq = 0;
        auto flush = [&]() {
            wave_sync();
            for (unsigned e0 = 0; e0 < nq; e0 += 64) {
                const unsigned e = e0 + lane;
                bool keep = false, ep = false, eq = false;
                int slot = 0;
                float d = 0.f;
                if (e < nq) {
                    slot = rq[w][e];
                    const int j = slot / C2, kk = slot - j * C2;
                    d = d_sqeuclidean(a.X + (int64_t)ids[w][j] * a.ld, a.X + (int64_t)ids[w][kk] * a.ld, a.dim);
                    ep = d < roots[w][j];
                    eq = d < roots[w][kk];
                    keep = ep || eq;
                }
                emit(keep, ep, eq, slot, d);
            }
            if (lane == 0) resc += nq;  // nq is wave-uniform: counted once (the lanes' counts are summed below)
            nq = 0;
            wave_sync();
        };
#pragma unroll
        for (int x = 0; x < TA; ++x)
#pragma unroll
            for (int y = 0; y < TB; ++y) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int j = 16 * x + (lane >> 4) * 4 + r;  // D[row][col] layout of 16x16 MFMA
                    const int kk = 16 * y + m;
                    bool keep = false, ep = false, eq = false, need = false;
                    float d = 0.f;
                    if (j < C && kk < C2) {
                        const int32_t p = ids[w][j], q = ids[w][kk];
                        const bool valid = p >= 0 && q >= 0 && (kk < C ? kk > j : p != q);
                        if (valid) {
                            const float dt = (nr[w][j] + nr[w][kk]) - 2.0f * acc[x][y][r];
                            if (EXACT) {  // no re-evaluation code in this instance: fewer registers
                                d = dt;
                                ep = d < roots[w][j];
                                eq = d < roots[w][kk];
                                keep = ep || eq;
                            } else {
                                const float thr = fmaxf(roots[w][j], roots[w][kk]);
                                if (!(thr < __builtin_inff())) {
                                    need = true;
                                } else {
                                    const float Db = thr / (1.0f - a.g_r);
                                    const float e = er[w][j] + er[w][kk];
                                    const float F = (2.0f * a.g_d + 4.0f * u24) * (nr[w][j] + nr[w][kk]) *
                                                    (1.0f + 2.0f * a.g_d);
                                    need = dt <= Db + 1.05f * (2.0f * e * sqrtf(Db) + e * e + F) + 1e-30f;
                                }
                            }
                        }
                    }
                    if (EXACT) {
                        emit(keep, ep, eq, j * C2 + kk, d);
                    } else {
                        const unsigned long long bq = __ballot(need);
                        if (need) rq[w][nq + __popcll(bq & lt)] = (uint16_t)(j * C2 + kk);
                        nq += (unsigned)__popcll(bq);
                    }
                }
                if (!EXACT && nq > kJoinQueue - 256) flush();  // wave-uniform
            }
        if (!EXACT && nq) flush();
        wave_sync();
        for (int o = 32; o > 0; o >>= 1) recs += __shfl_xor(recs, o);
        if (lane == 0) {
            a.out.kc[r0] = (unsigned long long)run;
            a.out.krec[r0] = (unsigned long long)recs;
            kept += run;
        }
        wave_sync();
    }
    for (int o = 32; o > 0; o >>= 1) resc += __shfl_xor(resc, o);
    if (lane == 0 && resc) atomicAdd(&a.rescored[(blockIdx.x * W + w) % kKeptSlots], resc);
    if (lane == 0 && kept) atomicAdd(&a.out.kept[(blockIdx.x * W + w) % kKeptSlots], kept);
}

// Leaf pair distances on MFMA for exact-mode data (initialize_from_leaves, core/nndescent/mod.rs:203-300):
// the leaf's bf16 rows are staged in LDS, one wave per 16 x 16 block (A <= O) of the leaf's Gram matrix
// over 16x16x32 bf16 MFMAs, d = |a|^2 + |o|^2 - 2 G.  In exact mode this is the reference's FP32 value
// bit for bit (see the header), so the scratch D gets exactly what k_leaf_dist would write.
__device__ __forceinline__ void tri_pair16(int q, int m, int& a, int& o) {
    const int r = m * (m - 1) / 2 - 1 - q;
    int k = (int)((sqrtf(8.0f * (float)r + 1.0f) - 1.0f) * 0.5f);
    while ((k + 1) * (k + 2) / 2 <= r) ++k;
    while (k * (k + 1) / 2 > r) --k;
    a = m - 2 - k;
    o = m - 1 - (r - k * (k + 1) / 2);
}

// One 16 x 16 block of a Gram matrix on gfx950's 16x16x32 bf16 MFMA (K = ldb, a multiple of 16; an odd
// 16-column tail takes one 16x16x16 step).  ra / ro point at this lane's A row (block row + lane % 16) and
// O row in LDS; lane group g = lane / 16 holds k = 8 g .. 8 g + 7 of each 32-wide step for both operands, so
// every product pairs the same k.  Result: acc[r] = G(16 A + 4 g + r, 16 O + lane % 16).
__device__ __forceinline__ v4f gram_block16(const uint16_t* ra, const uint16_t* ro, int ldb, int g) {
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    int k0 = 0;
    for (; k0 + 32 <= ldb; k0 += 32)
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(*reinterpret_cast<const v8bf*>(ra + k0 + 8 * g),
                                                      *reinterpret_cast<const v8bf*>(ro + k0 + 8 * g), acc, 0, 0, 0);
    if (k0 < ldb)
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(*reinterpret_cast<const v4s*>(ra + k0 + 4 * g),
                                                        *reinterpret_cast<const v4s*>(ro + k0 + 4 * g), acc, 0, 0, 0);
    return acc;
}

__global__ __launch_bounds__(256) void k_leaf_dist_mfma(const uint16_t* __restrict__ Xb, int64_t ldb,
                                                        const float* __restrict__ nrm, const int32_t* __restrict__ tidx,
                                                        const int64_t* __restrict__ leaves,
                                                        const int64_t* __restrict__ n_leaves_dev, int mcap,
                                                        float* __restrict__ D, int64_t Dl4, unsigned long long* pairs) {
    extern __shared__ __align__(16) uint16_t bsm16[];
    const int B = blockDim.x, t = threadIdx.x;
    const int w = t >> 6, lane = t & 63, nw = B >> 6;
    const int lds_ld = (int)ldb + 8;  // bf16 elements per staged row (16-byte aligned, skewed banks)
    float* nr = (float*)(bsm16 + (int64_t)round_up((int64_t)mcap, (int64_t)16) * lds_ld);
    int32_t* rid = (int32_t*)(nr + round_up((int64_t)mcap, (int64_t)16));
    const int64_t n_leaves = *n_leaves_dev;
    const int mrow = lane & 15, kq = (lane >> 4) * 4;
    unsigned long long np = 0;
    for (int64_t leaf = blockIdx.x; leaf < n_leaves; leaf += gridDim.x) {
        const int64_t s = leaves[2 * leaf];
        const int m = (int)(leaves[2 * leaf + 1] - s);
        if (m > mcap || m < 2) continue;  // uniform
        np += (unsigned long long)m * (m - 1);
        const int mp = round_up(m, 16);
        const int v8 = (int)(ldb >> 3);  // 16-byte chunks per row
        // member ids and norms first, so that the row loads below do not wait on an id load each
        for (int i = t; i < mp; i += B) {
            const int32_t id = i < m ? tidx[s + i] : -1;
            rid[i] = id;
            nr[i] = id >= 0 ? nrm[id] : 0.f;
        }
        __syncthreads();
        for (int i0 = 0; i0 < mp * v8; i0 += 8 * B) {  // eight 16-byte row loads in flight per thread
            uint4 v[8];
            bool pad[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int i = min(i0 + u * B + t, mp * v8 - 1);
                const int r = i / v8, e = i - r * v8;
                const int32_t id = rid[r];
                pad[u] = id < 0;  // rows m..mp-1 are zero
                v[u] = *reinterpret_cast<const uint4*>(Xb + (int64_t)(id < 0 ? 0 : id) * ldb + e * 8);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int i = i0 + u * B + t;
                if (i < mp * v8) {
                    const int r = i / v8, e = i - r * v8;
                    *reinterpret_cast<uint4*>(bsm16 + r * lds_ld + e * 8) = pad[u] ? make_uint4(0, 0, 0, 0) : v[u];
                }
            }
        }
        __syncthreads();
        const int nb = mp >> 4;
        const int T = nb * (nb + 1) / 2;
        for (int q = w; q < T; q += nw) {
            int A, O;
            tri_pair16(q, nb + 1, A, O);  // A < O' over nb + 1 blocks -> A <= O = O' - 1
            O -= 1;
            const v4f acc = gram_block16(bsm16 + (16 * A + mrow) * lds_ld, bsm16 + (16 * O + mrow) * lds_ld, (int)ldb,
                                         lane >> 4);
            const int o = 16 * O + mrow;  // D[row][col] layout of the 16x16 MFMA: rows 16 A + kq + r
            if (A < O) {  // every a < o: o's row gets its four columns 16 A + kq .. + 3 in one 16-byte store
                if (o < m) {
                    float d[4];
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int a = 16 * A + kq + r;
                        d[r] = (nr[a] + nr[o]) - 2.0f * acc[r];
                        D[(s + a) * Dl4 + o] = d[r];
                    }
                    *reinterpret_cast<float4*>(D + (s + o) * Dl4 + 16 * A + kq) = make_float4(d[0], d[1], d[2], d[3]);
                }
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int a = 16 * A + kq + r;
                    if (a < o && o < m) {
                        const float d = (nr[a] + nr[o]) - 2.0f * acc[r];
                        D[(s + a) * Dl4 + o] = d;
                        D[(s + o) * Dl4 + a] = d;
                    }
                }
            }
        }
        __syncthreads();
    }
    if (t == 0 && np) atomicAdd(pairs, np);
}

// The same pair distances for the trees after the first few, writing only the pairs that can enter a heap:
// a push of d onto p's heap changes it only if d < p's root, and the roots only fall while pushes are applied,
// so the pairs with d >= p's root at the time this kernel reads it (before any push of this tree) are dropped
// for p.  Each unordered pair is evaluated once (16 x 16 blocks A <= O of the leaf's Gram matrix, the upper
// triangle inside a diagonal block; exact-mode distances are symmetric) and tested against both endpoints'
// roots.  Everything is written point-major, so that one push pass can replay several trees per point: a
// survivor goes to its leaf column in the member's D row (row = the member's id), the member's mask (Msk, 8
// words, bit o = leaf member o) gets its bit (the masks are gathered in LDS with atomic ORs and written once
// per leaf), and info[id] = (leaf start, leaf size) for every member of every leaf, including the leaves this
// kernel does not evaluate (one member, or more than mcap: the push pass evaluates those itself).
// surv counts the survivors.
// Launched with 256 threads (four waves) and mcap <= 256; NK16 = ldb / 16.  Wave w owns the row-block pairs
// (P, nb - 1 - P), P = w, w + 4, ... (nb + 1 block-pair columns each, balanced), keeps a row block's operands
// and its rows' norms, roots and ids in registers and walks the column blocks O >= A two at a time (two
// independent MFMA chains).  PRE (rows of at most 128 bf16): the leaves are software-pipelined - wave w stages
// rows 64 w .. 64 w + 63 of a leaf, 16 lanes per row, taking the row ids from its own lanes by shuffle; while a
// leaf is evaluated, the next leaf's rows (ids loaded one leaf earlier) are already in flight in registers.
template <int NK16, bool PRE>
__global__ __launch_bounds__(256) void k_leaf_surv_mfma(const uint16_t* __restrict__ Xb,
                                                        const float* __restrict__ nrm, const int32_t* __restrict__ tidx,
                                                        const int64_t* __restrict__ leaves,
                                                        const int64_t* __restrict__ n_leaves_dev, int mcap,
                                                        const float* __restrict__ hdist, int k, float* __restrict__ D,
                                                        int64_t Dl4, uint32_t* __restrict__ Msk, int2* __restrict__ info,
                                                        unsigned long long* surv) {
    constexpr int ldb = 16 * NK16, NK32 = NK16 / 2, v8 = ldb / 8;
    extern __shared__ __align__(16) uint16_t ssm16[];
    const int B = blockDim.x, t = threadIdx.x;
    const int w = t >> 6, lane = t & 63, nw = B >> 6;
    constexpr int lds_ld = ldb + 8;  // bf16 elements per staged row (16-byte aligned, skewed banks)
    const int mp16 = (int)round_up((int64_t)mcap, (int64_t)16);
    float* nr = (float*)(ssm16 + (int64_t)mp16 * lds_ld);
    float* rt = nr + mp16;  // each member's heap root
    int32_t* rid = (int32_t*)(rt + mp16);
    uint32_t* lmask = (uint32_t*)(rid + mp16);  // mp16 x 8 survivor mask words
    const int64_t n_leaves = *n_leaves_dev;
    const int64_t G = gridDim.x;
    const int mrow = lane & 15, kqi = lane >> 4, kq = kqi * 4;
    auto geom = [&](int64_t L, int64_t& s, int& m) {  // wave-uniform; m = 0 for a leaf this kernel skips
        s = 0;
        m = 0;
        if (L < n_leaves) {
            s = leaves[2 * L];
            m = (int)(leaves[2 * L + 1] - s);
            if (m > mcap || m < 2) m = 0;
        }
    };
    auto own_id = [&](int64_t s, int m) -> int32_t {  // the id of row 64 w + lane
        const int r = 64 * w + lane;
        return r < m ? tidx[s + r] : -1;
    };
    // PRE: the staged leaf's row 64 w + lane (id, norm, root) and chunk lane % 16 of rows 64 w + 4 u + lane / 16
    int32_t pid = -1;
    float pnr = 0.f, prt = 0.f;
    uint4 pv[16];
    auto issue = [&](int32_t id) {
        pid = id;
        pnr = id >= 0 ? nrm[id] : 0.f;
        prt = id >= 0 ? hdist[(int64_t)id * k] : 0.f;
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int32_t ru = __shfl(id, 4 * u + (lane >> 4));
            pv[u] = (ru >= 0 && mrow < v8) ? *reinterpret_cast<const uint4*>(Xb + (int64_t)ru * ldb + mrow * 8)
                                           : make_uint4(0, 0, 0, 0);
        }
    };
    int64_t s0, s1;
    int m0, m1;
    geom(blockIdx.x, s0, m0);
    geom(blockIdx.x + G, s1, m1);
    int32_t id1 = -1;
    if (PRE) {
        id1 = own_id(s1, m1);
        issue(own_id(s0, m0));
    }
    unsigned long long ns = 0;
    for (int64_t leaf = blockIdx.x; leaf < n_leaves; leaf += G) {
        const int64_t s = s0;
        const int m = m0;
        const int mp = round_up(m, 16);
        if (m) {
            for (int i = t; i < mp * 8; i += B) lmask[i] = 0u;
            if (PRE) {
                const int r = 64 * w + lane;
                if (r < mp) {
                    rid[r] = pid;
                    nr[r] = pnr;
                    rt[r] = prt;
                }
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const int row = 64 * w + 4 * u + (lane >> 4);
                    if (row < mp && mrow < v8) *reinterpret_cast<uint4*>(ssm16 + row * lds_ld + mrow * 8) = pv[u];
                }
            } else {
                for (int i = t; i < mp; i += B) rid[i] = i < m ? tidx[s + i] : -1;
                lds_barrier();
                // then the rows (eight 16-byte loads in flight per thread), norms and roots in one round trip
                for (int i0 = 0; i0 < mp * v8; i0 += 8 * B) {
                    uint4 v[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int i = min(i0 + u * B + t, mp * v8 - 1);
                        const int r = i / v8, e = i - r * v8;
                        const int32_t id = rid[r];
                        v[u] = *reinterpret_cast<const uint4*>(Xb + (int64_t)(id < 0 ? 0 : id) * ldb + e * 8);
                    }
                    if (i0 == 0)
                        for (int i = t; i < mp; i += B) {
                            const int32_t id = rid[i];
                            nr[i] = id >= 0 ? nrm[id] : 0.f;
                            rt[i] = id >= 0 ? hdist[(int64_t)id * k] : 0.f;
                        }
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int i = i0 + u * B + t;
                        if (i < mp * v8) {
                            const int r = i / v8, e = i - r * v8;
                            *reinterpret_cast<uint4*>(ssm16 + r * lds_ld + e * 8) =
                                rid[r] < 0 ? make_uint4(0, 0, 0, 0) : v[u];  // rows m..mp-1 are zero
                        }
                    }
                }
            }
            lds_barrier();
        }
        // the next leaf's rows go in flight now (their ids were loaded one leaf ago), and the ids of the one after
        int64_t s2;
        int m2;
        geom(leaf + 2 * G, s2, m2);
        if (PRE) {
            issue(id1);
            id1 = own_id(s2, m2);
        }
        s0 = s1;
        m0 = m1;
        s1 = s2;
        m1 = m2;
        if (!m) {  // uniform: a leaf of one member or above mcap, evaluated by the push pass; its members' info
            if (leaf < n_leaves) {
                const int64_t ls = leaves[2 * leaf];
                const int64_t lm = leaves[2 * leaf + 1] - ls;
                for (int64_t i = t; i < lm; i += B) info[tidx[ls + i]] = make_int2((int)ls, (int)lm);
            }
            continue;
        }
        const int nb = mp >> 4;
        for (int P = w; P < ((nb + 1) >> 1); P += nw) {
            for (int h = 0; h < 2; ++h) {
                const int A = h ? nb - 1 - P : P;  // wave-uniform row block
                if (h && A == P) break;            // the middle block of an odd count
                v8bf a32[NK32 > 0 ? NK32 : 1];
                v4s a16;
                {
                    const uint16_t* ra = ssm16 + (16 * A + mrow) * lds_ld;
#pragma unroll
                    for (int kk = 0; kk < NK32; ++kk) a32[kk] = *reinterpret_cast<const v8bf*>(ra + 32 * kk + 8 * kqi);
                    if (NK16 & 1) a16 = *reinterpret_cast<const v4s*>(ra + 32 * NK32 + 4 * kqi);
                }
                const int x0 = 16 * A + kq;  // this lane's rows x0 .. x0 + 3
                const float4 nx4 = *reinterpret_cast<const float4*>(nr + x0);
                const float4 rx4 = *reinterpret_cast<const float4*>(rt + x0);
                const int4 ix4 = *reinterpret_cast<const int4*>(rid + x0);
                const float nx[4] = {nx4.x, nx4.y, nx4.z, nx4.w}, rx[4] = {rx4.x, rx4.y, rx4.z, rx4.w};
                const int32_t ix[4] = {ix4.x, ix4.y, ix4.z, ix4.w};  // the rows' ids: their D rows
                // B operands of column blocks (O, O + 1), the next pair's loaded while this pair's results are
                // tested (an odd last block repeats O; its results are unused)
                v8bf b0[NK32 > 0 ? NK32 : 1], b1[NK32 > 0 ? NK32 : 1];
                v4s t0, t1;
                float cn[2], cr[2];  // the two columns' norms and roots
                int32_t ci[2];       // and ids
                auto load_b = [&](int O) {
                    const int O1 = O + 1 < nb ? O + 1 : O;
                    cn[0] = nr[16 * O + mrow];
                    cr[0] = rt[16 * O + mrow];
                    ci[0] = rid[16 * O + mrow];
                    cn[1] = nr[16 * O1 + mrow];
                    cr[1] = rt[16 * O1 + mrow];
                    ci[1] = rid[16 * O1 + mrow];
                    const uint16_t* rb0 = ssm16 + (16 * O + mrow) * lds_ld;
                    const uint16_t* rb1 = ssm16 + (16 * O1 + mrow) * lds_ld;
#pragma unroll
                    for (int kk = 0; kk < NK32; ++kk) {
                        b0[kk] = *reinterpret_cast<const v8bf*>(rb0 + 32 * kk + 8 * kqi);
                        b1[kk] = *reinterpret_cast<const v8bf*>(rb1 + 32 * kk + 8 * kqi);
                    }
                    if (NK16 & 1) {
                        t0 = *reinterpret_cast<const v4s*>(rb0 + 32 * NK32 + 4 * kqi);
                        t1 = *reinterpret_cast<const v4s*>(rb1 + 32 * NK32 + 4 * kqi);
                    }
                };
                load_b(A);
                for (int O = A; O < nb; O += 2) {  // two column blocks per step: independent MFMA chains
                    v4f acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int kk = 0; kk < NK32; ++kk) {
                        acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a32[kk], b0[kk], acc0, 0, 0, 0);
                        acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a32[kk], b1[kk], acc1, 0, 0, 0);
                    }
                    if (NK16 & 1) {
                        acc0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a16, t0, acc0, 0, 0, 0);
                        acc1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a16, t1, acc1, 0, 0, 0);
                    }
                    const float cnow_n[2] = {cn[0], cn[1]}, cnow_r[2] = {cr[0], cr[1]};
                    const int32_t cnow_i[2] = {ci[0], ci[1]};
                    if (O + 2 < nb) load_b(O + 2);  // wave-uniform
#pragma unroll
                    for (int hb = 0; hb < 2; ++hb) {
                        const int Oh = O + hb;
                        if (Oh >= nb) break;  // wave-uniform
                        const v4f acc = hb ? acc1 : acc0;
                        const int o = 16 * Oh + mrow;  // this lane's column
                        const float no = cnow_n[hb], ro = cnow_r[hb];
                        float* const Do = D + (int64_t)cnow_i[hb] * Dl4;  // column o's D row
                        const bool col_ok = o < m;
                        bool sx[4], so[4];
                        float d[4];
                        uint64_t bx[4];
                        uint64_t any_so = 0;
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            d[r] = (nx[r] + no) - 2.0f * acc[r];
                            const bool pair = col_ok & (x0 + r < o);  // each unordered pair once, both directions
                            sx[r] = pair & (d[r] < rx[r]);            // o enters x's heap
                            so[r] = pair & (d[r] < ro);               // x enters o's heap
                            bx[r] = __ballot(sx[r]);
                            const uint64_t bo = __ballot(so[r]);
                            any_so |= bo;
                            ns += (unsigned)(__popcll(bx[r]) + __popcll(bo));  // wave-uniform count
                        }
                        uint32_t nib = 0u;  // rows x that survive for column o's member (o's mask bits)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int x = x0 + r;
                            if (bx[r]) {  // wave-uniform
                                if (sx[r]) D[(int64_t)ix[r] * Dl4 + o] = d[r];
                                const uint32_t half = (uint32_t)(bx[r] >> (16 * kqi)) & 0xffffu;
                                if ((mrow == 0) & (half != 0u))
                                    atomicOr(&lmask[x * 8 + (Oh >> 1)], half << (16 * (Oh & 1)));
                            }
                            nib |= (uint32_t)so[r] << r;
                        }
                        if (!any_so) continue;  // wave-uniform: no mask bits for any column this block
                        // o's D row, columns x0 .. x0 + 3.  Off the diagonal every one of them is a pair of this
                        // block and no other (A < Oh puts x0 + 3 below o), so a lane with a survivor writes all
                        // four at once: the extra values land in slots that only this pair owns and its mask bit
                        // leaves unread.  On the diagonal the pairs x < o are written one by one.
                        if (A < Oh) {
                            if (nib) *reinterpret_cast<float4*>(Do + x0) = make_float4(d[0], d[1], d[2], d[3]);
                        } else {
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                                if (so[r]) Do[x0 + r] = d[r];
                        }
                        // row o, columns 16 A + kq .. + 3: each lane group ORs its nibble (no return, no wait)
                        if (nib) atomicOr(&lmask[o * 8 + (A >> 1)], nib << (kq + 16 * (A & 1)));
                    }
                }
            }
        }
        lds_barrier();
        for (int i = t; i < m * 8; i += B) Msk[(int64_t)rid[i >> 3] * 8 + (i & 7)] = lmask[i];
        for (int i = t; i < m; i += B) info[rid[i]] = make_int2((int)s, m);
        lds_barrier();
    }
    if (lane == 0 && ns) atomicAdd(surv, ns);  // ns is the wave's count (the same in every lane)
}

size_t leaf_surv_mfma_lds(int64_t ldb, int mcap) {
    if (ldb > 256 || mcap > 256) return SIZE_MAX;  // instances cover rows of up to 256 bf16, leaves of up to 256
    const size_t mp16 = (size_t)round_up((int64_t)mcap, (int64_t)16);
    return mp16 * (ldb + 8) * 2 + mp16 * 12 + mp16 * 32;
}

void launch_leaf_surv_mfma(hipStream_t s, const Bf16Rows& R, const int32_t* tree_indices, const int64_t* leaves,
                           const int64_t* n_leaves_dev, int64_t max_leaves, int mcap, const float* hdist, int k,
                           float* D, int64_t Dl4, uint32_t* Msk, int2* info, unsigned long long* surv) {
    if (mcap > 256) throw Error{RINND_ERR_INVALID_ARG, "survivor leaf kernel: leaves above 256 members"};
    if (Dl4 % 4 || ((uintptr_t)D & 15))  // the kernel writes four D columns with one 16-byte store
        throw Error{RINND_ERR_INVALID_ARG, "survivor leaf kernel: D rows must be 16-byte aligned"};
    // rows of at most 128 bf16: persistent, two workgroups per CU (the LDS limit), each pipelining its leaves
    const unsigned grid = (unsigned)std::min<int64_t>(std::max<int64_t>(max_leaves, 1),
                                                      R.ldb <= 128 ? 2 * (int64_t)device_cus() : 4096);
    const size_t lds = leaf_surv_mfma_lds(R.ldb, mcap);
#define RINND_SURV_CASE(NKV)                                                                                     \
    case NKV:                                                                                                    \
        hipLaunchKernelGGL((k_leaf_surv_mfma<NKV, (NKV <= 8)>), dim3(grid), dim3(256), lds, s, R.Xb.p, R.nrm.p,   \
                           tree_indices, leaves, n_leaves_dev, mcap, hdist, k, D, Dl4, Msk, info, surv);         \
        break;
    switch (R.ldb / 16) {
        RINND_SURV_CASE(1) RINND_SURV_CASE(2) RINND_SURV_CASE(3) RINND_SURV_CASE(4) RINND_SURV_CASE(5)
        RINND_SURV_CASE(6) RINND_SURV_CASE(7) RINND_SURV_CASE(8) RINND_SURV_CASE(9) RINND_SURV_CASE(10)
        RINND_SURV_CASE(11) RINND_SURV_CASE(12) RINND_SURV_CASE(13) RINND_SURV_CASE(14) RINND_SURV_CASE(15)
        RINND_SURV_CASE(16)
        default: throw Error{RINND_ERR_INVALID_ARG, "survivor leaf kernel: row width above 256"};
    }
#undef RINND_SURV_CASE
}

size_t leaf_dist_mfma_lds(int64_t ldb, int mcap) {
    return (size_t)round_up((int64_t)mcap, (int64_t)16) * (ldb + 8) * 2 + (size_t)round_up((int64_t)mcap, (int64_t)16) * 8;
}

void launch_leaf_dist_mfma(hipStream_t s, const Bf16Rows& R, const int32_t* tree_indices, const int64_t* leaves,
                           const int64_t* n_leaves_dev, int64_t max_leaves, int mcap, float* D, int64_t Dl4,
                           unsigned long long* pairs) {
    if (Dl4 % 4 || ((uintptr_t)D & 15))  // the kernel writes four D columns with one 16-byte store
        throw Error{RINND_ERR_INVALID_ARG, "leaf distance kernel: D rows must be 16-byte aligned"};
    const unsigned grid = (unsigned)std::min<int64_t>(std::max<int64_t>(max_leaves, 1), 4096);
    hipLaunchKernelGGL(k_leaf_dist_mfma, dim3(grid), dim3(256), leaf_dist_mfma_lds(R.ldb, mcap), s, R.Xb.p, R.ldb,
                       R.nrm.p, tree_indices, leaves, n_leaves_dev, mcap, D, Dl4, pairs);
}

void prepare_bf16_rows(hipStream_t s, const float* X, int64_t ld, int64_t n, int dim, Bf16Rows& R) {
    R.aux.ensure(2 + (size_t)dim);
    int* flag = (int*)R.aux.p;
    unsigned long long* mx = (unsigned long long*)(R.aux.p + 1);
    double* sums = R.aux.p + 2;
    HIP_TRY(hipMemsetAsync(R.aux.p, 0, 16, s));
    R.ldb = round_up(dim, 16);
    R.Xb.ensure((size_t)std::max<int64_t>(n, 1) * R.ldb);
    R.nrm.ensure(std::max<int64_t>(n, 1));
    // the exactness test and the uncentred rows in one pass (kept when the data are exact)
    const unsigned cg = (unsigned)std::min<int64_t>(std::max<int64_t>(ceil_div(n, 16), 1), 1024);
    if (((uintptr_t)X & 15) == 0 && ld % 4 == 0)
        hipLaunchKernelGGL(k_check_to_bf16<true>, dim3(cg), dim3(256), 0, s, X, ld, n, dim, R.Xb.p, R.ldb, R.nrm.p,
                           flag, mx);
    else
        hipLaunchKernelGGL(k_check_to_bf16<false>, dim3(cg), dim3(256), 0, s, X, ld, n, dim, R.Xb.p, R.ldb, R.nrm.p,
                           flag, mx);
    int f = 0;
    unsigned long long mb = 0;
    HIP_TRY(hipMemcpyAsync(&f, flag, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&mb, mx, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    double maxnorm;
    std::memcpy(&maxnorm, &mb, 8);
    // integers in [-256, 256] with |x|^2 <= 2^22, or non-negative with |x|^2 <= 2^23
    R.exact = ((f & 1) == 0 && (maxnorm <= 4194304.0 || ((f & 2) == 0 && maxnorm <= 8388608.0))) ? 1 : 0;
    if (!R.exact) {  // centred rows
        HIP_TRY(hipMemsetAsync(sums, 0, dim * 8, s));
        R.mu.ensure(dim);
        hipLaunchKernelGGL(k_col_sums, dim3(1024), dim3(256), 0, s, X, ld, n, dim, sums);
        hipLaunchKernelGGL(k_col_mean, dim3(ceil_div(dim, 256)), dim3(256), 0, s, sums, n, dim, R.mu.p);
        HIP_TRY(hipStreamSynchronize(s));
        R.err.ensure(std::max<int64_t>(n, 1));
        const double u24 = 1.0 / 16777216.0;
        hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)std::min<int64_t>(ceil_div(n, 4), 1 << 16)), dim3(256), 0, s, X,
                           ld, n, dim, R.mu.p, R.Xb.p, R.ldb, R.nrm.p, R.err.p,
                           (float)(1.01 * dim * u24 / (1.0 - dim * u24)));
    }
    HIP_TRY(hipGetLastError());
}

template <bool EXACT>
static void launch_join_bf16_tiles(hipStream_t s, unsigned grid, int C, const JoinBF16Args& a) {
    if (C <= 8) hipLaunchKernelGGL((k_join_bf16<1, 1, EXACT>), dim3(grid), dim3(256), 0, s, a);
    else if (C <= 16) hipLaunchKernelGGL((k_join_bf16<1, 2, EXACT>), dim3(grid), dim3(256), 0, s, a);
    else if (C <= 32) hipLaunchKernelGGL((k_join_bf16<2, 4, EXACT>), dim3(grid), dim3(256), 0, s, a);
    else if (C <= 48) hipLaunchKernelGGL((k_join_bf16<3, 6, EXACT>), dim3(grid), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_join_bf16<4, 8, EXACT>), dim3(grid), dim3(256), 0, s, a);
}

void launch_join_bf16(hipStream_t s, const float* X, int64_t ld, int dim, const Bf16Rows& R,
                      const int32_t* new_idx, const int32_t* old_idx, int C, int64_t row_begin,
                      int64_t row_end, const float* heap_dist, int k, const JoinOut& out,
                      unsigned long long* rescored) {
    if (row_end <= row_begin) return;
    JoinBF16Args a{};
    a.X = X;
    a.ld = ld;
    a.dim = dim;
    a.Xb = R.Xb.p;
    a.ldb = R.ldb;
    a.nrm = R.nrm.p;
    a.err = R.exact ? nullptr : R.err.p;
    a.new_idx = new_idx;
    a.old_idx = old_idx;
    a.C = C;
    a.row_begin = row_begin;
    a.row_end = row_end;
    a.hd = heap_dist;
    a.k = k;
    a.exact = R.exact;
    const double u24 = 1.0 / 16777216.0;
    a.g_d = (float)(dim * u24 / (1.0 - dim * u24));
    a.g_r = (float)(1.01 * (dim + 8) * u24);
    a.out = out;
    a.rescored = rescored;
    const unsigned grid = (unsigned)std::min<int64_t>(ceil_div(row_end - row_begin, 4), 1 << 18);
    if (a.exact) launch_join_bf16_tiles<true>(s, grid, C, a);
    else launch_join_bf16_tiles<false>(s, grid, C, a);
}

}  // namespace rinnd
