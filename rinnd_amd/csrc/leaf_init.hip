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
#pragma unroll 4
    for (int j = 0; j < k; ++j)
        if (hi[j * B + c] == x) return;
    int pos = 0, child = 1;  // replace the root and sift down (neighbor_heap.rs:207-243)
    while (child < k) {
        const int right = child + 1;
        if (right < k && hd[child * B + c] < hd[right * B + c]) child = right;
        const float dc = hd[child * B + c];
        if (d >= dc) break;
        hd[pos * B + c] = dc;
        hi[pos * B + c] = hi[child * B + c];
        hf[pos * B + c] = hf[child * B + c];
        pos = child;
        child = 2 * pos + 1;
    }
    hd[pos * B + c] = d;
    hi[pos * B + c] = x;
    hf[pos * B + c] = 1;
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

static int init_block(int64_t leaf_size) {
    return (int)std::min<int64_t>(256, round_up(std::max<int64_t>(leaf_size, 1), (int64_t)64));
}
static size_t init_heap_bytes(int B, int k) {
    return (size_t)round_up((int64_t)B * k * 2 + (B * k + 3) / 4, 4) * 4;
}

bool leaf_init_supported(int k, int64_t leaf_size) {
    return init_heap_bytes(init_block(leaf_size), k) <= (size_t)kInitLdsBudget;
}

void launch_leaf_init(hipStream_t s, int metric, const float* X, int64_t ld, int dim, const int32_t* tree_indices,
                      const int64_t* leaves, const int64_t* n_leaves_dev, int64_t max_leaves, int64_t leaf_size,
                      int32_t* idx, float* dist, uint8_t* flags, int k) {
    if (max_leaves <= 0) return;
    const int B = init_block(leaf_size);
    const size_t heap = init_heap_bytes(B, k);
    const int64_t ldl = ld + 4;
    const int64_t cap_rows = ((int64_t)kInitLdsBudget - (int64_t)heap) / (ldl * 4);
    const int stage_cap = (int)std::max<int64_t>(0, std::min<int64_t>(cap_rows, std::max<int64_t>(leaf_size, 1)));
    const unsigned grid = (unsigned)std::min<int64_t>(max_leaves, 8192);  // persistent over the leaves
    if (stage_cap > 0)
        hipLaunchKernelGGL(k_leaf_init<true>, dim3(grid), dim3(B), heap + (size_t)stage_cap * ldl * 4, s, metric, X,
                           ld, dim, tree_indices, leaves, n_leaves_dev, stage_cap, ldl, idx, dist, flags, k);
    hipLaunchKernelGGL(k_leaf_init<false>, dim3(std::min<unsigned>(grid, 4096)), dim3(B), heap, s, metric, X, ld,
                       dim, tree_indices, leaves, n_leaves_dev, stage_cap, ldl, idx, dist, flags, k);
    HIP_TRY(hipGetLastError());
}

}  // namespace rinnd
