// kernels_topk.hip -- per-query k smallest cRMSD values of a score row.
//
// One work-group per query row.  Non-negative floats order like their bit
// patterns, so the k-th smallest key is found by a three-pass radix select
// (11 + 11 + 10 bits, LDS histograms).  Every key below the threshold is kept
// and keys equal to it are kept in index order until k are taken, so ties
// resolve to the lower library index.  The k survivors are sorted by
// (key, index) with a bitonic sort in LDS.
#include "kernels.hpp"

namespace bs {
namespace {

constexpr int kTopkMax = 1024;

__device__ inline uint32_t key_of(float v)
{
    uint32_t u = __float_as_uint(v);
    if (v != v) return 0x7FC00000u;          // NaN sorts last
    if (u & 0x80000000u) return 0u;          // -0 and negatives (never produced) first
    return u;
}

__global__ __launch_bounds__(256) void topk_kernel(const float* __restrict__ scores, int n,
                                                   int k, uint32_t* __restrict__ out_idx,
                                                   float* __restrict__ out_val)
{
    __shared__ uint32_t hist[2048];
    __shared__ unsigned long long cand[kTopkMax];
    __shared__ uint32_t s_bin, s_need, s_cnt, s_eq_taken;
    __shared__ uint32_t s_wave[4];
    const float* row = scores + (int64_t)blockIdx.x * n;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;

    uint32_t prefix = 0, pmask = 0, need = (uint32_t)k;
    for (int pass = 0; pass < 3; ++pass) {
        const int shift = pass == 0 ? 21 : (pass == 1 ? 10 : 0);
        const int nb = pass < 2 ? 2048 : 1024;
        for (int b = tid; b < nb; b += 256) hist[b] = 0;
        __syncthreads();
        for (int e = tid; e < n; e += 256) {
            const uint32_t key = key_of(row[e]);
            if ((key & pmask) == prefix) atomicAdd(&hist[(key >> shift) & (nb - 1)], 1u);
        }
        __syncthreads();
        if (tid == 0) {
            uint32_t cum = 0;
            int b = 0;
            for (; b < nb; ++b) {
                if (cum + hist[b] >= need) break;
                cum += hist[b];
            }
            s_bin = (uint32_t)b;
            s_need = need - cum;
        }
        __syncthreads();
        prefix |= s_bin << shift;
        pmask |= (uint32_t)(nb - 1) << shift;
        need = s_need;
        __syncthreads();
    }
    const uint32_t T = prefix;  // k-th smallest key; `need` copies of it are taken
    if (tid == 0) {
        s_cnt = 0;
        s_eq_taken = 0;
    }
    __syncthreads();
    for (int base = 0; base < n; base += 256) {
        const int e = base + tid;
        const uint32_t key = e < n ? key_of(row[e]) : 0xFFFFFFFFu;
        const bool lt = e < n && key < T;
        const bool eq = e < n && key == T;
        // rank of this equal key among equal keys of the chunk, in index order
        const uint64_t m = __ballot(eq);
        const uint32_t below = (uint32_t)__popcll(m & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
        if (lane == 0) s_wave[wv] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t off = s_eq_taken;
        for (int w = 0; w < wv; ++w) off += s_wave[w];
        const bool take = lt || (eq && off + below < need);
        if (take) {
            const uint32_t pos = atomicAdd(&s_cnt, 1u);
            if (pos < kTopkMax) cand[pos] = ((unsigned long long)key << 32) | (uint32_t)e;
        }
        __syncthreads();
        if (tid == 0) s_eq_taken += s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
        __syncthreads();
    }
    // bitonic sort of the k candidates (padded to a power of two)
    int P = 1;
    while (P < k) P <<= 1;
    for (int v = tid; v < P; v += 256)
        if (v >= k) cand[v] = ~0ull;
    __syncthreads();
    for (int size = 2; size <= P; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int v = tid; v < P; v += 256) {
                const int w = v ^ stride;
                if (w > v) {
                    const bool up = (v & size) == 0;
                    const unsigned long long x = cand[v], y = cand[w];
                    if ((x > y) == up) {
                        cand[v] = y;
                        cand[w] = x;
                    }
                }
            }
            __syncthreads();
        }
    for (int v = tid; v < k; v += 256) {
        const unsigned long long c = cand[v];
        out_idx[(int64_t)blockIdx.x * k + v] = (uint32_t)(c & 0xFFFFFFFFu);
        out_val[(int64_t)blockIdx.x * k + v] = __uint_as_float((uint32_t)(c >> 32));
    }
}

}  // namespace

void launch_topk(const float* scores, int n_q, int n_lib, int k, uint32_t* idx, float* val,
                 hipStream_t s)
{
    if (n_q <= 0) return;
    topk_kernel<<<n_q, 256, 0, s>>>(scores, n_lib, k, idx, val);
}

}  // namespace bs
