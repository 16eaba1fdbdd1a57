// kernels_topk.hip -- per-query k smallest cRMSD values of a score row, ties
// to the lower library index, sorted by (value, index).
//
// Two stages so that a request of a few queries still spreads over the chip:
//  1. topk_chunk_kernel: one work-group per (chunk of a row, query).  The
//     chunk's k-th smallest key is found by a three-pass radix select (11 + 11
//     + 10 bits; non-negative floats order like their bit patterns) with
//     per-wave LDS histograms and a block-wide scan over the bins; the keys
//     below it and, in index order, enough keys equal to it are written out as
//     (key << 32 | index) candidates.
//  2. topk_merge_kernel: one work-group per query sorts the row's candidates
//     (at most kMergeMax, chunk sizes are chosen so) with a bitonic sort in
//     LDS and writes the first k.
// The k smallest (value, index) pairs of a row are among the union of every
// chunk's k smallest, so the result is exact.
#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

namespace bs {
namespace {

constexpr int kTopkMax = 1024;
constexpr int kMergeMax = 8192;
constexpr int kChunkThreads = 256;
constexpr int kChunkWaves = kChunkThreads / 64;
constexpr int kMinChunk = 4096;
constexpr int kCachedKeys = kMinChunk / kChunkThreads;  // keys per thread kept in registers

__device__ inline uint32_t key_of(float v)
{
    uint32_t u = __float_as_uint(v);
    if (v != v) return 0x7FC00000u;  // NaN sorts after every number
    if (u & 0x80000000u) return 0u;  // -0 and negatives (never produced) first
    return u;
}

// Exclusive scan of one value per thread over the block (256 threads);
// returns the thread's prefix, *total gets the block sum.
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_wave, uint32_t* total)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_wave[w] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int k = 0; k < kChunkWaves; ++k) {
        const uint32_t t = s_wave[k];
        if (k < w) before += t;
        all += t;
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}

__global__ __launch_bounds__(kChunkThreads) void topk_chunk_kernel(const float* __restrict__ scores, int n,
                                                                   int k, int chunk,
                                                                   unsigned long long* __restrict__ cand)
{
    __shared__ uint32_t hist[kChunkWaves][2048];
    __shared__ uint32_t s_wave[kChunkWaves];
    __shared__ uint32_t s_bin, s_need, s_cnt;
    __shared__ uint32_t eq_cnt[kChunkWaves * 129];
    const int q = blockIdx.y, c = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int base = c * chunk;
    const int len = min(chunk, n - base);
    const int kc = min(k, len);
    const float* row = scores + (int64_t)q * n + base;
    unsigned long long* out = cand + ((int64_t)q * gridDim.x + c) * k;

    // the thread's first kCachedKeys keys (elements tid, tid + 256, ...) are
    // read once and kept in registers for every pass; a longer chunk reads
    // the rest from memory each time
    uint32_t keys[kCachedKeys];
#pragma unroll
    for (int j = 0; j < kCachedKeys; ++j) {
        const int e = j * kChunkThreads + tid;
        keys[j] = e < len ? key_of(row[e]) : 0xFFFFFFFFu;
    }
    auto key_at = [&](int j, int e) { return j < kCachedKeys ? keys[j] : key_of(row[e]); };
    const int iters = (len + kChunkThreads - 1) / kChunkThreads;  // <= 128 (chunk <= 32768)

    uint32_t prefix = 0, pmask = 0, need = (uint32_t)kc;
    for (int pass = 0; pass < 3; ++pass) {
        const int shift = pass == 0 ? 21 : (pass == 1 ? 10 : 0);
        const int nb = pass < 2 ? 2048 : 1024;
        for (int b = tid; b < kChunkWaves * 2048; b += kChunkThreads) (&hist[0][0])[b] = 0;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kCachedKeys; ++j) {
            const uint32_t key = keys[j];
            if (j * kChunkThreads + tid < len && (key & pmask) == prefix)
                atomicAdd(&hist[wv][(key >> shift) & (nb - 1)], 1u);
        }
        for (int e = kCachedKeys * kChunkThreads + tid; e < len; e += kChunkThreads) {
            const uint32_t key = key_of(row[e]);
            if ((key & pmask) == prefix) atomicAdd(&hist[wv][(key >> shift) & (nb - 1)], 1u);
        }
        __syncthreads();
        // thread t owns bins [8t, 8t + 8): the bin where the running count reaches `need`
        const int per = nb / kChunkThreads;
        uint32_t own[8], sum = 0;
        for (int i = 0; i < per; ++i) {
            const int b = tid * per + i;
            own[i] = hist[0][b] + hist[1][b] + hist[2][b] + hist[3][b];
            sum += own[i];
        }
        uint32_t total;
        const uint32_t before = block_exclusive_scan(sum, s_wave, &total);
        if (before < need && before + sum >= need) {
            uint32_t cum = before;
            int i = 0;
            for (; i < per - 1; ++i) {
                if (cum + own[i] >= need) break;
                cum += own[i];
            }
            s_bin = (uint32_t)(tid * per + i);
            s_need = need - cum;
        }
        __syncthreads();
        prefix |= s_bin << shift;
        pmask |= (uint32_t)(nb - 1) << shift;
        need = s_need;
        __syncthreads();
    }
    const uint32_t T = prefix;  // the kc-th smallest key; `need` copies of it are taken
    // rank the keys equal to T in index order: counts per (iteration, wave), one scan
    for (int j = 0; j < iters; ++j) {
        const int e = j * kChunkThreads + tid;
        const bool eq = e < len && key_at(j, e) == T;
        const uint64_t m = __ballot(eq);
        if (lane == 0) eq_cnt[j * kChunkWaves + wv] = (uint32_t)__popcll(m);
    }
    if (tid == 0) s_cnt = 0;
    __syncthreads();
    {
        // exclusive prefix over the (iteration, wave) counts, two entries per thread
        const int n_e = iters * kChunkWaves, i0 = 2 * tid;
        const uint32_t a = i0 < n_e ? eq_cnt[i0] : 0u, b = i0 + 1 < n_e ? eq_cnt[i0 + 1] : 0u;
        uint32_t total;
        const uint32_t before = block_exclusive_scan(a + b, s_wave, &total);
        if (i0 < n_e) eq_cnt[i0] = before;
        if (i0 + 1 < n_e) eq_cnt[i0 + 1] = before + a;
    }
    __syncthreads();
    for (int j = 0; j < iters; ++j) {
        const int e = j * kChunkThreads + tid;
        const uint32_t key = e < len ? key_at(j, e) : 0xFFFFFFFFu;
        const bool lt = e < len && key < T;
        const bool eq = e < len && key == T;
        const uint64_t m = __ballot(eq);
        const uint32_t below = (uint32_t)__popcll(m & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
        const bool take = lt || (eq && eq_cnt[j * kChunkWaves + wv] + below < need);
        if (take) {
            const uint32_t pos = atomicAdd(&s_cnt, 1u);
            if (pos < (uint32_t)kc) out[pos] = ((unsigned long long)key << 32) | (uint32_t)(base + e);
        }
    }
    for (int v = kc + tid; v < k; v += kChunkThreads) out[v] = ~0ull;  // short last chunk
}

__global__ __launch_bounds__(1024) void topk_merge_kernel(const unsigned long long* __restrict__ cand, int m,
                                                          int k, uint32_t* __restrict__ out_idx,
                                                          float* __restrict__ out_val)
{
    __shared__ unsigned long long s[kMergeMax];
    const int q = blockIdx.x, tid = threadIdx.x;
    const unsigned long long* in = cand + (int64_t)q * m;
    int P = 1;
    while (P < m) P <<= 1;
    for (int v = tid; v < P; v += 1024) s[v] = v < m ? in[v] : ~0ull;
    __syncthreads();
    for (int size = 2; size <= P; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int v = tid; v < P; v += 1024) {
                const int w = v ^ stride;
                if (w > v) {
                    const bool up = (v & size) == 0;
                    const unsigned long long x = s[v], y = s[w];
                    if ((x > y) == up) {
                        s[v] = y;
                        s[w] = x;
                    }
                }
            }
            __syncthreads();
        }
    for (int v = tid; v < k; v += 1024) {
        const unsigned long long c = s[v];
        out_idx[(int64_t)q * k + v] = (uint32_t)(c & 0xFFFFFFFFu);
        out_val[(int64_t)q * k + v] = __uint_as_float((uint32_t)(c >> 32));
    }
}

// Chunk length: at least kMinChunk, a multiple of 256, and few enough chunks
// that a row's candidates (chunks x k) fit the merge.
int topk_chunk_len(int n_lib, int k)
{
    const int max_chunks = std::max(1, kMergeMax / std::max(k, 1));
    int c = (n_lib + max_chunks - 1) / max_chunks;
    c = std::max(c, kMinChunk);
    return (c + kChunkThreads - 1) / kChunkThreads * kChunkThreads;
}

}  // namespace

size_t topk_scratch_bytes(int n_q, int n_lib, int k)
{
    if (n_q <= 0 || k <= 0) return 0;
    const int chunk = topk_chunk_len(n_lib, k);
    const int n_chunks = (n_lib + chunk - 1) / chunk;
    return (size_t)n_q * n_chunks * k * 8;
}

bool topk_supported(int n_lib, int k)
{
    if (k <= 0 || k > kTopkMax || k > n_lib) return false;
    const int chunk = topk_chunk_len(n_lib, k);
    const int n_chunks = (n_lib + chunk - 1) / chunk;
    return chunk <= 128 * kChunkThreads && n_chunks * k <= kMergeMax;
}

void launch_topk(const float* scores, int n_q, int n_lib, int k, uint32_t* idx, float* val, void* scratch,
                 hipStream_t s)
{
    if (n_q <= 0) return;
    if (k <= 0 || k > kTopkMax || k > n_lib) fail(BS_ERR_INVALID_ARGUMENT, "top-k: k = %d outside [1, min(%d, %d)]",
                                                  k, kTopkMax, n_lib);
    if (!topk_supported(n_lib, k))
        fail(BS_ERR_UNSUPPORTED, "top-k: %d library members with k = %d exceed the selection's range", n_lib, k);
    const int chunk = topk_chunk_len(n_lib, k);
    const int n_chunks = (n_lib + chunk - 1) / chunk;
    auto* cand = static_cast<unsigned long long*>(scratch);
    topk_chunk_kernel<<<dim3(n_chunks, n_q), kChunkThreads, 0, s>>>(scores, n_lib, k, chunk, cand);
    topk_merge_kernel<<<n_q, 1024, 0, s>>>(cand, n_chunks * k, k, idx, val);
}

}  // namespace bs
