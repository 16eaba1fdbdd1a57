// kernels_batch.hip -- many ensembles of uneven size (n conformations, L
// atoms each) in a fixed sequence of launches, independent of how many
// ensembles there are: every preparation step is one launch over all the
// batch's conformations (or reference samples, mean slices, coordinates),
// the Gram step is one persistent launch over a work list of (ensemble, tile)
// items, and the FP64 re-evaluation is one launch over all flag bitmaps.  The host captures the
// sequence in a HIP graph.
//
// Batch-wide index -> (ensemble, local index) goes through the prefix sums of
// BatchArgs by binary search (ensembles are a few hundred; the search is a
// handful of scalar loads per wave).
#include "gram_tile.hpp"
#include "prep_device.hpp"

#include <algorithm>

namespace bs {
namespace {

// Ensemble e with prefix[e] <= g < prefix[e + 1] (empty ranges are skipped).
template <typename T>
__device__ inline int find_ens(const T* __restrict__ prefix, int n_ens, T g)
{
    int lo = 0, hi = n_ens;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= g)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// ------------------------------------------------------------------ prep
// One work-group per slice of the reference mean (fit0_slice).
__global__ __launch_bounds__(256) void batch_fit0_kernel(BatchArgs b)
{
    __shared__ Fit0Lds lds;
    const int g = blockIdx.x;
    const int e = find_ens(b.slice_prefix, b.n_ens, g);
    const BatchEns& E = b.ens[e];
    fit0_slice(E.xyz, E.n, E.n_ref, E.L, g - b.slice_prefix[e], E.partial, lds);
}

// One thread per reference coordinate, as in the single-ensemble kernel.
__global__ __launch_bounds__(256) void batch_mean_final_kernel(BatchArgs b)
{
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= b.total_coords) return;
    const int e = find_ens(b.coord_prefix, b.n_ens, g);
    const BatchEns& E = b.ens[e];
    mean_final_coord(E.partial, E.n_slices, E.n_ref, E.L, g - b.coord_prefix[e], E.ref);
}

__global__ __launch_bounds__(256) void batch_ref_gram_kernel(BatchArgs b)
{
    __shared__ double part[4][6];
    const BatchEns& E = b.ens[blockIdx.x];
    ref_gram_block(E.ref, E.L, part);
}

template <int MODE>
__global__ __launch_bounds__(256) void batch_fit_pack_kernel(BatchArgs b)
{
    __shared__ FitPackLds lds;
    const int g = xcd_contiguous_block(b.total_fit);
    if (g < 0) return;
    const int e = find_ens(b.fit_prefix, b.n_ens, g);
    const BatchEns& E = b.ens[e];
    fit_pack_block<MODE>(E.xyz, E.ref, E.n, E.L, E.n_kblk, (g - b.fit_prefix[e]) * kFitConfs, E.rot,
                         E.cent, static_cast<uint4*>(E.panels), E.consts, E.n_pad, lds);
}

__global__ void batch_zero_kernel(uint4* p, int64_t n4)
{
    for (int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x; v < n4; v += (int64_t)gridDim.x * 256)
        p[v] = make_uint4(0u, 0u, 0u, 0u);
}

// ------------------------------------------------------------------ Gram
__device__ inline uint64_t uniform64(uint64_t v)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// The view of ensemble E that the tile code takes.
__device__ inline GramArgs ens_args(const BatchArgs& b, const BatchEns& E)
{
    GramArgs a{};
    a.panels_a = E.panels;
    a.panels_b = E.panels;
    a.consts_a = E.consts;
    a.consts_b = E.consts;
    a.lda = E.n_pad;
    a.ldb = E.n_pad;
    a.n_rows = E.n;
    a.n_cols = E.n;
    a.L = E.L;
    a.n_kblk = E.n_kblk;
    a.triangle = 1;
    a.out_ld = 0;
    a.out = E.out;
    a.out_f64 = 0;
    a.flag_bits = E.bits;
    a.lim = E.lim;
    return a;
}

constexpr int kBatchLdsBytes = 3 * 4 * kPanelBlockU4 * 16;  // 72 KiB: the 64-tile stages
static_assert(kBatchLdsBytes >= 4 * epilogue_lds_bytes<32>(), "four 32-tile waves must fit");
static_assert(kBatchLdsBytes >= epilogue_lds_bytes<64>(), "64-tile epilogue must fit");

// Persistent: n_wg work-groups (2 per CU), each pulls items from its XCD's
// list, then from the others in turn, until every list is drained.  Every
// work-group leaves when its last probe of list x + 7 comes back empty.
template <int MODE>
__global__ __launch_bounds__(256, 2) void batch_gram_kernel(BatchArgs b)
{
    __shared__ __align__(16) char lds[kBatchLdsBytes];
    __shared__ uint32_t s_item;
    const int xcd = blockIdx.x % kXcd;
    int probe = 0;  // lists xcd, xcd + 1, ... (mod 8) in turn
    while (true) {
        if (threadIdx.x == 0) {
            uint32_t item = 0xFFFFFFFFu;
            while (probe < kXcd) {
                const int x = (xcd + probe) % kXcd;
                const int len = b.list_start[x + 1] - b.list_start[x];
                if (len > 0) {
                    const uint32_t k = atomicAdd(b.counters + x, 1u);
                    if (k < (uint32_t)len) {
                        item = b.items[b.list_start[x] + (int)k];
                        break;
                    }
                }
                ++probe;
            }
            s_item = item;
        }
        __syncthreads();
        // uniform values in scalar registers: the ensemble's fields are then
        // scalar loads, not per-lane copies
        const uint32_t item = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();  // all have read s_item before thread 0 writes the next one
        if (item == 0xFFFFFFFFu) break;
        // an opaque copy of the thread index per item, so the compiler cannot
        // hoist thread-derived addresses out of the loop and keep them live
        // across items (that costs ~50 VGPRs and spills)
        int thread = threadIdx.x;
        asm volatile("" : "+v"(thread));
        if (!(item & 0x80000000u)) {
            const uint64_t code = uniform64(b.tiles64[item]);
            const BatchEns& E = b.ens[(int)(code >> 32)];
            const GramArgs a = ens_args(b, E);
            gram_tile64<MODE, false, 1>(a, (int)((code >> 16) & 0xFFFFu), (int)(code & 0xFFFFu),
                              *reinterpret_cast<uint4(*)[3][4][kPanelBlockU4]>(lds), thread);
            __syncthreads();  // the next item reuses the LDS
        } else {
            const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
            const uint64_t code = uniform64(b.tiles32[4 * (int64_t)(item & 0x7FFFFFFFu) + w]);
            if (code != ~0ull) {
                const BatchEns& E = b.ens[(int)(code >> 32)];
                const GramArgs a = ens_args(b, E);
                gram_tile32<MODE>(a, (int)((code >> 16) & 0xFFFFu), (int)(code & 0xFFFFu),
                                  lds + w * epilogue_lds_bytes<32>(), thread);
            }
        }
    }
}

// ------------------------------------------------------------------ FP64 re-evaluation
// Flagged pair of ensemble E at output o: one wave, FP64 textbook QCP.
__device__ inline void batch_eval_one(const BatchEns& E, int64_t o, int lane)
{
    const int i = tri_row(o, E.n);
    const int j = (int)(o - ((int64_t)i * E.n - (int64_t)i * (i + 1) / 2)) + i + 1;
    const double r = fallback_pair(E.xyz + (int64_t)i * E.L * 3, E.cent + 3 * (int64_t)i,
                                   E.xyz + (int64_t)j * E.L * 3, E.cent + 3 * (int64_t)j, E.L,
                                   lane);
    if (lane == 0) E.out[o] = (float)r;
}

// The queued entries (batch-wide bit indices), one wave each.
__global__ __launch_bounds__(256) void batch_flag_eval_kernel(BatchArgs b)
{
    const int lane = threadIdx.x & 63;
    const int64_t m = min((int64_t)b.counters[kXcd], (int64_t)kFlagListCap);
    for (int64_t k = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); k < m; k += (int64_t)gridDim.x * 4) {
        const int64_t gb = b.flag_list[k];
        const int64_t gw = gb >> 5;
        const int e = find_ens(b.word_prefix, b.n_ens, gw);
        const BatchEns& E = b.ens[e];
        const int64_t o = (gw - b.word_prefix[e]) * 32 + (gb & 31);
        if (o < (int64_t)E.n * (E.n - 1) / 2) batch_eval_one(E, o, lane);
    }
}

// Bits the queue could not hold: each wave takes 64 consecutive words of the
// batch-wide bitmap (one per lane) and re-evaluates their set bits one pair
// at a time.
__global__ __launch_bounds__(256) void batch_fallback_kernel(BatchArgs b)
{
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int64_t n_waves = (int64_t)gridDim.x * 4;
    const uint32_t* bits0 = b.ens_bits0;
    for (int64_t w0 = wave * 64; w0 < b.total_words; w0 += n_waves * 64) {
        const int64_t wi = w0 + lane;
        uint32_t word = wi < b.total_words ? bits0[wi] : 0u;
        while (true) {
            const uint64_t m = __ballot(word != 0u);
            if (!m) break;
            const int src = __ffsll((unsigned long long)m) - 1;
            const uint32_t wsrc = __shfl(word, src);
            const int bit = __ffs(wsrc) - 1;
            if (lane == src) word &= word - 1u;  // clear the lowest set bit
            const int64_t gw = w0 + src;
            const int e = find_ens(b.word_prefix, b.n_ens, gw);
            const BatchEns& E = b.ens[e];
            const int64_t o = (gw - b.word_prefix[e]) * 32 + bit;
            if (o >= (int64_t)E.n * (E.n - 1) / 2) continue;
            batch_eval_one(E, o, lane);
        }
    }
}

}  // namespace

void launch_batch_prep(const BatchArgs& b, hipStream_t s)
{
    batch_fit0_kernel<<<b.total_slices, 256, 0, s>>>(b);
    batch_mean_final_kernel<<<(b.total_coords + 255) / 256, 256, 0, s>>>(b);
    batch_ref_gram_kernel<<<b.n_ens, 256, 0, s>>>(b);
    const int pw = xcd_contiguous_grid(b.total_fit);
    if (b.mode == kGramBF16x3)
        batch_fit_pack_kernel<kGramBF16x3><<<pw, 256, 0, s>>>(b);
    else
        batch_fit_pack_kernel<kGramFP32><<<pw, 256, 0, s>>>(b);
}

void launch_batch_gram(const BatchArgs& b, uint32_t* zero_base, int64_t zero_words, hipStream_t s)
{
    // bitmaps and list counters are one contiguous, 16-byte-aligned region
    const int64_t n4 = (zero_words + 3) / 4;
    const int64_t zb = std::min<int64_t>(2048, std::max<int64_t>(1, (n4 + 255) / 256));
    batch_zero_kernel<<<(int)zb, 256, 0, s>>>(reinterpret_cast<uint4*>(zero_base), n4);
    if (b.mode == kGramBF16x3)
        batch_gram_kernel<kGramBF16x3><<<b.n_wg, 256, 0, s>>>(b);
    else
        batch_gram_kernel<kGramFP32><<<b.n_wg, 256, 0, s>>>(b);
}

void launch_batch_fallback(const BatchArgs& b, hipStream_t s)
{
    launch_flag_compact(b.ens_bits0, b.total_words, b.flag_list, kFlagListCap, b.counters + kXcd, s);
    batch_flag_eval_kernel<<<1024, 256, 0, s>>>(b);
    const int64_t waves = (b.total_words + 63) / 64;
    const int64_t blocks = std::min<int64_t>(2048, std::max<int64_t>(1, (waves + 3) / 4));
    batch_fallback_kernel<<<(int)blocks, 256, 0, s>>>(b);
}

}  // namespace bs
