// kernels_gram.hip -- covariance as a Gram contraction on MFMA, fused with the
// FP64 shifted-QCP epilogue, plus the FP64 re-evaluation of flagged pairs.
//
// For a block of pairs (i, j) the nine residual cross-covariances
//     E_cd(i, j) = sum_k e_i,k,c * e_j,k,d          (c, d in {x, y, z})
// are nine GEMM tiles that share three row panels (e_x, e_y, e_z of the rows)
// and three column panels.  A wave owns a 32 x 32 block of pairs and keeps the
// nine 32x32 FP32 accumulators of v_mfma_f32_32x32x16_bf16 (or _32x32x2f32);
// lane l then holds column j = l % 32 and rows 8(r/4) + 4(l/32) + r%4 of all
// nine tiles, so the 3x3 matrix of every pair it owns is lane-local and the
// QCP epilogue runs without any exchange.
//
// BF16x3: operands are split e = hi + lo (two bf16) and each tile accumulates
// hi*hi + hi*lo + lo*hi in FP32.  A work-group (4 waves) covers 64 x 64 pairs;
// operand blocks (32 conformations x 16 atoms, 6 KiB) are double-buffered
// through LDS.  Work-group ids are dealt to the 8 XCDs round-robin, so the
// logical tile is taken from a contiguous per-XCD slice of a super-blocked
// order: each XCD's L2 keeps re-using the same row and column panels.
#include "gram_tile.hpp"

#include <algorithm>

namespace bs {
namespace {



// One work-group per 64 x 64 tile.  Work-group ids are dealt to the 8 XCDs
// round-robin, so the logical tile comes from a contiguous per-XCD slice of
// the super-blocked order.  Two work-groups share each CU; the ones in the
// second resident batch (ids [n_cu, 2 n_cu)) start about half a tile late, so
// that on every SIMD one work-group's MFMA phase overlaps the other's FP64
// epilogue (VALU).  Tiles cost the same, so refilled slots keep the offset.
template <int MODE, bool STATS>
__global__ __launch_bounds__(256, 2) void gram_kernel(GramArgs a)
{
    __shared__ uint4 lds[3][4][kPanelBlockU4];  // 3 stages x (A0, A1, B0, B1) x 6 KiB
    const int wg = blockIdx.x;
    const int per = (a.n_tiles + kXcd - 1) / kXcd;
    const int t = (wg % kXcd) * per + (wg / kXcd);
    if (t >= a.n_tiles) return;
    const uint32_t code = a.tiles[t];
    gram_tile64<MODE, true, 2, STATS>(a, (int)(code >> 16), (int)(code & 0xFFFFu), lds, (int)threadIdx.x);
}

__global__ void zero_u32_kernel(uint32_t* p, int n)
{
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v < n) p[v] = 0u;
}

__global__ void zero_words_kernel(uint4* p, int64_t n4)
{
    for (int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x; v < n4; v += (int64_t)gridDim.x * 256)
        p[v] = make_uint4(0u, 0u, 0u, 0u);
}

// Queues the set bits of the flag bitmap: one thread per word; a wave scan of
// the words' bit counts and one atomic per wave give each word its slots.
// Queued bits are cleared; bits that do not fit in the list (cap entries)
// stay set for a bitmap scan.  *count receives the number of set bits.  (The
// batched pipeline's flag path; a single Gram launch queues directly.)
__global__ __launch_bounds__(256) void flag_compact_kernel(uint32_t* __restrict__ bits,
                                                           int64_t n_words,
                                                           int64_t* __restrict__ list, int cap,
                                                           uint32_t* __restrict__ count)
{
    const int lane = threadIdx.x & 63;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t base_w = (int64_t)blockIdx.x * 256 + (threadIdx.x & ~63); base_w < n_words;
         base_w += stride) {
        const int64_t w = base_w + lane;
        uint32_t word = w < n_words ? bits[w] : 0u;
        const int p = __popc(word);
        int incl = p;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(incl, o);
            if (lane >= o) incl += t;
        }
        const int total = __shfl(incl, 63);
        if (total == 0) continue;
        uint32_t start = 0;
        if (lane == 63) start = atomicAdd(count, (uint32_t)total);
        start = __shfl(start, 63);
        int64_t pos = (int64_t)start + (incl - p);
        if (p) {
            uint32_t left = word;
            while (word) {
                const int b = __ffs(word) - 1;
                word &= word - 1u;
                if (pos < cap) {
                    list[pos] = w * 32 + b;
                    left &= ~(1u << b);
                }
                ++pos;
            }
            bits[w] = left;
        }
    }
}

// Decodes output o of a triangle (n x n, i < j) or of a dense n_rows x out_ld block.
__device__ inline void decode_output(int64_t o, int triangle, int n, int64_t out_ld, int& i, int& j)
{
    if (triangle) {
        i = tri_row(o, n);
        j = (int)(o - ((int64_t)i * n - (int64_t)i * (i + 1) / 2)) + i + 1;
    } else {
        i = (int)(o / out_ld);
        j = (int)(o - (int64_t)i * out_ld);
    }
}

// FP64 re-evaluation of one flagged output o (lane 0 stores the result and,
// with st_min set, folds the stored f32 value into both elements' keys).
__device__ inline void resolve_output(int64_t o, const double* __restrict__ xa, const double* __restrict__ ca,
                                      const double* __restrict__ xb, const double* __restrict__ cb, int L,
                                      int triangle, int n, int64_t out_ld, void* __restrict__ out, int out_f64,
                                      unsigned* st_min, unsigned* st_max, int lane)
{
    int i, j;
    decode_output(o, triangle, n, out_ld, i, j);
    const double r = fallback_pair(xa + (int64_t)i * L * 3, ca + 3 * (int64_t)i, xb + (int64_t)j * L * 3,
                                   cb + 3 * (int64_t)j, L, lane);
    if (lane == 0) {
        if (out_f64) {
            static_cast<double*>(out)[o] = r;
        } else {
            const float v = (float)r;
            static_cast<float*>(out)[o] = v;
            if (st_min && v == v) {
                const unsigned k = fkey(__float_as_uint(v));
                atomicMin(st_min + i, k);
                atomicMin(st_min + j, k);
                atomicMax(st_max + i, k);
                atomicMax(st_max + j, k);
            }
        }
    }
}

// The flagged outputs of one Gram launch: one wave per queued entry, then,
// only when more were flagged than the list holds, a scan of the bitmap in
// which each wave takes 64 words (one per lane), processes their set bits one
// pair at a time (the 64 lanes splitting the atoms) and clears the words.
// ctl = {count, report, done}: the last work-group to finish publishes the
// count in report and zeroes count and done for the next launch.
__global__ __launch_bounds__(256) void flag_resolve_kernel(const double* __restrict__ xa,
                                                           const double* __restrict__ ca,
                                                           const double* __restrict__ xb,
                                                           const double* __restrict__ cb, int L,
                                                           uint32_t* __restrict__ ctl,
                                                           const int64_t* __restrict__ list, int cap,
                                                           uint32_t* __restrict__ bits, int64_t n_out,
                                                           int triangle, int n, int64_t out_ld,
                                                           void* __restrict__ out, int out_f64,
                                                           uint32_t* h_report_a, uint32_t* h_report_b,
                                                           unsigned* st_min, unsigned* st_max)
{
    const int lane = threadIdx.x & 63;
    const uint32_t cnt = ctl[0];  // read before this group counts itself done
    const int64_t wave = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int64_t n_waves = (int64_t)gridDim.x * 4;
    const int64_t m = min((int64_t)cnt, (int64_t)cap);
    for (int64_t k = wave; k < m; k += n_waves) {
        const int64_t o = list[k];
        if (o >= 0 && o < n_out)
            resolve_output(o, xa, ca, xb, cb, L, triangle, n, out_ld, out, out_f64, st_min, st_max, lane);
    }
    if ((int64_t)cnt > (int64_t)cap) {  // uniform: the bitmap holds the rest
        const int64_t n_words = (n_out + 31) >> 5;
        for (int64_t w0 = wave * 64; w0 < n_words; w0 += n_waves * 64) {
            const int64_t wi = w0 + lane;
            uint32_t word = wi < n_words ? bits[wi] : 0u;
            if (word) bits[wi] = 0u;
            while (true) {
                const uint64_t msk = __ballot(word != 0u);
                if (!msk) break;
                const int src = __ffsll((unsigned long long)msk) - 1;
                const uint32_t wsrc = __shfl(word, src);
                const int bit = __ffs(wsrc) - 1;
                const int64_t o = (w0 + src) * 32 + bit;
                if (lane == src) word &= word - 1u;  // clear the lowest set bit
                if (o >= n_out) continue;
                resolve_output(o, xa, ca, xb, cb, L, triangle, n, out_ld, out, out_f64, st_min, st_max, lane);
            }
        }
    }
    // Every thread of the group has used its copy of the count (no fence:
    // the reset only needs the reads done, and the next launch sees it).
    __syncthreads();
    if (threadIdx.x == 0 && atomicAdd(&ctl[2], 1u) == gridDim.x - 1) {
        ctl[1] = cnt;
        ctl[0] = 0u;
        ctl[2] = 0u;
        if (h_report_a) __hip_atomic_store(h_report_a, cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        if (h_report_b) __hip_atomic_store(h_report_b, cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

}  // namespace

int build_tile_order(int tiles_r, int tiles_c, bool triangle, int super, uint32_t* out)
{
    int n = 0;
    const int sr_n = (tiles_r + super - 1) / super, sc_n = (tiles_c + super - 1) / super;
    for (int sr = 0; sr < sr_n; ++sr)
        for (int sc = triangle ? sr : 0; sc < sc_n; ++sc)
            for (int r = sr * super; r < min((sr + 1) * super, tiles_r); ++r)
                for (int c = sc * super; c < min((sc + 1) * super, tiles_c); ++c)
                    if (!triangle || r <= c) out[n++] = ((uint32_t)r << 16) | (uint32_t)c;
    return n;
}

void launch_gram(const GramArgs& a, GramMode mode, hipStream_t s)
{
    if (a.n_tiles <= 0) return;
    const int per = (a.n_tiles + kXcd - 1) / kXcd;
    const int grid = per * kXcd;
    const bool st = a.stats_part != nullptr;
    if (mode == kGramBF16x3 && st)
        gram_kernel<kGramBF16x3, true><<<grid, 256, 0, s>>>(a);
    else if (mode == kGramBF16x3)
        gram_kernel<kGramBF16x3, false><<<grid, 256, 0, s>>>(a);
    else if (st)
        gram_kernel<kGramFP32, true><<<grid, 256, 0, s>>>(a);
    else
        gram_kernel<kGramFP32, false><<<grid, 256, 0, s>>>(a);
}

void launch_zero_u32(uint32_t* p, int n, hipStream_t s)
{
    zero_u32_kernel<<<(n + 255) / 256, 256, 0, s>>>(p, n);
}

void launch_zero_words(uint32_t* p, int64_t n_words, hipStream_t s)
{
    // callers allocate bitmaps in whole 16-byte units
    const int64_t n4 = (n_words + 3) / 4;
    const int64_t blocks = std::min<int64_t>(4096, std::max<int64_t>(1, (n4 + 255) / 256));
    zero_words_kernel<<<(int)blocks, 256, 0, s>>>(reinterpret_cast<uint4*>(p), n4);
}

void launch_flag_resolve(const double* xyz_a, const double* cent_a, const double* xyz_b,
                         const double* cent_b, int L, uint32_t* ctl, const int64_t* list, int cap,
                         uint32_t* bits, int64_t n_out, int triangle, int n, int64_t out_ld, void* out,
                         int out_f64, uint32_t* h_report_a, uint32_t* h_report_b, unsigned* st_min,
                         unsigned* st_max, hipStream_t s)
{
    // 1024 waves: a typical list (tens to thousands of pairs) is one pass, a
    // full list or a bitmap scan takes several.  A grid sized for the bitmap
    // (up to 1024 groups) cost 9 us more per call in dispatch and in the
    // done-counter atomics of the idle groups.
    flag_resolve_kernel<<<256, 256, 0, s>>>(
        xyz_a, cent_a, xyz_b, cent_b, L, ctl, list, cap, bits, n_out, triangle, n, out_ld, out, out_f64,
        h_report_a, h_report_b, st_min, st_max);
}

void launch_flag_compact(uint32_t* bits, int64_t n_words, int64_t* list, int cap, uint32_t* count,
                         hipStream_t s)
{
    const int64_t blocks = std::min<int64_t>(4096, std::max<int64_t>(1, (n_words + 255) / 256));
    flag_compact_kernel<<<(int)blocks, 256, 0, s>>>(bits, n_words, list, cap, count);
}

}  // namespace bs
