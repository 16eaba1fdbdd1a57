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
template <int MODE>
__global__ __launch_bounds__(256, 2) void gram_kernel(GramArgs a)
{
    __shared__ uint4 lds[2][4][kPanelBlockU4];  // 2 stages x (A0, A1, B0, B1) x 6 KiB
    const int wg = blockIdx.x;
    const int per = (a.n_tiles + kXcd - 1) / kXcd;
    const int t = (wg % kXcd) * per + (wg / kXcd);
    if (t >= a.n_tiles) return;
    if (a.skew_sleeps > 0 && wg >= a.max_wg / 2 && wg < a.max_wg)
        for (int k = 0; k < a.skew_sleeps; ++k) __builtin_amdgcn_s_sleep(127);
    const uint32_t code = a.tiles[t];
    gram_tile64<MODE>(a, (int)(code >> 16), (int)(code & 0xFFFFu), lds, (int)threadIdx.x);
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

// FP64 re-evaluation of flagged outputs: each wave takes 64 bitmap words (one
// per lane) and processes their set bits one pair at a time, the 64 lanes
// splitting the atoms; textbook QCP with the near-degenerate polish.
__global__ __launch_bounds__(256) void fallback_kernel(const double* __restrict__ xa,
                                                       const double* __restrict__ ca,
                                                       const double* __restrict__ xb,
                                                       const double* __restrict__ cb, int L,
                                                       const uint32_t* __restrict__ bits,
                                                       int64_t n_out, int triangle, int n,
                                                       int64_t out_ld, void* __restrict__ out,
                                                       int out_f64, uint32_t* __restrict__ count)
{
    const int lane = threadIdx.x & 63;
    const int64_t n_words = (n_out + 31) >> 5;
    const int64_t wave = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int64_t n_waves = (int64_t)gridDim.x * 4;
    for (int64_t w0 = wave * 64; w0 < n_words; w0 += n_waves * 64) {
        const int64_t wi = w0 + lane;
        uint32_t word = wi < n_words ? bits[wi] : 0u;
        uint32_t done = 0;
        while (true) {
            const uint64_t m = __ballot(word != 0u);
            if (!m) break;
            const int src = __ffsll((unsigned long long)m) - 1;
            const uint32_t wsrc = __shfl(word, src);
            const int bit = __ffs(wsrc) - 1;
            const int64_t o = (w0 + src) * 32 + bit;
            if (lane == src) word &= word - 1u;  // clear the lowest set bit
            if (o >= n_out) continue;
            int i, j;
            if (triangle) {
                i = tri_row(o, n);
                j = (int)(o - ((int64_t)i * n - (int64_t)i * (i + 1) / 2)) + i + 1;
            } else {
                i = (int)(o / out_ld);
                j = (int)(o - (int64_t)i * out_ld);
            }
            const double r = fallback_pair(xa + (int64_t)i * L * 3, ca + 3 * (int64_t)i,
                                           xb + (int64_t)j * L * 3, cb + 3 * (int64_t)j, L, lane);
            if (lane == 0) {
                if (out_f64)
                    static_cast<double*>(out)[o] = r;
                else
                    static_cast<float*>(out)[o] = (float)r;
            }
            ++done;
        }
        if (count && lane == 0 && done) atomicAdd(count, done);
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
    if (mode == kGramBF16x3)
        gram_kernel<kGramBF16x3><<<grid, 256, 0, s>>>(a);
    else
        gram_kernel<kGramFP32><<<grid, 256, 0, s>>>(a);
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

void launch_fallback(const double* xyz_a, const double* cent_a, const double* xyz_b,
                     const double* cent_b, int L, const uint32_t* bits, int64_t n_out,
                     int triangle, int n, int64_t out_ld, void* out, int out_f64, uint32_t* count,
                     hipStream_t s)
{
    const int64_t n_words = (n_out + 31) / 32;
    const int64_t blocks = std::min<int64_t>(2048, std::max<int64_t>(1, (n_words + 255) / 256));
    fallback_kernel<<<(int)blocks, 256, 0, s>>>(xyz_a, cent_a, xyz_b, cent_b, L, bits, n_out,
                                                triangle, n, out_ld, out, out_f64, count);
}

}  // namespace bs
