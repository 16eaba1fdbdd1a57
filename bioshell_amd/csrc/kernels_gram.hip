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
#include "gram_common.hpp"

#include <algorithm>

namespace bs {
namespace {


// One 64 x 64 tile of pairs: K loop over atom blocks, then the epilogue.
template <int MODE>
__device__ __forceinline__ void gram_tile(const GramArgs& a, int tr, int tc,
                                          uint4 (&lds)[2][4][kPanelBlockU4])
{
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wr = wv >> 1, wc = wv & 1;

    const uint4* __restrict__ pa = static_cast<const uint4*>(a.panels_a);
    const uint4* __restrict__ pb = static_cast<const uint4*>(a.panels_b);
    const int64_t stride_cb = (int64_t)a.n_kblk * kPanelBlockU4;
    const uint4* srcA0 = pa + (int64_t)(2 * tr) * stride_cb;
    const uint4* srcB0 = pb + (int64_t)(2 * tc) * stride_cb;

    // Six 16-byte loads per thread bring one stage (A0, A1, B0, B1 blocks) in.
    auto src_of = [&](int u) -> const uint4* {
        const int idx = tid + 256 * u;
        const int blk = idx / kPanelBlockU4, off = idx - blk * kPanelBlockU4;
        return (blk < 2 ? srcA0 + blk * stride_cb : srcB0 + (blk - 2) * stride_cb) + off;
    };
    const uint4* s0p = src_of(0);
    const uint4* s1p = src_of(1);
    const uint4* s2p = src_of(2);
    const uint4* s3p = src_of(3);
    const uint4* s4p = src_of(4);
    const uint4* s5p = src_of(5);
    uint4 st0, st1, st2, st3, st4, st5;
#define BS_GLOAD(KB)                                          \
    {                                                         \
        const int64_t o_ = (int64_t)(KB) * kPanelBlockU4;     \
        st0 = s0p[o_]; st1 = s1p[o_]; st2 = s2p[o_];          \
        st3 = s3p[o_]; st4 = s4p[o_]; st5 = s5p[o_];          \
    }
#define BS_LSTORE(ST)                                         \
    {                                                         \
        uint4* d_ = &lds[(ST)][0][0] + tid;                   \
        d_[0] = st0; d_[256] = st1; d_[512] = st2;            \
        d_[768] = st3; d_[1024] = st4; d_[1280] = st5;        \
    }

    f32x16 acc[9];
#pragma unroll
    for (int q = 0; q < 9; ++q)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[q][r] = 0.0f;

    BS_GLOAD(0)
    BS_LSTORE(0)
    __syncthreads();
    for (int kb = 0; kb < a.n_kblk; ++kb) {
        const int cur = kb & 1;
        const bool more = kb + 1 < a.n_kblk;
        if (more) {
            BS_GLOAD(kb + 1)
        }
        const uint4* A = lds[cur][wr];
        const uint4* B = lds[cur][2 + wc];
        mfma_block<MODE>(A, B, lane, acc);
        if (more) {
            BS_LSTORE(cur ^ 1)
        }
        __syncthreads();
    }

#undef BS_GLOAD
#undef BS_LSTORE

    // ---------------------------------------------------------------- epilogue
    // The LDS stage buffers are free now.  Layout (bytes):
    //   [0, 12288)       constants of the tile's 64 rows and 64 columns, 12 doubles each
    //   [12288, 49152)   per wave, 4 accumulator slots x 9 values x 64 lanes (f32)
    // Accumulators go through LDS in 4 rounds of 4 slots so that one copy of
    // the FP64 solve serves all 16 pairs a lane owns.
    double* sconst = reinterpret_cast<double*>(&lds[0][0][0]);
    float* sacc = reinterpret_cast<float*>(reinterpret_cast<char*>(&lds[0][0][0]) + 12288) +
                  wv * (4 * 9 * 64);
    for (int v = tid; v < kNumConsts * 128; v += 256) {
        const int q = v / 128, r = v % 128;
        const bool is_row = r < 64;
        const int conf = is_row ? tr * 64 + r : tc * 64 + (r - 64);
        const double val = is_row ? (conf < a.n_rows ? a.consts_a[(int64_t)q * a.lda + conf] : 0.0)
                                  : (conf < a.n_cols ? a.consts_b[(int64_t)q * a.ldb + conf] : 0.0);
        sconst[r * 12 + q] = val;
    }
    __syncthreads();

    const int jl = wc * 32 + (lane & 31);  // column within the tile
    const int j = tc * 64 + jl;
    const bool jok = j < a.n_cols;
    double cj[kNumConsts];
#pragma unroll
    for (int q = 0; q < kNumConsts; ++q) cj[q] = sconst[(64 + jl) * 12 + q];
    const int il0 = wr * 32 + 4 * (lane >> 5);  // first row within the tile of this lane
    const double Ld = (double)a.L;
    const double invL = 1.0 / Ld;
    const double tau2L = a.tau * a.tau * Ld;

#define BS_STASH(BASE)                                                   \
    _Pragma("unroll") for (int s_ = 0; s_ < 4; ++s_)                     \
        _Pragma("unroll") for (int q_ = 0; q_ < 9; ++q_)                 \
            sacc[(s_ * 9 + q_) * 64 + lane] = acc[q_][(BASE) + s_];

#pragma unroll
    for (int round = 0; round < 4; ++round) {
        BS_STASH(4 * round)
#pragma unroll 1
        for (int sl = 0; sl < 4; sl += 2) {
            // accumulator slot r = 4 round + sl holds row 8 (r/4) + r%4 of the
            // lane's block; slots sl and sl + 1 are solved together
            const int il = il0 + 8 * round + sl;
            const int i0 = tr * 64 + il, i1 = i0 + 1;
            double E0[9], E1[9], c0[kNumConsts], c1[kNumConsts];
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                E0[q] = (double)sacc[(sl * 9 + q) * 64 + lane];
                E1[q] = (double)sacc[((sl + 1) * 9 + q) * 64 + lane];
            }
#pragma unroll
            for (int q = 0; q < kNumConsts; ++q) {
                c0[q] = sconst[il * 12 + q];
                c1[q] = sconst[(il + 1) * 12 + q];
            }
            ShiftedOut o[2];
            qcp_shifted2(E0, c0, E1, c1, cj, o[0], o[1]);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int i = h ? i1 : i0;
                const double* ci = h ? c1 : c0;
                const bool ok = jok && i < a.n_rows && (!a.triangle || i < j);
                bool flag = false;
                if (ok) {
                    const double sgg = ci[10] * cj[10];
                    const double err = a.acc_coef * sgg;
                    // flag: unconverged / near-double root, near-duplicate
                    // residuals, estimated error above tau (compared squared:
                    // no FP64 sqrt), or residuals too large for the split
                    flag = !o[h].converged || (o[h].H < a.kappa_dup * a.u_rep * sgg) ||
                           (err * err > tau2L * o[h].rmsd2_L) ||
                           (2.0 * (ci[9] + cj[9]) > a.rep_g_max);
                    const int64_t oi =
                        a.triangle ? tri_index(a.n_rows, i, j) : (int64_t)i * a.out_ld + j;
                    if (a.out_f64)
                        static_cast<double*>(a.out)[oi] = sqrt(o[h].rmsd2_L * invL);
                    else
                        static_cast<float*>(a.out)[oi] = sqrtf((float)(o[h].rmsd2_L * invL));
                }
                append_flagged(a, flag, i, j, lane);
            }
        }
    }
#undef BS_STASH
}

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
    gram_tile<MODE>(a, (int)(code >> 16), (int)(code & 0xFFFFu), lds);
}

__global__ void zero_u32_kernel(uint32_t* p, int n)
{
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v < n) p[v] = 0u;
}

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// One wave per flagged pair: FP64 covariance from the centred coordinates and
// the textbook QCP (the same arithmetic as the FP64 per-pair kernel).
__global__ __launch_bounds__(256) void fallback_kernel(const double* __restrict__ ca,
                                                       const double* __restrict__ cbp, int L,
                                                       const uint32_t* __restrict__ pairs,
                                                       const uint32_t* __restrict__ count,
                                                       int max_pairs, int triangle, int n,
                                                       int64_t out_ld, void* __restrict__ out,
                                                       int out_f64)
{
    const int lane = threadIdx.x & 63;
    const int cnt = min((int)*count, max_pairs);
    for (int p = blockIdx.x * 4 + (threadIdx.x >> 6); p < cnt; p += gridDim.x * 4) {
        const int i = (int)pairs[2 * (int64_t)p], j = (int)pairs[2 * (int64_t)p + 1];
        const double* x = ca + (int64_t)i * L * 3;
        const double* y = cbp + (int64_t)j * L * 3;
        double m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, ga = 0.0, gb = 0.0;
        for (int k = lane; k < L; k += 64) {
            const double a0 = x[3 * k], a1 = x[3 * k + 1], a2 = x[3 * k + 2];
            const double b0 = y[3 * k], b1 = y[3 * k + 1], b2 = y[3 * k + 2];
            m[0] += a0 * b0; m[1] += a0 * b1; m[2] += a0 * b2;
            m[3] += a1 * b0; m[4] += a1 * b1; m[5] += a1 * b2;
            m[6] += a2 * b0; m[7] += a2 * b1; m[8] += a2 * b2;
            ga += a0 * a0 + a1 * a1 + a2 * a2;
            gb += b0 * b0 + b1 * b1 + b2 * b2;
        }
#pragma unroll
        for (int q = 0; q < 9; ++q) m[q] = wave_sum(m[q]);
        ga = wave_sum(ga);
        gb = wave_sum(gb);
        if (lane == 0) {
            const double r = qcp_rmsd<double>(m, ga, gb, L);
            const int64_t o = triangle ? tri_index(n, i, j) : (int64_t)i * out_ld + j;
            if (out_f64)
                static_cast<double*>(out)[o] = r;
            else
                static_cast<float*>(out)[o] = (float)r;
        }
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

void launch_fallback(const double* centred_a, const double* centred_b, int L,
                     const uint32_t* pairs, const uint32_t* count, int max_pairs, int triangle,
                     int n, int64_t out_ld, void* out, int out_f64, hipStream_t s)
{
    fallback_kernel<<<1024, 256, 0, s>>>(centred_a, centred_b, L, pairs, count, max_pairs,
                                         triangle, n, out_ld, out, out_f64);
}

}  // namespace bs
