// kernels_gram_pipe.hip -- the Gram kernel with the FP64 epilogue of one tile
// software-pipelined under the MFMA K loop of the next.
//
// Counters on the non-pipelined kernel show MFMA busy cycles equal to the
// issued-MFMA ideal but only ~37 % of the kernel time: its MFMA phase and its
// FP64 epilogue (VALU) run back to back.  Here each work-group is persistent
// (one per CU, four waves, one per SIMD with the whole 512-register file) and
// walks a sequence of 64 x 64 tiles.  While a wave accumulates tile t on the
// matrix cores, the same straight-line blocks carry the epilogue of tile t-1
// on the VALU: every K-loop step issues one atom block of MFMAs plus the QCP
// solve of one of the previous tile's 16 accumulator slots.
//
// The previous tile's accumulators are kept half in LDS (slots 0..7, then
// 8..15 once the first half is consumed) and half in registers; its row and
// column constants sit in a second LDS buffer.  The epilogue slot is branch
// free: one Newton step and the verification that the root reached is the
// largest one and converged (qcp.hpp); a pair that fails it is flagged for
// the FP64 re-evaluation like any other flagged pair.
#include "gram_common.hpp"

#include <algorithm>

namespace bs {
namespace {

constexpr int kStash = 8;  // accumulator slots per wave held in LDS

template <int MODE, int OUT64>
__global__ __launch_bounds__(256, 1) void gram_pipe_kernel(GramArgs a)
{
    __shared__ uint4 lds[2][4][kPanelBlockU4];    // 48 KiB: operand stages
    __shared__ double sconst[2][128 * 12];         // 24 KiB: constants of two tiles
    __shared__ float sacc[4][kStash * 9 * 64];     // 72 KiB: previous tile's slots

    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wr = wv >> 1, wc = wv & 1;
    const int slots = gridDim.x / kXcd, x = blockIdx.x % kXcd, slot = blockIdx.x / kXcd;
    const int per = (a.n_tiles + kXcd - 1) / kXcd;
    const int t_end = min(a.n_tiles, (x + 1) * per);
    int t = x * per + slot;

    const uint4* __restrict__ pa = static_cast<const uint4*>(a.panels_a);
    const uint4* __restrict__ pb = static_cast<const uint4*>(a.panels_b);
    const int nk = a.n_kblk;
    const int64_t stride_cb = (int64_t)nk * kPanelBlockU4;
    float* my = sacc[wv];
    const int il0 = wr * 32 + 4 * (lane >> 5);  // first row (within a tile) of this lane
    const int jl = wc * 32 + (lane & 31);       // column within a tile
    const double Ld = (double)a.L;
    const double invL = 1.0 / Ld;
    const double tau2L = a.tau * a.tau * Ld;
    const double dup_coef = a.kappa_dup * a.u_rep;

    f32x16 acc[9];
    float keep[9][8];           // previous tile, slots 8..15
    double cjp[kNumConsts];     // previous tile, this lane's column constants
#pragma unroll
    for (int q = 0; q < kNumConsts; ++q) cjp[q] = 0.0;
#pragma unroll
    for (int q = 0; q < 9; ++q)
#pragma unroll
        for (int r = 0; r < 8; ++r) keep[q][r] = 0.0f;
    int ptr = 0, ptc = 0, pbuf = 0;
    bool have_prev = false;

    // per-thread offsets of the six 16-byte words that make up one stage
    int blk_of[6], off_of[6];
#pragma unroll
    for (int u = 0; u < 6; ++u) {
        const int idx = tid + 256 * u;
        blk_of[u] = idx / kPanelBlockU4;
        off_of[u] = idx - blk_of[u] * kPanelBlockU4;
    }

    while (true) {
        const bool have_cur = t < t_end;
        if (!have_cur && !have_prev) break;
        int tr = 0, tc = 0;
        if (have_cur) {
            const uint32_t code = a.tiles[t];
            tr = (int)(code >> 16);
            tc = (int)(code & 0xFFFFu);
        }
        const int cbuf = pbuf ^ 1;
        if (have_cur) {
            for (int v = tid; v < kNumConsts * 128; v += 256) {
                const int q = v / 128, r = v % 128;
                const bool is_row = r < 64;
                const int conf = is_row ? tr * 64 + r : tc * 64 + (r - 64);
                const double val =
                    is_row ? (conf < a.n_rows ? a.consts_a[(int64_t)q * a.lda + conf] : 0.0)
                           : (conf < a.n_cols ? a.consts_b[(int64_t)q * a.ldb + conf] : 0.0);
                sconst[cbuf][r * 12 + q] = val;
            }
        }
#pragma unroll
        for (int q = 0; q < 9; ++q)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[q][r] = 0.0f;

        const uint4* sp[6];
#pragma unroll
        for (int u = 0; u < 6; ++u)
            sp[u] = (blk_of[u] < 2 ? pa + (int64_t)(2 * tr + blk_of[u]) * stride_cb
                                   : pb + (int64_t)(2 * tc + blk_of[u] - 2) * stride_cb) +
                    off_of[u];
        uint4 st[6];
        const int nkc = have_cur ? nk : 0;
        if (have_cur) {
#pragma unroll
            for (int u = 0; u < 6; ++u) st[u] = sp[u][0];
#pragma unroll
            for (int u = 0; u < 6; ++u) (&lds[0][0][0])[tid + 256 * u] = st[u];
        }
        __syncthreads();

        // ---- previous tile's epilogue context
        const int pj = ptc * 64 + jl;
        const bool pjok = pj < a.n_cols;
        const double* pconst = sconst[pbuf];

        // One K step: prefetch the next atom block (the last one is re-read at
        // the end, which keeps the step branch free), MFMAs, stage store, barrier.
#define BS_KSTEP(KB)                                                                        \
    {                                                                                       \
        const int kb_ = (KB);                                                               \
        const int64_t o_ = (int64_t)min(kb_ + 1, nk - 1) * kPanelBlockU4;                   \
        _Pragma("unroll") for (int u = 0; u < 6; ++u) st[u] = sp[u][o_];                    \
        mfma_block<MODE>(lds[kb_ & 1][wr], lds[kb_ & 1][2 + wc], lane, acc);                \
        _Pragma("unroll") for (int u = 0; u < 6; ++u)(&lds[(kb_ & 1) ^ 1][0][0])[tid + 256 * u] = \
            st[u];                                                                          \
        __syncthreads();                                                                    \
    }

        // One epilogue slot of the previous tile (straight-line FP64).
#define BS_ESLOT(S)                                                                          \
    {                                                                                        \
        const int s_ = (S);                                                                  \
        const int pos_ = s_ & (kStash - 1);                                                  \
        double E[9], ci[kNumConsts];                                                         \
        _Pragma("unroll") for (int q = 0; q < 9; ++q) E[q] = (double)my[(pos_ * 9 + q) * 64 + lane]; \
        const int il = il0 + 8 * (s_ >> 2) + (s_ & 3);                                       \
        _Pragma("unroll") for (int q = 0; q < kNumConsts; ++q) ci[q] = pconst[il * 12 + q];  \
        const ShiftedPoly sp_ = shifted_poly(E, ci, cjp);                                    \
        ShiftedIter st_{shifted_start(sp_), 0.0, 0.0, 0, false, false, false, false};        \
        shifted_newton(sp_, st_);                                                            \
        shifted_verify(sp_, st_);                                                            \
        const ShiftedOut so_ = shifted_finish(sp_, st_);                                     \
        const bool good = so_.converged;                                                     \
        const double r2 = so_.rmsd2_L;                                                       \
        const int i_ = ptr * 64 + il;                                                        \
        const bool ok = pjok & (i_ < a.n_rows) & ((a.triangle == 0) | (i_ < pj));            \
        const double sgg = ci[10] * cjp[10];                                                 \
        const double err = a.acc_coef * sgg;                                                 \
        const bool flag = ok & (!good | (sp_.H < dup_coef * sgg) | (err * err > tau2L * r2) | \
                                (2.0 * (ci[9] + cjp[9]) > a.rep_g_max));                     \
        if (ok) {                                                                            \
            const int64_t oi =                                                               \
                a.triangle ? tri_index(a.n_rows, i_, pj) : (int64_t)i_ * a.out_ld + pj;      \
            if constexpr (OUT64)                                                             \
                static_cast<double*>(a.out)[oi] = sqrt(r2 * invL);                           \
            else                                                                             \
                static_cast<float*>(a.out)[oi] = sqrtf((float)(r2 * invL));                  \
            flag_output(a, flag, oi);                                                        \
        }                                                                                    \
    }

        const int nmix = have_prev ? min(nkc, 16) : 0;
        int kb = 0, s = 0;
        // slots 0..7 (from the LDS stash), each under one K step while steps remain
        const int mix_a = min(nmix, 8);
        for (; s < mix_a; ++s, ++kb) {
            BS_KSTEP(kb)
            BS_ESLOT(s)
        }
        if (have_prev) {
            for (; s < 8; ++s) BS_ESLOT(s)
            // the first half is consumed: bring slots 8..15 into the stash
#pragma unroll
            for (int r = 0; r < 8; ++r)
#pragma unroll
                for (int q = 0; q < 9; ++q) my[(r * 9 + q) * 64 + lane] = keep[q][r];
        }
        const int mix_b = nmix > 8 ? nmix - 8 : 0;
        for (int m = 0; m < mix_b; ++m, ++s, ++kb) {
            BS_KSTEP(kb)
            BS_ESLOT(s)
        }
        if (have_prev)
            for (; s < 16; ++s) BS_ESLOT(s)
        for (; kb < nkc; ++kb) BS_KSTEP(kb)
#undef BS_KSTEP
#undef BS_ESLOT

        // the current tile becomes the previous one
        if (have_cur) {
#pragma unroll
            for (int r = 0; r < 8; ++r)
#pragma unroll
                for (int q = 0; q < 9; ++q) my[(r * 9 + q) * 64 + lane] = acc[q][r];
#pragma unroll
            for (int q = 0; q < 9; ++q)
#pragma unroll
                for (int r = 0; r < 8; ++r) keep[q][r] = acc[q][8 + r];
#pragma unroll
            for (int q = 0; q < kNumConsts; ++q) cjp[q] = sconst[cbuf][(64 + jl) * 12 + q];
        }
        ptr = tr;
        ptc = tc;
        pbuf = cbuf;
        have_prev = have_cur;
        t += slots;
        __syncthreads();
    }
}

}  // namespace

void launch_gram_pipe(const GramArgs& a, GramMode mode, hipStream_t s)
{
    if (a.n_tiles <= 0) return;
    const int per = (a.n_tiles + kXcd - 1) / kXcd;
    const int slots = std::max(1, std::min(per, (a.max_wg / 2) / kXcd));
    const int grid = slots * kXcd;
    if (mode == kGramBF16x3) {
        if (a.out_f64)
            gram_pipe_kernel<kGramBF16x3, 1><<<grid, 256, 0, s>>>(a);
        else
            gram_pipe_kernel<kGramBF16x3, 0><<<grid, 256, 0, s>>>(a);
    } else {
        if (a.out_f64)
            gram_pipe_kernel<kGramFP32, 1><<<grid, 256, 0, s>>>(a);
        else
            gram_pipe_kernel<kGramFP32, 0><<<grid, 256, 0, s>>>(a);
    }
}

}  // namespace bs
