// gram_tile.hpp -- one tile of the Gram kernels: K loop and FP64 epilogue.
#pragma once

#include "gram_common.hpp"

namespace bs {

// LDS bytes the epilogue of a TILE x TILE tile needs: the constants of the
// tile's rows and columns, 12 doubles each.
template <int TILE>
constexpr int epilogue_lds_bytes()
{
    return 2 * TILE * 12 * 8;
}

// LDS bytes of the epilogue's row-statistics keys (64 x 64 tile, rows padded to 65).
constexpr int kStatsKeysBytes = 64 * 65 * 4;

// The FP64 QCP epilogue of a TILE x TILE tile whose accumulators the calling
// waves hold: TILE 64 is four waves of a work-group (wave w owns rows
// 32 (w/2) .. +31 and columns 32 (w%2) .. +31), TILE 32 is one wave on its own.
// The rows' and columns' constants go to LDS at `lds` (which the caller no
// longer needs); each lane then solves its 16 pairs straight from its
// accumulator registers, P at a time.  Accumulator slot r of a lane holds row
// 8 (r/4) + 4 (lane/32) + r%4 of the wave's block and column lane % 32.
//
// The solve is Newton on the secular form of the QCP polynomial from mu = 0
// (qcp.hpp), with a second step in the waves where a bracket is still wide;
// a pair whose S = -B is not positive definite or whose bracket stays wide
// is flagged, like every pair the error model cannot vouch for, and is
// re-evaluated in FP64 from the coordinates.
//
// STATS (64 x 64 tiles of a triangle, f32 output): every pair also leaves
// the order-preserving key of its stored value in an LDS tile (all-ones for a
// pair outside the triangle or flagged: the resolve kernel adds the flagged
// pairs' final values), and the work-group folds the tile's rows and columns
// into the 256 keys of row_stats_tile_kernel's layout at a.stats_part.
template <int TILE, int P, bool STATS>
__device__ __forceinline__ void epilogue(const GramArgs& a, int tr, int tc, f32x16 (&acc)[9],
                                         char* lds, int thread)
{
    static_assert(P == 1 || P == 2, "pairs per block: 1 or 2 (rows of one group of four)");
    static_assert(!STATS || TILE == 64, "row statistics: 64 x 64 tiles");
    unsigned* skeys = reinterpret_cast<unsigned*>(lds + epilogue_lds_bytes<TILE>());
    constexpr int kWaves = (TILE / 32) * (TILE / 32);
    constexpr int kThreads = 64 * kWaves;
    const int lane = thread & 63;
    const int tid = kWaves == 1 ? lane : thread;
    const int wv = kWaves == 1 ? 0 : (thread >> 6);
    const int wr = wv >> 1, wc = wv & 1;
    double* sconst = reinterpret_cast<double*>(lds);
    if constexpr (kWaves == 1) wave_lds_sync();  // earlier reads of this region are done
    for (int v = tid; v < kNumConsts * 2 * TILE; v += kThreads) {
        const int q = v / (2 * TILE), r = v % (2 * TILE);
        const bool is_row = r < TILE;
        const int conf = is_row ? tr * TILE + r : tc * TILE + (r - TILE);
        const double val = is_row ? (conf < a.n_rows ? a.consts_a[(int64_t)q * a.lda + conf] : 0.0)
                                  : (conf < a.n_cols ? a.consts_b[(int64_t)q * a.ldb + conf] : 0.0);
        sconst[r * 12 + q] = val;
    }
    if constexpr (kWaves == 1)
        wave_lds_sync();
    else
        __syncthreads();

    const int jl = wc * 32 + (lane & 31);  // column within the tile
    const int j = tc * TILE + jl;
    const bool jok = j < a.n_cols;
    double cj[kNumConsts];
#pragma unroll
    for (int q = 0; q < kNumConsts; ++q) cj[q] = sconst[(TILE + jl) * 12 + q];
    const int il0 = wr * 32 + 4 * (lane >> 5);  // first row within the tile of this lane
    const EpilogueLimits& lim = a.lim;  // uniform: scalar registers
    uint32_t flagged = 0;               // bit r: pair r goes to FP64 re-evaluation
    // P = 2 solves pairs r and r + 1 (rows il and il + 1) together: with both
    // rows' constants read from LDS at once and no branch between their two
    // solves, the scheduler interleaves the two independent FP64 chains, where
    // one pair per block leaves each chain's latency exposed.  It needs about
    // a dozen more registers than P = 1; four pairs per block spill while the
    // accumulators of the later pairs are still live.
#pragma unroll
    for (int r = 0; r < 16; r += P) {
        const int il = il0 + 8 * (r / 4) + (r % 4);  // P = 2: r % 4 is 0 or 2, row il + 1 holds pair r + 1
        const int i = tr * TILE + il;
        double c[P][kNumConsts];
        SecularSystem y[P];
        SecularBracket br[P];
        PairSolve o[P];
#pragma unroll
        for (int p = 0; p < P; ++p) {
#pragma unroll
            for (int q = 0; q < kNumConsts; ++q) c[p][q] = sconst[(il + p) * 12 + q];
        }
#pragma unroll
        for (int p = 0; p < P; ++p) {
            double E[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) E[q] = (double)acc[q][r + p];
            y[p] = secular_system(E, c[p], cj);
        }
        bool again = false;
#pragma unroll
        for (int p = 0; p < P; ++p) {
            br[p] = secular_step(y[p], 0.0);
            o[p] = secular_out(br[p], lim.solve_tol2L);
            again |= !o[p].converged & br[p].pd;
        }
        if (again) {  // rare for a few hundred atoms; taken by waves that need it
#pragma unroll
            for (int p = 0; p < P; ++p)
                if (!o[p].converged & br[p].pd) {
                    br[p] = secular_step(y[p], br[p].lo);
                    o[p] = secular_out(br[p], lim.solve_tol2L);
                }
        }
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const bool ok = jok && i + p < a.n_rows && (!a.triangle || i + p < j);
            const double sgg = c[p][10] * cj[10];
            const double err = lim.acc_coef * sgg;
            // flag: no verified root, near-duplicate residuals, estimated
            // error above tau (compared squared: no FP64 sqrt), or residuals
            // too large for the split
            const bool flag = !o[p].converged || (o[p].H < lim.dup_coef * sgg) ||
                              (err * err > lim.tau2L * o[p].rmsd2_L) || (c[p][9] + cj[9] > lim.rep_h_max);
            const int64_t oi =
                a.triangle ? tri_index(a.n_rows, i + p, j) : (int64_t)(i + p) * a.out_ld + j;
            if constexpr (STATS) {  // f32 output
                const float v = __builtin_amdgcn_sqrtf((float)(o[p].rmsd2_L * lim.inv_L));
                if (ok) static_cast<float*>(a.out)[oi] = v;
                skeys[(il + p) * 65 + jl] = ok && !flag && v == v ? fkey(__float_as_uint(v)) : 0xFFFFFFFFu;
            } else if (ok) {
                if (a.out_f64)
                    static_cast<double*>(a.out)[oi] = sqrt(o[p].rmsd2_L * lim.inv_L);
                else  // v_sqrt_f32 (1 ulp) instead of the 17-instruction correctly rounded sqrtf
                    static_cast<float*>(a.out)[oi] =
                        __builtin_amdgcn_sqrtf((float)(o[p].rmsd2_L * lim.inv_L));
            }
            flagged |= (uint32_t)(ok & flag) << (r + p);
        }
    }
    // flagged pairs are rare: one branch per tile rather than a queueing
    // branch per pair (those cost about a sixth of the epilogue)
    if (flagged) {
#pragma unroll 1
        for (int r = 0; r < 16; ++r) {
            if (!((flagged >> r) & 1u)) continue;
            const int i = tr * TILE + il0 + 8 * (r / 4) + (r % 4);
            flag_output(a, true, a.triangle ? tri_index(a.n_rows, i, j) : (int64_t)i * a.out_ld + j);
        }
    }
    if constexpr (STATS) {
        __syncthreads();  // every wave's keys are in LDS
        // tile row r and tile column r: four threads of 16 entries each, then two shuffles
        const int r = thread >> 2, q = thread & 3;
        unsigned lo = 0xFFFFFFFFu, hi = 0u, clo = 0xFFFFFFFFu, chi = 0u;
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            const unsigned kr = skeys[r * 65 + q * 16 + c];
            const unsigned kc = skeys[(q * 16 + c) * 65 + r];
            lo = min(lo, kr);
            hi = max(hi, kr == 0xFFFFFFFFu ? 0u : kr);
            clo = min(clo, kc);
            chi = max(chi, kc == 0xFFFFFFFFu ? 0u : kc);
        }
#pragma unroll
        for (int m = 1; m <= 2; m <<= 1) {
            lo = min(lo, (unsigned)__shfl_xor((int)lo, m));
            hi = max(hi, (unsigned)__shfl_xor((int)hi, m));
            clo = min(clo, (unsigned)__shfl_xor((int)clo, m));
            chi = max(chi, (unsigned)__shfl_xor((int)chi, m));
        }
        if (q == 0) {
            const int64_t b = (int64_t)tr * a.stats_T - (int64_t)tr * (tr - 1) / 2 + (tc - tr);
            unsigned* out = a.stats_part + b * 256;
            out[r] = lo;
            out[64 + r] = hi;
            out[128 + r] = clo;
            out[192 + r] = chi;
        }
    }
}

// s_waitcnt with the vector-memory counter limited to N (gfx9 encoding:
// vmcnt[3:0] | expcnt 7 << 4 | lgkmcnt << 8 | vmcnt[5:4] << 14); the _LDS0
// form also waits for every LDS access of the wave (lgkmcnt 0).
#define BS_WAIT_VMCNT(N) __builtin_amdgcn_s_waitcnt(0x0F70 | ((N) & 15) | (((N) >> 4) << 14))
#define BS_WAIT_VMCNT_LDS0(N) __builtin_amdgcn_s_waitcnt(0x0070 | ((N) & 15) | (((N) >> 4) << 14))
// lgkmcnt(0) alone
#define BS_WAIT_LDS0() __builtin_amdgcn_s_waitcnt(0xC07F)

// Direct global -> LDS 16-byte loads: lane l of the wave writes dst + 16 l.
#define BS_GLOAD_LDS(SRC, DST)                                                                  \
    __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1)))*)(SRC),     \
                                     (void __attribute__((address_space(3)))*)(DST), 16, 0, 0)

// Operand fragments of one atom block for one wave, in registers: the six
// 16-byte words of its row block (A) and of its column block (B).  BF16x3:
// word 2c / 2c+1 are the hi / lo planes of coordinate c; FP32: the two words
// of coordinate c's plane.
struct OperandRegs {
    uint4 a[6], b[6];
};

template <int MODE>
__device__ __forceinline__ void load_operands(const uint4* __restrict__ A, const uint4* __restrict__ B,
                                              int lane, OperandRegs& r)
{
#pragma unroll
    for (int u = 0; u < 6; ++u) {
        const int w = MODE == kGramBF16x3 ? u * 64 + lane : (u >> 1) * 128 + 2 * lane + (u & 1);
        r.a[u] = A[w];
        r.b[u] = B[w];
    }
}

// The MFMAs of one atom block on operands in registers (the same products,
// in the same order, as mfma_block).
template <int MODE>
__device__ __forceinline__ void mfma_regs(const OperandRegs& r, f32x16 (&acc)[9])
{
    if constexpr (MODE == kGramBF16x3) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int d = 0; d < 3; ++d)
                acc[3 * c + d] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                    as_bf16x8(r.a[2 * c]), as_bf16x8(r.b[2 * d]), acc[3 * c + d], 0, 0, 0);
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int d = 0; d < 3; ++d)
                acc[3 * c + d] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                    as_bf16x8(r.a[2 * c]), as_bf16x8(r.b[2 * d + 1]), acc[3 * c + d], 0, 0, 0);
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int d = 0; d < 3; ++d)
                acc[3 * c + d] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                    as_bf16x8(r.a[2 * c + 1]), as_bf16x8(r.b[2 * d]), acc[3 * c + d], 0, 0, 0);
    } else {
        float fa[3][8], fb[3][8];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const uint4 a0 = r.a[2 * c], a1 = r.a[2 * c + 1], b0 = r.b[2 * c], b1 = r.b[2 * c + 1];
            fa[c][0] = __uint_as_float(a0.x); fa[c][1] = __uint_as_float(a0.y);
            fa[c][2] = __uint_as_float(a0.z); fa[c][3] = __uint_as_float(a0.w);
            fa[c][4] = __uint_as_float(a1.x); fa[c][5] = __uint_as_float(a1.y);
            fa[c][6] = __uint_as_float(a1.z); fa[c][7] = __uint_as_float(a1.w);
            fb[c][0] = __uint_as_float(b0.x); fb[c][1] = __uint_as_float(b0.y);
            fb[c][2] = __uint_as_float(b0.z); fb[c][3] = __uint_as_float(b0.w);
            fb[c][4] = __uint_as_float(b1.x); fb[c][5] = __uint_as_float(b1.y);
            fb[c][6] = __uint_as_float(b1.z); fb[c][7] = __uint_as_float(b1.w);
        }
#pragma unroll
        for (int s = 0; s < 8; ++s)
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int d = 0; d < 3; ++d)
                    acc[3 * c + d] =
                        __builtin_amdgcn_mfma_f32_32x32x2f32(fa[c][s], fb[d][s], acc[3 * c + d], 0, 0, 0);
    }
}

// BF16x3 operands of one atom block for one wave, one plane (0: hi, 1: lo):
// words 2c + plane of its row block (a) and of its column block (b).
struct PlaneRegs {
    uint4 a[3], b[3];
};

__device__ __forceinline__ void load_plane(const uint4* __restrict__ A, const uint4* __restrict__ B, int lane,
                                           int plane, PlaneRegs& r)
{
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        r.a[c] = A[(2 * c + plane) * 64 + lane];
        r.b[c] = B[(2 * c + plane) * 64 + lane];
    }
}

// The hi * hi products of one atom block (the first 9 of mfma_regs' 27).
__device__ __forceinline__ void mfma_hh(const PlaneRegs& h, f32x16 (&acc)[9])
{
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d)
            acc[3 * c + d] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(h.a[c]), as_bf16x8(h.b[d]),
                                                                     acc[3 * c + d], 0, 0, 0);
}

// Cross product m (0..17) of one atom block: hi * lo for m < 9, then lo * hi,
// in mfma_regs' order, so the accumulation order is the same.
__device__ __forceinline__ void mfma_cross(const PlaneRegs& h, const PlaneRegs& l, f32x16 (&acc)[9], int m)
{
    const int q = m % 9, c = q / 3, d = q % 3;
    if (m < 9)
        acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(h.a[c]), as_bf16x8(l.b[d]), acc[q], 0, 0, 0);
    else
        acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(l.a[c]), as_bf16x8(h.b[d]), acc[q], 0, 0, 0);
}

// One 64 x 64 tile of pairs by a 4-wave work-group.  K loop over atom
// blocks with a three-stage LDS ring fed by direct global->LDS loads (no
// registers held for staging): a stage holds the four operand blocks (A0,
// A1, B0, B1: 6 KiB each) of one atom block; wave w streams block w.  Loads
// past the last block re-read the last block (in bounds, unused); the loop
// drains them before the epilogue reuses the LDS.  Steps are unrolled by six
// so that every LDS stage and register set is named statically.
//
// BF16x3.  Only the hi plane of the next block is read ahead (24 VGPRs; a
// whole second operand set does not fit beside 144 accumulators, and the
// compiler then sinks the reads to their use, exposing LDS latency).  Step k
// holds hi_k in registers and:
//   reads lo_k from its stage, and runs the 9 hi*hi products (which cover
//   those reads);
//   waits for its own loads of block k+1 (vmcnt: block k+2's 6 may still be
//   in flight) and its LDS reads, and meets the others at a barrier: block k
//   is then in every wave's registers and block k+1 in LDS;
//   reads hi_{k+1}, then issues block k+3 into block k's stage one load per
//   3 of the 18 cross products (back to back, the loads' setup stalled the
//   MFMA stream: 4 % of the launch);
//   waits for hi_{k+1} (lgkmcnt 0, long landed).  The explicit wait matters:
//   while LDS-DMA loads are in flight the compiler turns any LDS-result wait
//   into lgkmcnt(0), which before the hi*hi products would also wait for
//   lo_k just issued.
// The LDS reads come before the DMA loads in each group: read after DMA, the
// compiler cannot tell the two apart and waits vmcnt(0).
//
// FP32, and BF16x3 with SPLIT false (the persistent batch kernel, whose
// own loop state leaves no room for the split schedule: it would spill).
// The operands of block k are read into registers during step k-1; step k
// waits for its own loads of block k+1 and its LDS reads, meets the others
// at a barrier, issues block k+3 into block k's stage, reads block k+1 into
// the other register set, and runs the 27 (or 72) MFMAs of block k.
//
// PAIRS: pairs per block of the epilogue, 2 where the registers allow (the
// Gram kernel), 1 in the batch kernel.  STATS: the epilogue's row statistics.
template <int MODE, bool SPLIT, int PAIRS, bool STATS = false>
__device__ __forceinline__ void gram_tile64(const GramArgs& a, int tr, int tc,
                                            uint4 (&lds)[3][4][kPanelBlockU4], int thread)
{
    const int lane = thread & 63, wv = thread >> 6;
    const int wr = wv >> 1, wc = wv & 1;
    const uint4* __restrict__ pa = static_cast<const uint4*>(a.panels_a);
    const uint4* __restrict__ pb = static_cast<const uint4*>(a.panels_b);
    const int64_t stride_cb = (int64_t)a.n_kblk * kPanelBlockU4;
    const uint4* src = (wv < 2 ? pa + (int64_t)(2 * tr + wv) * stride_cb
                               : pb + (int64_t)(2 * tc + wv - 2) * stride_cb) + lane;
    const int n_kblk = a.n_kblk, last = n_kblk - 1;
    // the wave index as a uniform value: the loads' LDS targets (M0) are then
    // formed in SGPRs, not by a v_readfirstlane per load (1-2 % of the launch)
    const int wvu = __builtin_amdgcn_readfirstlane(wv);

#define BS_ISSUE(KB, ST)                                                          \
    {                                                                             \
        const uint4* g_ = src + (int64_t)min((KB), last) * kPanelBlockU4;         \
        _Pragma("unroll") for (int u_ = 0; u_ < 6; ++u_)                          \
            BS_GLOAD_LDS(g_ + u_ * 64, &lds[ST][wvu][u_ * 64]);                   \
    }
#define BS_STEP(KB, SCUR, SNEXT, RCUR, RNEXT)                                     \
    {                                                                             \
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");                    \
        BS_WAIT_VMCNT_LDS0(6);                                                    \
        __builtin_amdgcn_s_barrier();                                             \
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");                    \
        BS_ISSUE((KB) + 3, SCUR)                                                  \
        load_operands<MODE>(lds[SNEXT][wr], lds[SNEXT][2 + wc], lane, RNEXT);    \
        mfma_regs<MODE>(RCUR, acc);                                               \
    }

#define BS_STEP3(KB, SCUR, SNEXT, HCUR, HNEXT)                                    \
    {                                                                             \
        PlaneRegs lo_;                                                            \
        load_plane(lds[SCUR][wr], lds[SCUR][2 + wc], lane, 1, lo_);               \
        __builtin_amdgcn_sched_barrier(0);                                        \
        mfma_hh(HCUR, acc);                                                       \
        __builtin_amdgcn_sched_barrier(0);                                        \
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");                    \
        BS_WAIT_VMCNT_LDS0(6);                                                    \
        __builtin_amdgcn_s_barrier();                                             \
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");                    \
        load_plane(lds[SNEXT][wr], lds[SNEXT][2 + wc], lane, 0, HNEXT);           \
        __builtin_amdgcn_sched_barrier(0);                                        \
        {                                                                         \
            const uint4* g_ = src + (int64_t)min((KB) + 3, last) * kPanelBlockU4;  \
            _Pragma("unroll") for (int u_ = 0; u_ < 6; ++u_) {                    \
                BS_GLOAD_LDS(g_ + u_ * 64, &lds[SCUR][wvu][u_ * 64]);             \
                __builtin_amdgcn_sched_barrier(0);                                \
                mfma_cross(HCUR, lo_, acc, 3 * u_);                               \
                mfma_cross(HCUR, lo_, acc, 3 * u_ + 1);                           \
                mfma_cross(HCUR, lo_, acc, 3 * u_ + 2);                           \
                __builtin_amdgcn_sched_barrier(0);                                \
            }                                                                     \
        }                                                                         \
        BS_WAIT_LDS0();                                                           \
        __builtin_amdgcn_sched_barrier(0);                                        \
    }

    f32x16 acc[9];
#pragma unroll
    for (int q = 0; q < 9; ++q)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[q][r] = 0.0f;

    // the K loop runs at a raised wave priority and the FP64 epilogue at the
    // normal one: with two work-groups per CU in different phases, the
    // sequencer then favours the MFMA stream
    __builtin_amdgcn_s_setprio(2);
    BS_ISSUE(0, 0)
    BS_ISSUE(1, 1)
    BS_ISSUE(2, 2)
    BS_WAIT_VMCNT(12);  // own loads of block 0 landed
    __syncthreads();
    if constexpr (MODE == kGramBF16x3 && SPLIT) {
        PlaneRegs ha, hb;
        load_plane(lds[0][wr], lds[0][2 + wc], lane, 0, ha);
        BS_WAIT_LDS0();
        for (int kb = 0;; kb += 6) {
            BS_STEP3(kb, 0, 1, ha, hb)
            if (kb + 1 >= n_kblk) break;
            BS_STEP3(kb + 1, 1, 2, hb, ha)
            if (kb + 2 >= n_kblk) break;
            BS_STEP3(kb + 2, 2, 0, ha, hb)
            if (kb + 3 >= n_kblk) break;
            BS_STEP3(kb + 3, 0, 1, hb, ha)
            if (kb + 4 >= n_kblk) break;
            BS_STEP3(kb + 4, 1, 2, ha, hb)
            if (kb + 5 >= n_kblk) break;
            BS_STEP3(kb + 5, 2, 0, hb, ha)
            if (kb + 6 >= n_kblk) break;
        }
    } else {
        OperandRegs ra, rb;
        load_operands<MODE>(lds[0][wr], lds[0][2 + wc], lane, ra);
        for (int kb = 0;; kb += 6) {
            BS_STEP(kb, 0, 1, ra, rb)
            if (kb + 1 >= n_kblk) break;
            BS_STEP(kb + 1, 1, 2, rb, ra)
            if (kb + 2 >= n_kblk) break;
            BS_STEP(kb + 2, 2, 0, ra, rb)
            if (kb + 3 >= n_kblk) break;
            BS_STEP(kb + 3, 0, 1, rb, ra)
            if (kb + 4 >= n_kblk) break;
            BS_STEP(kb + 4, 1, 2, ra, rb)
            if (kb + 5 >= n_kblk) break;
            BS_STEP(kb + 5, 2, 0, rb, ra)
            if (kb + 6 >= n_kblk) break;
        }
    }
    BS_WAIT_VMCNT(0);  // no load may still write the LDS the epilogue reuses
    __syncthreads();
#undef BS_ISSUE
#undef BS_STEP
#undef BS_STEP3

    __builtin_amdgcn_s_setprio(0);
    static_assert(!STATS || sizeof(lds) >= epilogue_lds_bytes<64>() + kStatsKeysBytes,
                  "the statistics keys fit in the ring's LDS");
    epilogue<64, PAIRS, STATS>(a, tr, tc, acc, reinterpret_cast<char*>(&lds[0][0][0]), thread);
}

// One 32 x 32 tile of pairs by a single wave: operand fragments are read
// straight from the panels (L2), no LDS staging and no barriers in the K loop,
// then the epilogue in this wave's own LDS region (epilogue_lds_bytes<32>()
// bytes at `lds`).  Used for small ensembles, where 64 x 64 tiles would be
// mostly padding.
template <int MODE>
__device__ __forceinline__ void gram_tile32(const GramArgs& a, int tr, int tc, char* lds, int thread)
{
    const int lane = thread & 63;
    const uint4* __restrict__ pa = static_cast<const uint4*>(a.panels_a);
    const uint4* __restrict__ pb = static_cast<const uint4*>(a.panels_b);
    const int64_t stride_cb = (int64_t)a.n_kblk * kPanelBlockU4;
    const uint4* A = pa + (int64_t)tr * stride_cb;
    const uint4* B = pb + (int64_t)tc * stride_cb;
    f32x16 acc[9];
#pragma unroll
    for (int q = 0; q < 9; ++q)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[q][r] = 0.0f;
    __builtin_amdgcn_s_setprio(2);  // as in gram_tile64: the K loop ahead of other waves' epilogues
    for (int kb = 0; kb < a.n_kblk; ++kb)
        mfma_block<MODE>(A + (int64_t)kb * kPanelBlockU4, B + (int64_t)kb * kPanelBlockU4, lane, acc);
    __builtin_amdgcn_s_setprio(0);
    epilogue<32, 1, false>(a, tr, tc, acc, lds, thread);
}

}  // namespace bs
