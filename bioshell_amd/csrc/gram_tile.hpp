// gram_tile.hpp -- one tile of the Gram kernels: K loop and FP64 epilogue.
#pragma once

#include "gram_common.hpp"

namespace bs {

// LDS bytes the epilogue of a TILE x TILE tile needs (see below).
template <int TILE>
constexpr int epilogue_lds_bytes()
{
    return 2 * TILE * 12 * 8 + (TILE / 32) * (TILE / 32) * 4 * 9 * 64 * 4;
}

// The FP64 QCP epilogue of a TILE x TILE tile whose accumulators the calling
// waves hold: TILE 64 is four waves of a work-group (wave w owns rows
// 32 (w/2) .. +31 and columns 32 (w%2) .. +31), TILE 32 is one wave on its own.
// LDS layout (bytes) at `lds`, which the caller no longer needs:
//   [0, 2 TILE 96)        constants of the tile's rows and columns, 12 doubles each
//   then, per wave        4 accumulator slots x 9 values x 64 lanes (f32)
// Accumulators go through LDS in 4 rounds of 4 slots so that one copy of the
// FP64 solve serves all 16 pairs a lane owns.
template <int TILE>
__device__ __forceinline__ void epilogue(const GramArgs& a, int tr, int tc, f32x16 (&acc)[9],
                                         char* lds, int thread)
{
    constexpr int kWaves = (TILE / 32) * (TILE / 32);
    constexpr int kThreads = 64 * kWaves;
    const int lane = thread & 63;
    const int tid = kWaves == 1 ? lane : thread;
    const int wv = kWaves == 1 ? 0 : (thread >> 6);
    const int wr = wv >> 1, wc = wv & 1;
    constexpr int kConstBytes = 2 * TILE * 12 * 8;
    double* sconst = reinterpret_cast<double*>(lds);
    float* sacc = reinterpret_cast<float*>(lds + kConstBytes) + wv * (4 * 9 * 64);
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
    const double Ld = (double)a.L;
    const double invL = 1.0 / Ld;
    const double tau2L = a.tau * a.tau * Ld;
    // the solve's own error budget: a quarter of the flagging tolerance
    const double sec_tol2L = 0.0625 * tau2L;

#define BS_STASH(BASE)                                                   \
    _Pragma("unroll") for (int s_ = 0; s_ < 4; ++s_)                     \
        _Pragma("unroll") for (int q_ = 0; q_ < 9; ++q_)                 \
            sacc[(s_ * 9 + q_) * 64 + lane] = acc[q_][(BASE) + s_];

#pragma unroll
    for (int round = 0; round < 4; ++round) {
        BS_STASH(4 * round)  // each lane reads back only its own slots: no sync
#pragma unroll 1
        for (int sl = 0; sl < 4; sl += 2) {
            // accumulator slot r = 4 round + sl holds row 8 (r/4) + r%4 of the
            // lane's block; slots sl and sl + 1 are solved together
            const int il = il0 + 8 * round + sl;
            const int i0 = tr * TILE + il, i1 = i0 + 1;
            // the rows' constants are read from LDS where they are used
            const double* c0 = sconst + il * 12;
            const double* c1 = sconst + (il + 1) * 12;
            // Newton on the secular form of the QCP polynomial from mu = 0 for
            // both pairs in lockstep (straight-line code: the two chains
            // interleave); a second step where the bracket is still wide, and
            // the quartic where S = -B is not positive definite, run only in
            // waves that need them
            SecularBracket br[2];
            {
                double E0[9], E1[9];
#pragma unroll
                for (int q = 0; q < 9; ++q) {
                    E0[q] = (double)sacc[(sl * 9 + q) * 64 + lane];
                    E1[q] = (double)sacc[((sl + 1) * 9 + q) * 64 + lane];
                }
                br[0] = secular_step(secular_system(E0, c0, cj), 0.0);
                br[1] = secular_step(secular_system(E1, c1, cj), 0.0);
            }
            ShiftedOut o[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                o[h] = secular_out(br[h], sec_tol2L);
                if (!o[h].converged) {  // rebuilt from the accumulators in LDS
                    double Eh[9];
#pragma unroll
                    for (int q = 0; q < 9; ++q) Eh[q] = (double)sacc[((sl + h) * 9 + q) * 64 + lane];
                    const double* ch = h ? c1 : c0;
                    if (br[h].pd) {
                        br[h] = secular_step(secular_system(Eh, ch, cj), br[h].lo);
                        o[h] = secular_out(br[h], sec_tol2L);
                    }
                    if (!o[h].converged) o[h] = qcp_shifted(Eh, ch, cj);
                }
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int i = h ? i1 : i0;
                const double* ci = h ? c1 : c0;
                const bool ok = jok && i < a.n_rows && (!a.triangle || i < j);
                if (ok) {
                    const double sgg = ci[10] * cj[10];
                    const double err = a.acc_coef * sgg;
                    // flag: unconverged / near-double root, near-duplicate
                    // residuals, estimated error above tau (compared squared:
                    // no FP64 sqrt), or residuals too large for the split
                    const bool flag = !o[h].converged || (o[h].H < a.kappa_dup * a.u_rep * sgg) ||
                                      (err * err > tau2L * o[h].rmsd2_L) ||
                                      (2.0 * (ci[9] + cj[9]) > a.rep_g_max);
                    const int64_t oi =
                        a.triangle ? tri_index(a.n_rows, i, j) : (int64_t)i * a.out_ld + j;
                    if (a.out_f64)
                        static_cast<double*>(a.out)[oi] = sqrt(o[h].rmsd2_L * invL);
                    else
                        static_cast<float*>(a.out)[oi] = sqrtf((float)(o[h].rmsd2_L * invL));
                    flag_output(a, flag, oi);
                }
            }
        }
    }
#undef BS_STASH
}

// s_waitcnt with the vector-memory counter limited to N (gfx9 encoding:
// vmcnt[3:0] | expcnt 7 << 4 | lgkmcnt << 8 | vmcnt[5:4] << 14); the _LDS0
// form also waits for every LDS access of the wave (lgkmcnt 0).
#define BS_WAIT_VMCNT(N) __builtin_amdgcn_s_waitcnt(0x0F70 | ((N) & 15) | (((N) >> 4) << 14))
#define BS_WAIT_VMCNT_LDS0(N) __builtin_amdgcn_s_waitcnt(0x0070 | ((N) & 15) | (((N) >> 4) << 14))

// Direct global -> LDS 16-byte loads: lane l of the wave writes dst + 16 l.
#define BS_GLOAD_LDS(SRC, DST)                                                                  \
    __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1)))*)(SRC),     \
                                     (void __attribute__((address_space(3)))*)(DST), 16, 0, 0)

// One 64 x 64 tile of pairs by a 4-wave work-group.  K loop over atom
// blocks with a three-stage LDS ring fed by direct global->LDS loads (no
// registers held for staging): stage k holds the four operand blocks (A0, A1,
// B0, B1: 6 KiB each) of atom block k; wave w streams block w.  Step k waits
// for its own loads of stage k (vmcnt: the 6 loads of stage k+1 may still be
// in flight) and for its own LDS reads of stage k-1 (lgkmcnt 0), meets the
// others at a barrier (after which stage k-1 is free), issues atom block k+2
// into that stage, and runs the 27 (or 72) MFMAs of block k.  Steps are unrolled by three so that every LDS access names its
// stage statically; the compiler can then see that reads of stage k do not
// depend on the loads still in flight into the other stages.  Loads past the
// last block re-read the last block (in bounds, unused); the loop drains
// them before the epilogue reuses the LDS.
template <int MODE>
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

#define BS_ISSUE(KB, ST)                                                          \
    {                                                                             \
        const uint4* g_ = src + (int64_t)min((KB), last) * kPanelBlockU4;         \
        _Pragma("unroll") for (int u_ = 0; u_ < 6; ++u_)                          \
            BS_GLOAD_LDS(g_ + u_ * 64, &lds[ST][wv][u_ * 64]);                    \
    }
#define BS_STEP(KB, SC, SN)                                                       \
    {                                                                             \
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");                    \
        BS_WAIT_VMCNT_LDS0(6);                                                    \
        __builtin_amdgcn_s_barrier();                                             \
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");                    \
        BS_ISSUE((KB) + 2, SN)                                                    \
        if ((KB) < n_kblk) mfma_block<MODE>(lds[SC][wr], lds[SC][2 + wc], lane, acc); \
    }

    f32x16 acc[9];
#pragma unroll
    for (int q = 0; q < 9; ++q)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[q][r] = 0.0f;

    // the K loop runs at a raised wave priority and the FP64 epilogue at the
    // normal one: with two work-groups per CU in different phases, the
    // sequencer then favours the MFMA stream (measured 2 % faster than equal
    // priorities, 0.6 % for the opposite choice)
    __builtin_amdgcn_s_setprio(2);
    BS_ISSUE(0, 0)
    BS_ISSUE(1, 1)
    for (int kb = 0; kb < n_kblk; kb += 3) {
        BS_STEP(kb, 0, 2)
        BS_STEP(kb + 1, 1, 0)
        BS_STEP(kb + 2, 2, 1)
    }
    BS_WAIT_VMCNT(0);  // no load may still write the LDS the epilogue reuses
    __syncthreads();
#undef BS_ISSUE
#undef BS_STEP

    __builtin_amdgcn_s_setprio(0);
    epilogue<64>(a, tr, tc, acc, reinterpret_cast<char*>(&lds[0][0][0]), thread);
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
    epilogue<32>(a, tr, tc, acc, lds, thread);
}

}  // namespace bs
