// gram_common.hpp -- pieces shared by the Gram kernels.
#pragma once

#include "kernels.hpp"
#include "qcp.hpp"
#include "wave.hpp"

namespace bs {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kXcd = 8;  // MI355X: work-group ids are dealt to 8 XCDs round-robin

__device__ inline int64_t tri_index(int64_t n, int64_t i, int64_t j)
{
    return i * n - i * (i + 1) / 2 + (j - i - 1);
}

__device__ inline bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

// One atom block of the nine 32x32 accumulators of a wave: rows from the A
// stage block (32 conformations), columns from the B stage block.
template <int MODE>
__device__ __forceinline__ void mfma_block(const uint4* __restrict__ A, const uint4* __restrict__ B,
                                           int lane, f32x16 (&acc)[9])
{
    if constexpr (MODE == kGramBF16x3) {
        bf16x8 ah[3], al[3], bh[3], bl[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            ah[c] = as_bf16x8(A[(2 * c) * 64 + lane]);
            al[c] = as_bf16x8(A[(2 * c + 1) * 64 + lane]);
            bh[c] = as_bf16x8(B[(2 * c) * 64 + lane]);
            bl[c] = as_bf16x8(B[(2 * c + 1) * 64 + lane]);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int d = 0; d < 3; ++d)
                acc[3 * c + d] =
                    __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[c], bh[d], acc[3 * c + d], 0, 0, 0);
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int d = 0; d < 3; ++d)
                acc[3 * c + d] =
                    __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[c], bl[d], acc[3 * c + d], 0, 0, 0);
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int d = 0; d < 3; ++d)
                acc[3 * c + d] =
                    __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[c], bh[d], acc[3 * c + d], 0, 0, 0);
    } else {
        float fa[3][8], fb[3][8];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const uint4 a0 = A[c * 128 + 2 * lane], a1 = A[c * 128 + 2 * lane + 1];
            const uint4 b0 = B[c * 128 + 2 * lane], b1 = B[c * 128 + 2 * lane + 1];
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
                    acc[3 * c + d] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[c][s], fb[d][s],
                                                                          acc[3 * c + d], 0, 0, 0);
    }
}

// Row of packed-triangle position o (row-major upper triangle of n, i < j).
__device__ inline int tri_row(int64_t o, int n)
{
    const double b = 2.0 * n - 1.0;
    int i = (int)floor((b - sqrt(b * b - 8.0 * (double)o)) * 0.5);
    auto start = [n](int64_t r) { return r * n - r * (r + 1) / 2; };
    while (i > 0 && start(i) > o) --i;
    while (start(i + 1) <= o) ++i;
    return i;
}

// FP64 cRMSD of one pair from raw coordinates x, y (L atoms each) and their
// centroids cx, cy, by one wave: the 64 lanes split the atoms, centring on the
// fly, then the textbook QCP with the near-degenerate polish.  Every lane
// returns the result.
__device__ inline double fallback_pair(const double* __restrict__ x, const double* __restrict__ cx,
                                       const double* __restrict__ y, const double* __restrict__ cy,
                                       int L, int lane)
{
    const double x0 = cx[0], x1 = cx[1], x2 = cx[2], y0 = cy[0], y1 = cy[1], y2 = cy[2];
    double m9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, ga = 0.0, gb = 0.0;
    for (int k = lane; k < L; k += 64) {
        const double a0 = x[3 * k] - x0, a1 = x[3 * k + 1] - x1, a2 = x[3 * k + 2] - x2;
        const double b0 = y[3 * k] - y0, b1 = y[3 * k + 1] - y1, b2 = y[3 * k + 2] - y2;
        m9[0] += a0 * b0; m9[1] += a0 * b1; m9[2] += a0 * b2;
        m9[3] += a1 * b0; m9[4] += a1 * b1; m9[5] += a1 * b2;
        m9[6] += a2 * b0; m9[7] += a2 * b1; m9[8] += a2 * b2;
        ga += a0 * a0 + a1 * a1 + a2 * a2;
        gb += b0 * b0 + b1 * b1 + b2 * b2;
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) m9[q] = wave_sum(m9[q]);
    ga = wave_sum(ga);
    gb = wave_sum(gb);
    return qcp_rmsd<double>(m9, ga, gb, L);
}

// Marks output o for FP64 re-evaluation: queued in the list while it has
// room, else a bit in the bitmap.
__device__ __forceinline__ void flag_output(const GramArgs& a, bool flag, int64_t o)
{
    if (!flag) return;
    if (a.flag_list) {
        const uint32_t pos = atomicAdd(a.flag_count, 1u);
        if (pos < (uint32_t)a.flag_cap) {
            a.flag_list[pos] = o;
            return;
        }
    }
    atomicOr(a.flag_bits + (o >> 5), 1u << (unsigned)(o & 31));
}

}  // namespace bs
