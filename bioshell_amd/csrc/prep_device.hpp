// prep_device.hpp -- per-conformation preparation for the Gram path, FP64,
// as device functions shared by the single-ensemble kernels
// (kernels_prep.hip) and the batched ones (kernels_batch.hip).
//
//   centre   x_i = a_i - (sum a_i)/n                   (substructures.rs:27-42)
//   fit      rotation U_i that superposes x_i onto a reference r (Horn's
//            quaternion: top eigenvector of the 4x4 key matrix, Jacobi)
//   mean     r <- mean_i U_i x_i   (one Procrustes-mean step; r stays centred)
//   pack     e_i = U_i x_i - r split into MFMA operand panels, and the FP64
//            per-conformation constants of the shifted QCP frame (qcp.hpp)
//
// All of it is O(n L) and runs once per ensemble; none of it is per pair.
#pragma once

#include "kernels.hpp"
#include "qcp.hpp"
#include "wave.hpp"

namespace bs {

// Lane q < 9 of the wave stores m[q] to dst[q].
__device__ inline void store9(const double (&m)[9], double* dst, int lane)
{
    if (lane < 9) {
        double v = m[0];
#pragma unroll
        for (int q = 1; q < 9; ++q)
            if (lane == q) v = m[q];
        dst[lane] = v;
    }
}

// One wave centres conformation p (L atoms) into q.  With COV it also
// reduces sum (x_k - c) p0_k^T, the covariance with conformation p0 taken
// uncentred (equal to the centred one because sum (x_k - c) = 0), into cov9.
template <bool COV>
__device__ inline void centre_one(const double* __restrict__ p, int L, double* __restrict__ q,
                                  const double* __restrict__ p0, double* __restrict__ cov9,
                                  int lane)
{
    double s0 = 0, s1 = 0, s2 = 0;
    for (int k = lane; k < L; k += 64) {
        s0 += p[3 * k];
        s1 += p[3 * k + 1];
        s2 += p[3 * k + 2];
    }
    s0 = wave_sum(s0) / L;
    s1 = wave_sum(s1) / L;
    s2 = wave_sum(s2) / L;
    double m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = lane; k < L; k += 64) {
        const double b0 = p[3 * k] - s0, b1 = p[3 * k + 1] - s1, b2 = p[3 * k + 2] - s2;
        q[3 * k] = b0;
        q[3 * k + 1] = b1;
        q[3 * k + 2] = b2;
        if constexpr (COV) {
            const double a0 = p0[3 * k], a1 = p0[3 * k + 1], a2 = p0[3 * k + 2];
            m[0] += b0 * a0; m[1] += b0 * a1; m[2] += b0 * a2;
            m[3] += b1 * a0; m[4] += b1 * a1; m[5] += b1 * a2;
            m[6] += b2 * a0; m[7] += b2 * a1; m[8] += b2 * a2;
        }
    }
    if constexpr (COV) {
#pragma unroll
        for (int qq = 0; qq < 9; ++qq) m[qq] = wave_sum(m[qq]);
        store9(m, cov9, lane);
    }
}

// One wave reduces M = sum x_k r_k^T of a centred conformation into cov9.
__device__ inline void fit_cov_one(const double* __restrict__ x, int L,
                                   const double* __restrict__ ref, double* __restrict__ cov9,
                                   int lane)
{
    double m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = lane; k < L; k += 64) {
        const double b0 = x[3 * k], b1 = x[3 * k + 1], b2 = x[3 * k + 2];
        const double a0 = ref[3 * k], a1 = ref[3 * k + 1], a2 = ref[3 * k + 2];
        m[0] += b0 * a0; m[1] += b0 * a1; m[2] += b0 * a2;
        m[3] += b1 * a0; m[4] += b1 * a1; m[5] += b1 * a2;
        m[6] += b2 * a0; m[7] += b2 * a1; m[8] += b2 * a2;
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) m[q] = wave_sum(m[q]);
    store9(m, cov9, lane);
}

// Rotation that moves b onto a, from M = sum b a^T (row-major).
__device__ inline void rotation_from_cov(const double* M, double* R, double* lambda)
{
    const double xx = M[0], xy = M[1], xz = M[2], yx = M[3], yy = M[4], yz = M[5], zx = M[6],
                 zy = M[7], zz = M[8];
    double K[4][4] = {{xx + yy + zz, yz - zy, zx - xz, xy - yx},
                      {yz - zy, xx - yy - zz, xy + yx, zx + xz},
                      {zx - xz, xy + yx, -xx + yy - zz, yz + zy},
                      {xy - yx, zx + xz, yz + zy, -xx - yy + zz}};
    double q[4];
    top_eigvec4(K, q, lambda);
    const double q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
    R[0] = q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3;
    R[1] = 2.0 * (q1 * q2 - q0 * q3);
    R[2] = 2.0 * (q1 * q3 + q0 * q2);
    R[3] = 2.0 * (q1 * q2 + q0 * q3);
    R[4] = q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3;
    R[5] = 2.0 * (q2 * q3 - q0 * q1);
    R[6] = 2.0 * (q1 * q3 - q0 * q2);
    R[7] = 2.0 * (q2 * q3 + q0 * q1);
    R[8] = q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3;
}

// One thread turns the covariance in rot9 into the rotation, in place.
__device__ inline void rot_one(double* __restrict__ rot9)
{
    double m[9], R[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) m[q] = rot9[q];
    rotation_from_cov(m, R, nullptr);
#pragma unroll
    for (int q = 0; q < 9; ++q) rot9[q] = R[q];
}

// Reference mean, stage 1: slice s (conformations 32 s .. 32 s + 31) of one
// ensemble, each of the 256 threads one coordinate of the rotated sum at a
// time.  sR is a 32 x 9 LDS array.
__device__ inline void mean_partial_block(const double* __restrict__ centred,
                                          const double* __restrict__ rot, int n, int L, int s,
                                          double* __restrict__ partial, double (*sR)[9])
{
    const int i0 = s * kMeanSliceConfs;
    const int cnt = min(kMeanSliceConfs, n - i0);
    for (int v = threadIdx.x; v < cnt * 9; v += 256) sR[v / 9][v % 9] = rot[(int64_t)i0 * 9 + v];
    __syncthreads();
    for (int v = threadIdx.x; v < 3 * L; v += 256) {
        const int k = v / 3, c = v - 3 * (v / 3);
        double acc = 0.0;
        for (int t = 0; t < cnt; ++t) {
            const double* x = centred + ((int64_t)(i0 + t) * L + k) * 3;
            acc += sR[t][3 * c] * x[0] + sR[t][3 * c + 1] * x[1] + sR[t][3 * c + 2] * x[2];
        }
        partial[(int64_t)s * 3 * L + v] = acc;
    }
}

// Stage 2: one wave adds coordinate v over the slices in a fixed order
// (deterministic, no atomics).
__device__ inline void mean_final_wave(const double* __restrict__ partial, int n_slices, int n,
                                       int L, int v, double* __restrict__ ref, int lane)
{
    double acc = 0.0;
    for (int s = lane; s < n_slices; s += 64) acc += partial[(int64_t)s * 3 * L + v];
    acc = wave_sum(acc);
    if (lane == 0) ref[v] = acc / n;
}

// R0 = sum_k r_k r_k^T, stored as (xx, yy, zz, xy, xz, yz) after the 3L
// values; one 256-thread block, part is a 4 x 6 LDS array.
__device__ inline void ref_gram_block(double* __restrict__ ref, int L, double (*part)[6])
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int k = threadIdx.x; k < L; k += 256) {
        const double x = ref[3 * k], y = ref[3 * k + 1], z = ref[3 * k + 2];
        s[0] += x * x; s[1] += y * y; s[2] += z * z;
        s[3] += x * y; s[4] += x * z; s[5] += y * z;
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) s[q] = wave_sum(s[q]);
    if (lane == 0)
        for (int q = 0; q < 6; ++q) part[w][q] = s[q];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int q = threadIdx.x;
        ref[3 * L + q] = ((part[0][q] + part[1][q]) + part[2][q]) + part[3][q];
    }
}

__device__ inline uint16_t bf16_rne(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline double bf16_to_double(uint16_t h)
{
    return (double)__uint_as_float(((uint32_t)h) << 16);
}

// Pack: one wave per conformation i of one ensemble, atoms in chunks of
// kPackChunk atom blocks.  Phase 1: lanes walk consecutive atoms (coalesced
// reads), form e = U x - r, split it, accumulate g and C, and drop each
// operand value into the wave's LDS image of the panel words.  Phase 2: lanes
// copy those words out as 16-byte stores.  Rows i >= n (padding) get zero
// panels and zero constants.  Only wave-level synchronisation: the waves of a
// block may belong to different ensembles.
constexpr int kPackChunk = 16;                 // atom blocks per chunk (256 atoms)
constexpr int kPackWordsU4 = kPackChunk * 12;  // 16-byte words per chunk and conformation

template <int MODE>
__device__ inline void pack_one(const double* __restrict__ centred, const double* __restrict__ rot,
                                const double* __restrict__ ref, int n, int L, int n_kblk, int i,
                                uint4* __restrict__ panels, double* __restrict__ consts, int ld,
                                uint4* img, int lane)
{
    const bool live = i < n;
    double R[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) R[q] = live ? rot[(int64_t)i * 9 + q] : 0.0;
    const double* x = centred + (int64_t)(live ? i : 0) * L * 3;
    const int cb = i / kConfBlock, crow = i % kConfBlock;
    uint16_t* img16 = reinterpret_cast<uint16_t*>(img);
    float* img32 = reinterpret_cast<float*>(img);

    double g = 0.0;
    double C[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // sum r e~^T
    for (int kb0 = 0; kb0 < n_kblk; kb0 += kPackChunk) {
        const int nkb = min(kPackChunk, n_kblk - kb0);
        for (int kk = lane; kk < nkb * kAtomBlock; kk += 64) {
            const int k = kb0 * kAtomBlock + kk;
            const int kbl = kk / kAtomBlock, w = kk % kAtomBlock;
            double ev[3] = {0.0, 0.0, 0.0}, rv[3] = {0.0, 0.0, 0.0};
            if (live && k < L) {
                const double p0 = x[3 * k], p1 = x[3 * k + 1], p2 = x[3 * k + 2];
                rv[0] = ref[3 * k];
                rv[1] = ref[3 * k + 1];
                rv[2] = ref[3 * k + 2];
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    ev[c] = (R[3 * c] * p0 + R[3 * c + 1] * p1 + R[3 * c + 2] * p2) - rv[c];
            }
            double et[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if constexpr (MODE == kGramBF16x3) {
                    // lane-half h = w / 8, element w % 8 of planes 2c (hi), 2c+1 (lo)
                    const uint16_t hi = bf16_rne((float)ev[c]);
                    const double dhi = bf16_to_double(hi);
                    const uint16_t lo = bf16_rne((float)(ev[c] - dhi));
                    et[c] = dhi + bf16_to_double(lo);
                    const int h = w >> 3, e = w & 7;
                    img16[((kbl * 6 + 2 * c) * 2 + h) * 8 + e] = hi;
                    img16[((kbl * 6 + 2 * c + 1) * 2 + h) * 8 + e] = lo;
                } else {
                    // lane-half h = w % 2, step s = w / 2 of plane c
                    const float f = (float)ev[c];
                    et[c] = (double)f;
                    const int h = w & 1, st = w >> 1;
                    img32[((kbl * 3 + c) * 2 + h) * 8 + st] = f;
                }
            }
            g += et[0] * et[0] + et[1] * et[1] + et[2] * et[2];
#pragma unroll
            for (int p = 0; p < 3; ++p)
#pragma unroll
                for (int q = 0; q < 3; ++q) C[3 * p + q] += rv[p] * et[q];
        }
        wave_lds_sync();
        for (int u = lane; u < nkb * 12; u += 64) {
            const int kbl = u / 12, r = u % 12;
            uint4* blk = panels + ((int64_t)cb * n_kblk + kb0 + kbl) * kPanelBlockU4;
            if constexpr (MODE == kGramBF16x3) {
                const int p = r >> 1, h = r & 1;  // plane p, lane-half h
                blk[p * 64 + crow + 32 * h] = img[(kbl * 6 + p) * 2 + h];
            } else {
                const int c = r >> 2, h = (r >> 1) & 1, half = r & 1;  // 8 floats = 2 words
                blk[c * 128 + (crow + 32 * h) * 2 + half] = img[((kbl * 3 + c) * 2 + h) * 2 + half];
            }
        }
        wave_lds_sync();
    }
    g = wave_sum(g);
#pragma unroll
    for (int q = 0; q < 9; ++q) C[q] = wave_sum(C[q]);
    if (lane == 0) {
        const double* R0 = ref + 3 * L;  // xx yy zz xy xz yz
        const double trR0 = R0[0] + R0[1] + R0[2];
        const double trC = C[0] + C[4] + C[8];
        double cst[kNumConsts];
        cst[0] = (R0[0] - trR0) + 2.0 * (C[0] - trC);
        cst[1] = (R0[1] - trR0) + 2.0 * (C[4] - trC);
        cst[2] = (R0[2] - trR0) + 2.0 * (C[8] - trC);
        cst[3] = R0[3] + (C[1] + C[3]);
        cst[4] = R0[4] + (C[2] + C[6]);
        cst[5] = R0[5] + (C[5] + C[7]);
        cst[6] = C[5] - C[7];
        cst[7] = C[6] - C[2];
        cst[8] = C[1] - C[3];
        cst[9] = 0.5 * g;
        cst[10] = sqrt(g);
        if (!live)
            for (int q = 0; q < kNumConsts; ++q) cst[q] = 0.0;
#pragma unroll
        for (int q = 0; q < kNumConsts; ++q) consts[(int64_t)q * ld + i] = cst[q];
    }
}

}  // namespace bs
