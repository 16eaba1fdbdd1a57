// prep_device.hpp -- per-conformation preparation for the Gram path, FP64,
// as device functions shared by the single-ensemble kernels
// (kernels_prep.hip) and the batched ones (kernels_batch.hip).
//
//   fit0     for the S reference samples (kernels.hpp: ref_sample_count):
//            c_i = (sum x_i)/n (substructures.rs:27-42) and the rotation U_i
//            that superposes x_i - c_i onto x_0 - c_0
//   mean     r = mean over the samples of U_i (x_i - c_i): one
//            Procrustes-mean step, r centred; then R0 = sum_k r_k r_k^T
//   fit+pack U_i onto r from M = sum_k x_k r_k^T (with r centred, raw
//            coordinates give the centred covariance), then
//            e_i = U_i (x_i - c_i) - r split into MFMA operand panels, and
//            the FP64 per-conformation constants of the shifted QCP frame
//            (qcp.hpp); c_i is stored for the FP64 re-evaluation, which
//            centres on the fly
//
// Rotations come from Horn's quaternion (rotation_fast): the largest root of
// the key matrix's characteristic quartic by Newton from above, the
// eigenvector from the adjugate of K - lambda I; every lane of the wave
// computes it from the wave-reduced covariance, so no extra pass is needed.
// The first two passes read the S samples (24 bytes per atom each), the fit
// and pack reads all n conformations and writes 12 bytes per atom; all of it
// is O(n L) and runs once per ensemble, none of it per pair.
#pragma once

#include "kernels.hpp"
#include "qcp.hpp"
#include "wave.hpp"

namespace bs {

// Lane 0 of the wave stores m[0..8] to dst (a lane-indexed select would be
// lowered to a scratch array).
__device__ inline void store9(const double (&m)[9], double* dst, int lane)
{
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < 9; ++q) dst[q] = m[q];
    }
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

// Rotation matrix of a unit quaternion (row-major).
BS_HD void quat_to_rot(double q0, double q1, double q2, double q3, double* R)
{
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

// Rotation that moves b onto a, from M = sum b a^T (row-major), without
// iterating on the 4x4 matrix: lambda_max of Horn's K is the largest root of
// l^4 + c2 l^2 + c1 l + c0 (c2 = -2|M|^2, c1 = -8 det M, c0 = det K), reached
// by Newton from sqrt(3)|M|_F >= sigma1 + sigma2 + sigma3 >= lambda_max
// (monotone from above: K is symmetric, the quartic real-rooted); for a simple
// root every column of adj(K - lambda I) is a multiple of the eigenvector, and
// the one with the largest diagonal entry is taken.  Near-degenerate tops
// (tiny adjugate) fall back to power iteration.  Any rotation serves the
// shifted frame exactly (only residual sizes, hence flags, depend on it);
// this one is the optimal fit to rounding.
BS_HD void rotation_fast(const double* M, double* R)
{
    const double xx = M[0], xy = M[1], xz = M[2], yx = M[3], yy = M[4], yz = M[5], zx = M[6],
                 zy = M[7], zz = M[8];
    const double f2 = xx * xx + xy * xy + xz * xz + yx * yx + yy * yy + yz * yz + zx * zx +
                      zy * zy + zz * zz;
    const double c2 = -2.0 * f2, c1 = -8.0 * det3(M), c0 = det_key(M);
    double l = sqrt(3.0 * f2);
    for (int it = 0; it < 60; ++it) {
        const double l2 = l * l;
        const double p = (l2 + c2) * l2 + c1 * l + c0;
        const double dp = (4.0 * l2 + 2.0 * c2) * l + c1;
        if (!(dp > 0.0)) break;
        const double st = p / dp;
        l -= st;
        if (fabs(st) <= 1e-14 * fabs(l) + 1e-300) break;
    }
    const double a00 = xx + yy + zz - l, a01 = yz - zy, a02 = zx - xz, a03 = xy - yx;
    const double a11 = xx - yy - zz - l, a12 = xy + yx, a13 = zx + xz;
    const double a22 = -xx + yy - zz - l, a23 = yz + zy;
    const double a33 = -xx - yy + zz - l;
    // 2x2 minors of rows (0,1) and (2,3) of the symmetric A = K - l I
    const double s0 = a00 * a11 - a01 * a01, s1 = a00 * a12 - a01 * a02, s2 = a00 * a13 - a01 * a03;
    const double s3 = a01 * a12 - a11 * a02, s4 = a01 * a13 - a11 * a03;
    const double t5 = a22 * a33 - a23 * a23, t4 = a12 * a33 - a13 * a23, t3 = a12 * a23 - a13 * a22;
    const double t2 = a02 * a33 - a03 * a23, t1 = a02 * a23 - a03 * a22, t0 = a02 * a13 - a03 * a12;
    // adjugate (symmetric): rows j = 0..3
    const double d00 = a11 * t5 - a12 * t4 + a13 * t3;
    const double d01 = -a01 * t5 + a12 * t2 - a13 * t1;
    const double d02 = a01 * t4 - a11 * t2 + a13 * t0;
    const double d03 = -a01 * t3 + a11 * t1 - a12 * t0;
    const double d11 = a00 * t5 - a02 * t2 + a03 * t1;
    const double d12 = -a00 * t4 + a01 * t2 - a03 * t0;
    const double d13 = a00 * t3 - a01 * t1 + a02 * t0;
    const double d22 = a03 * s4 - a13 * s2 + a33 * s0;
    const double d23 = -(a02 * s4 - a12 * s2 + a23 * s0);
    const double d33 = a02 * s3 - a12 * s1 + a22 * s0;
    // the column with the largest diagonal entry (adj = k v v^T, k of fixed sign)
    const double e0 = fabs(d00), e1 = fabs(d11), e2 = fabs(d22), e3 = fabs(d33);
    double q0 = d00, q1 = d01, q2 = d02, q3 = d03, best = e0;
    if (e1 > best) { q0 = d01; q1 = d11; q2 = d12; q3 = d13; best = e1; }
    if (e2 > best) { q0 = d02; q1 = d12; q2 = d22; q3 = d23; best = e2; }
    if (e3 > best) { q0 = d03; q1 = d13; q2 = d23; q3 = d33; best = e3; }
    const double nq = q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3;
    const double sc = sqrt(3.0 * f2);
    if (!(best > 1e-12 * sc * sc * sc) || !(nq > 0.0)) {
        // (near-)repeated top eigenvalue (or M = 0): power iteration on
        // K + sc I (positive semi-definite) from (1, 1, 1, 1) gives a vector
        // in or near the top eigenspace; any such quaternion is an optimal or
        // near-optimal rotation
        const double k00 = xx + yy + zz + sc, k01 = yz - zy, k02 = zx - xz, k03 = xy - yx;
        const double k11 = xx - yy - zz + sc, k12 = xy + yx, k13 = zx + xz;
        const double k22 = -xx + yy - zz + sc, k23 = yz + zy;
        const double k33 = -xx - yy + zz + sc;
        q0 = q1 = q2 = q3 = 0.5;
        for (int it = 0; it < 32; ++it) {
            const double v0 = k00 * q0 + k01 * q1 + k02 * q2 + k03 * q3;
            const double v1 = k01 * q0 + k11 * q1 + k12 * q2 + k13 * q3;
            const double v2 = k02 * q0 + k12 * q1 + k22 * q2 + k23 * q3;
            const double v3 = k03 * q0 + k13 * q1 + k23 * q2 + k33 * q3;
            const double nv = v0 * v0 + v1 * v1 + v2 * v2 + v3 * v3;
            if (!(nv > 0.0)) break;  // M = 0: keep the current unit quaternion
            const double iv = 1.0 / sqrt(nv);
            q0 = v0 * iv; q1 = v1 * iv; q2 = v2 * iv; q3 = v3 * iv;
        }
        quat_to_rot(q0, q1, q2, q3, R);
        return;
    }
    const double inv = 1.0 / sqrt(nq);
    quat_to_rot(q0 * inv, q1 * inv, q2 * inv, q3 * inv, R);
}

// Wave-level loops over the atoms of one conformation read 4 atoms per lane
// per trip, so that 12 loads are in flight before the first is used.
#define BS_ATOM_LOOP _Pragma("unroll 4") for

// One wave, conformation x of an ensemble whose conformation 0 is x0: the
// centroid c of x into cent3 and, from the covariance of the centred sets,
// sum (x_k - c)(x0_k - c0)^T = sum x_k x0_k^T - L c c0^T, the rotation of
// x - c onto x0 - c0 into rot9.
__device__ inline void fit0_one(const double* __restrict__ x, const double* __restrict__ x0, int L,
                                double* __restrict__ rot9, double* __restrict__ cent3, int lane)
{
    double s[3] = {0, 0, 0}, s0[3] = {0, 0, 0};
    double m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    BS_ATOM_LOOP (int k = lane; k < L; k += 64) {
        const double b0 = x[3 * k], b1 = x[3 * k + 1], b2 = x[3 * k + 2];
        const double a0 = x0[3 * k], a1 = x0[3 * k + 1], a2 = x0[3 * k + 2];
        s[0] += b0; s[1] += b1; s[2] += b2;
        s0[0] += a0; s0[1] += a1; s0[2] += a2;
        m[0] += b0 * a0; m[1] += b0 * a1; m[2] += b0 * a2;
        m[3] += b1 * a0; m[4] += b1 * a1; m[5] += b1 * a2;
        m[6] += b2 * a0; m[7] += b2 * a1; m[8] += b2 * a2;
    }
    double c[3], c0[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        c[q] = wave_sum(s[q]) / L;
        c0[q] = wave_sum(s0[q]) / L;
    }
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int q = 0; q < 3; ++q) m[3 * p + q] = wave_sum(m[3 * p + q]) - L * c[p] * c0[q];
    double R[9];
    rotation_fast(m, R);
    store9(R, rot9, lane);
    if (lane == 0) {
        cent3[0] = c[0];
        cent3[1] = c[1];
        cent3[2] = c[2];
    }
}

// Reference mean, stage 1: slice s (samples 8 s .. 8 s + 7 of the S) of an
// ensemble of n conformations, each of the 256 threads one coordinate of the
// rotated centred sum at a time.  Sample t's rotation and centroid are
// entry t of rot and cent.  sR (8 x 9), sC (8 x 3) and sX (8 sample
// offsets) are LDS arrays.
__device__ inline void mean_partial_block(const double* __restrict__ xyz,
                                          const double* __restrict__ cent,
                                          const double* __restrict__ rot, int n, int S, int L,
                                          int s, double* __restrict__ partial, double (*sR)[9],
                                          double (*sC)[3], int64_t* sX)
{
    const int i0 = s * kMeanSliceConfs;
    const int cnt = min(kMeanSliceConfs, S - i0);
    for (int v = threadIdx.x; v < cnt * 9; v += 256) sR[v / 9][v % 9] = rot[(int64_t)i0 * 9 + v];
    for (int v = threadIdx.x; v < cnt * 3; v += 256) sC[v / 3][v % 3] = cent[(int64_t)i0 * 3 + v];
    if (threadIdx.x < cnt) sX[threadIdx.x] = (int64_t)ref_sample_index(i0 + threadIdx.x, n, S) * L * 3;
    __syncthreads();
    for (int v = threadIdx.x; v < 3 * L; v += 256) {
        const int k = v / 3, c = v - 3 * (v / 3);
        double acc = 0.0;
        for (int t = 0; t < cnt; ++t) {
            const double* x = xyz + sX[t] + 3 * k;
            acc += sR[t][3 * c] * (x[0] - sC[t][0]) + sR[t][3 * c + 1] * (x[1] - sC[t][1]) +
                   sR[t][3 * c + 2] * (x[2] - sC[t][2]);
        }
        partial[(int64_t)s * 3 * L + v] = acc;
    }
}

// Stage 2: coordinate v of r by one thread, in a fixed order (no atomics;
// the single-ensemble and batched kernels share it, so both give the same
// r): eight running sums over the slices s = q mod 8, fed by up to 32 loads
// in flight at a time, then added pairwise.
__device__ inline void mean_final_coord(const double* __restrict__ partial, int n_slices, int S,
                                        int L, int v, double* __restrict__ ref)
{
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const double* p = partial + v;
    const int64_t ld = 3 * (int64_t)L;
    int s = 0;
    for (; s + 32 <= n_slices; s += 32) {  // 32 loads in flight, added in the same order
        double x[32];
#pragma unroll
        for (int u = 0; u < 32; ++u) x[u] = p[(s + u) * ld];
#pragma unroll
        for (int u = 0; u < 32; ++u) acc[u & 7] += x[u];
    }
    for (; s + 8 <= n_slices; s += 8)
#pragma unroll
        for (int q = 0; q < 8; ++q) acc[q] += p[(s + q) * ld];
#pragma unroll
    for (int q = 0; q < 8; ++q)
        if (s + q < n_slices) acc[q] += p[(s + q) * ld];
    ref[v] = (((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]))) / S;
}

// Work-group ids are dealt to the 8 XCDs in turn.  A pack grid of 8 * per
// work-groups maps id b to the logical block (b % 8) * per + b / 8, so that
// each XCD packs a contiguous range of conformations: the 16-byte panel words
// of one conformation block (32 conformations, 8 work-groups of 4 waves) are
// then written through one L2, which merges them into whole lines, instead of
// arriving as partial lines from eight.  Returns -1 for the padding ids.
__device__ inline int xcd_contiguous_block(int n_blocks)
{
    const int per = (n_blocks + 7) / 8;
    const int lb = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
    return lb < n_blocks ? lb : -1;
}
inline int xcd_contiguous_grid(int n_blocks) { return (n_blocks + 7) / 8 * 8; }

// Fit + pack: one wave per conformation i of one ensemble.  Pass 1 reduces
// sum x_k and M = sum x_k r_k^T (raw coordinates; r is centred) and turns M
// into U_i (every lane).  Pass 2 (a second read of x, mostly from cache)
// walks the atoms in chunks of kPackChunk atom blocks: lanes take consecutive
// atoms, form e = U_i (x - c_i) - r, split it, accumulate g and
// C = sum r e~^T, and drop each operand value into the wave's LDS image of
// the panel words; then lanes copy the words out as 16-byte stores.  Rows
// i >= n (padding) get zero panels and constants.  Only wave-level
// synchronisation: the waves of a block may belong to different ensembles.
constexpr int kPackChunk = 16;                 // atom blocks per chunk (256 atoms)
constexpr int kPackWordsU4 = kPackChunk * 12;  // 16-byte words per chunk and conformation

template <int MODE>
__device__ inline void fit_pack_one(const double* __restrict__ xyz, const double* __restrict__ ref,
                                    int n, int L, int n_kblk, int i, double* __restrict__ rot,
                                    double* __restrict__ cent, uint4* __restrict__ panels,
                                    double* __restrict__ consts, int ld, uint4* img, int lane)
{
    const bool live = i < n;
    const double* x = xyz + (int64_t)(live ? i : 0) * L * 3;
    double R[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double c0 = 0.0, c1 = 0.0, c2 = 0.0;
    if (live) {
        double m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        BS_ATOM_LOOP (int k = lane; k < L; k += 64) {
            const double b0 = x[3 * k], b1 = x[3 * k + 1], b2 = x[3 * k + 2];
            const double a0 = ref[3 * k], a1 = ref[3 * k + 1], a2 = ref[3 * k + 2];
            c0 += b0; c1 += b1; c2 += b2;
            m[0] += b0 * a0; m[1] += b0 * a1; m[2] += b0 * a2;
            m[3] += b1 * a0; m[4] += b1 * a1; m[5] += b1 * a2;
            m[6] += b2 * a0; m[7] += b2 * a1; m[8] += b2 * a2;
        }
        c0 = wave_sum(c0) / L;
        c1 = wave_sum(c1) / L;
        c2 = wave_sum(c2) / L;
#pragma unroll
        for (int q = 0; q < 9; ++q) m[q] = wave_sum(m[q]);
        rotation_fast(m, R);
        store9(R, rot + (int64_t)i * 9, lane);
        if (lane == 0) {
            cent[3 * (int64_t)i] = c0;
            cent[3 * (int64_t)i + 1] = c1;
            cent[3 * (int64_t)i + 2] = c2;
        }
    }
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
                const double p0 = x[3 * k] - c0, p1 = x[3 * k + 1] - c1, p2 = x[3 * k + 2] - c2;
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
#pragma unroll
            for (int q = 0; q < kNumConsts; ++q) cst[q] = 0.0;
#pragma unroll
        for (int q = 0; q < kNumConsts; ++q) consts[(int64_t)q * ld + i] = cst[q];
    }
}

}  // namespace bs
