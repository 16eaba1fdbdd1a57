// prep_device.hpp -- per-conformation preparation for the Gram path, FP64,
// as device functions shared by the single-ensemble kernels
// (kernels_prep.hip) and the batched ones (kernels_batch.hip).
//
//   fit0     for the S reference samples (kernels.hpp: ref_sample_count):
//            c_i = (sum x_i)/n (substructures.rs:27-42) and the rotation U_i
//            that superposes x_i - c_i onto x_0 - c_0
//   mean     r = mean over the samples of U_i (x_i - c_i): one
//            Procrustes-mean step, r centred (partial sums of 16 samples in
//            the fit0 pass, then a fixed-order sum); then R0 = sum_k r_k r_k^T
//   fit+pack U_i onto r from M = sum_k x_k r_k^T (with r centred, raw
//            coordinates give the centred covariance), then
//            e_i = U_i (x_i - c_i) - r split into MFMA operand panels, and
//            the FP64 per-conformation constants of the shifted QCP frame
//            (qcp.hpp); c_i is stored for the FP64 re-evaluation, which
//            centres on the fly
//
// Rotations come from Horn's quaternion (rotation_fast): the largest root of
// the key matrix's characteristic quartic by Newton from above, the
// eigenvector from the adjugate of K - lambda I, one thread per conformation
// from the work-group's reduced sums.  The sample fits read the S samples
// twice (the fit, then the rotated sum), the fit and pack reads all n
// conformations twice and writes 12 bytes per atom; all of it is O(n L) and
// runs once per ensemble, none of it per pair.
#pragma once

#include "kernels.hpp"
#include "qcp.hpp"
#include "wave.hpp"

namespace bs {

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
// this one is the optimal fit to rounding.  upper, when positive, is a
// tighter bound to start from: E0 = (G_a + G_b) / 2 >= lambda_max, within
// L rmsd^2 / 2 of it, which saves most of the steps for sets that fit well.
BS_HD void rotation_fast(const double* M, double* R, double upper = 0.0)
{
    const double xx = M[0], xy = M[1], xz = M[2], yx = M[3], yy = M[4], yz = M[5], zx = M[6],
                 zy = M[7], zz = M[8];
    const double f2 = xx * xx + xy * xy + xz * xz + yx * yx + yy * yy + yz * yz + zx * zx +
                      zy * zy + zz * zz;
    const double c2 = -2.0 * f2, c1 = -8.0 * det3(M), c0 = det_key(M);
    double l = sqrt(3.0 * f2);
    if (upper > 0.0 && upper < l) l = upper;
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

// ------------------------------------------------------ per-conformation fits
// A work-group of 256 threads fits kFitConfs = 16 conformations at a time,
// kFitLanes = 16 consecutive threads per conformation; thread g of a
// conformation takes its atoms g, g + 16, g + 32, ...  Pass 1 issues its
// loads in batches of 8 atoms per thread; the second passes load chunk j + 1
// while chunk j is processed (the first chunk while the rotations are
// solved), so that a chain costs few exposed memory round trips.  Sums over
// the atoms are reduced through LDS, transposed: every thread leaves its
// partial sums, then one thread per (conformation, value) adds the 16
// partials in lane order.  The rotation of each conformation is solved by one
// thread, so a wave solves up to 64 at once: with a wave per conformation,
// the reductions and the solve (a few hundred FP64 instructions) would be
// issued once per conformation and would bound the kernels.
constexpr int kFitChunk = 4;                           // atom blocks per image chunk
constexpr int kFitImageU4 = kFitChunk * 12 * kFitConfs;  // 24 KiB of operand words
constexpr int kFitMeanAtoms = 64;                      // atoms per transpose of the reference mean
static_assert(kFitConfs * kFitLanes == 256 && kFitLanes == kAtomBlock, "fit work-group shape");

// LDS of one fit work-group summing up to NV values per conformation.
template <int NV>
struct FitLds {
    static constexpr int kValues = NV;
    static constexpr int kStride = NV * kFitConfs + 1;  // one lane's partials (odd: spreads banks)
    union {  // the reductions, the image and the mean's transposes are separated by barriers
        double red[kFitLanes * kStride];
        uint4 img[kFitImageU4];
        double ybuf[3 * kFitMeanAtoms * kFitConfs];
    };
    double tot[kFitConfs][NV];  // reduced sums
    double rc[kFitConfs][12];   // rotation (9) and centroid (3)
};
using FitPackLds = FitLds<13>;
using Fit0Lds = FitLds<17>;

// tot[c][q] = sum over the 16 lanes of conformation c of v[q], q < NV.
// Starts and ends with the work-group at one point (barriers inside).
template <int NV, class Lds>
__device__ inline void fit_reduce(const double (&v)[NV], Lds& lds, int c, int g)
{
    constexpr int S = Lds::kStride;
    static_assert(NV <= Lds::kValues, "fit reduction shape");
#pragma unroll
    for (int q = 0; q < NV; ++q) lds.red[g * S + q * kFitConfs + c] = v[q];
    __syncthreads();
    for (int t = threadIdx.x; t < NV * kFitConfs; t += 256) {  // NV * 16 may exceed 256
        double a = lds.red[t];
#pragma unroll
        for (int u = 1; u < kFitLanes; ++u) a += lds.red[u * S + t];
        lds.tot[t % kFitConfs][t / kFitConfs] = a;
    }
    __syncthreads();
}

// Reference mean, stage 1 -- slice s of an ensemble of n conformations:
// samples 16 s .. 16 s + 15 (those below S).  For each sample, the centroid
// c and, from the covariance of the centred sets,
// sum (x_k - c)(x0_k - c0)^T = sum x_k x0_k^T - L c c0^T (x0 = conformation
// 0 = xyz), the rotation U of x - c onto x0 - c0 (Newton from
// E0 = (G_x + G_x0) / 2); then a second read of the samples forms U (x_k - c)
// and sums it over the slice into partial[s] (3L values), kFitMeanAtoms
// atoms at a time through LDS.
__device__ inline void fit0_slice(const double* __restrict__ xyz, int n, int S, int L, int s,
                                  double* __restrict__ partial, Fit0Lds& lds)
{
    const int tid = threadIdx.x, c = tid / kFitLanes, g = tid % kFitLanes;
    const int t0 = s * kFitConfs;
    const bool live = t0 + c < S;
    const double* x = xyz + (int64_t)(live ? ref_sample_index(t0 + c, n, S) : 0) * L * 3;
    {
        double v[17];
#pragma unroll
        for (int q = 0; q < 17; ++q) v[q] = 0.0;
        if (live) {
#pragma unroll 8
            for (int k = g; k < L; k += kFitLanes) {
                const double b0 = x[3 * k], b1 = x[3 * k + 1], b2 = x[3 * k + 2];
                const double a0 = xyz[3 * k], a1 = xyz[3 * k + 1], a2 = xyz[3 * k + 2];
                v[0] += b0; v[1] += b1; v[2] += b2;
                v[3] += a0; v[4] += a1; v[5] += a2;
                v[6] += b0 * a0; v[7] += b0 * a1; v[8] += b0 * a2;
                v[9] += b1 * a0; v[10] += b1 * a1; v[11] += b1 * a2;
                v[12] += b2 * a0; v[13] += b2 * a1; v[14] += b2 * a2;
                v[15] += b0 * b0 + b1 * b1 + b2 * b2;
                v[16] += a0 * a0 + a1 * a1 + a2 * a2;
            }
        }
        fit_reduce(v, lds, c, g);
    }
    // the second pass's chunks: kPer atoms per thread
    constexpr int kPer = kFitMeanAtoms / kFitLanes;
    double xc[kPer][3], xn[kPer][3];
    // every call issues the same 3 kPer loads (addresses clamped to atom 0,
    // values zeroed): the wait for one register set never drains the other's
    auto load = [&](int k0, double (&dst)[kPer][3]) {
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int k = k0 + g + kFitLanes * u;
            const bool in = live && k < L;
            const int kk = in ? k : 0;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const double v = x[3 * kk + q];
                dst[u][q] = in ? v : 0.0;
            }
        }
    };
    load(0, xc);  // in flight while the rotations are solved
    __builtin_amdgcn_sched_barrier(0);
    if (tid < kFitConfs) {
        const double* sm = lds.tot[tid];
        double cc[3] = {0, 0, 0}, c0[3], m[9], R[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            cc[q] = sm[q] / L;
            c0[q] = sm[3 + q] / L;
        }
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int q = 0; q < 3; ++q) m[3 * p + q] = sm[6 + 3 * p + q] - L * cc[p] * c0[q];
        const double gx = sm[15] - L * (cc[0] * cc[0] + cc[1] * cc[1] + cc[2] * cc[2]);
        const double g0 = sm[16] - L * (c0[0] * c0[0] + c0[1] * c0[1] + c0[2] * c0[2]);
        if (t0 + tid < S) rotation_fast(m, R, 0.5 * (gx + g0));
#pragma unroll
        for (int q = 0; q < 9; ++q) lds.rc[tid][q] = R[q];
#pragma unroll
        for (int q = 0; q < 3; ++q) lds.rc[tid][9 + q] = cc[q];
    }
    __syncthreads();
    double R[9], cc[3];
#pragma unroll
    for (int q = 0; q < 9; ++q) R[q] = lds.rc[c][q];
#pragma unroll
    for (int q = 0; q < 3; ++q) cc[q] = lds.rc[c][9 + q];
    double* out = partial + (int64_t)s * 3 * L;
    auto chunk = [&](int k0, const double (&xd)[kPer][3]) {
        double y[kPer][3];
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int k = k0 + g + kFitLanes * u;
            y[u][0] = y[u][1] = y[u][2] = 0.0;
            if (live && k < L) {
                const double p0 = xd[u][0] - cc[0], p1 = xd[u][1] - cc[1], p2 = xd[u][2] - cc[2];
#pragma unroll
                for (int q = 0; q < 3; ++q) y[u][q] = R[3 * q] * p0 + R[3 * q + 1] * p1 + R[3 * q + 2] * p2;
            }
        }
#pragma unroll
        for (int u = 0; u < kPer; ++u)
#pragma unroll
            for (int q = 0; q < 3; ++q) lds.ybuf[((g + kFitLanes * u) * 3 + q) * kFitConfs + c] = y[u][q];
        __syncthreads();
        if (tid < 3 * kFitMeanAtoms && k0 + tid / 3 < L) {
            const double* yb = lds.ybuf + tid * kFitConfs;
            double a = yb[0];
#pragma unroll
            for (int w = 1; w < kFitConfs; ++w) a += yb[w];
            out[3 * k0 + tid] = a;
        }
        __syncthreads();
    };
    // two chunks per trip, ping-ponging between the register sets
    for (int k0 = 0; k0 < L; k0 += 2 * kFitMeanAtoms) {
        load(k0 + kFitMeanAtoms, xn);
        __builtin_amdgcn_sched_barrier(0);  // keep the prefetch ahead of the chunk's work
        chunk(k0, xc);
        if (k0 + kFitMeanAtoms >= L) break;
        load(k0 + 2 * kFitMeanAtoms, xc);
        __builtin_amdgcn_sched_barrier(0);
        chunk(k0 + kFitMeanAtoms, xn);
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

// Work-group ids are dealt to the 8 XCDs in turn.  A grid of 8 * per
// work-groups maps id b to the logical block (b % 8) * per + b / 8, so that
// each XCD packs a contiguous range of conformations: the panel words of one
// conformation block (32 conformations, 2 work-groups)
// are then written through one L2, which merges them into whole lines,
// instead of arriving as partial lines from several.  Returns -1 for the
// padding ids.
__device__ inline int xcd_contiguous_block(int n_blocks)
{
    const int per = (n_blocks + 7) / 8;
    const int lb = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
    return lb < n_blocks ? lb : -1;
}
inline int xcd_contiguous_grid(int n_blocks) { return (n_blocks + 7) / 8 * 8; }

// One split operand value set: e (3 coordinates) of atom w of the image's
// atom block kbl for conformation c, as BF16 hi/lo pairs or FP32; et
// receives what the panels hold.  Image words: (kbl, plane * 2 + lane-half,
// c) in BF16x3 mode (8 bf16 each), (kbl, plane * 2 + lane-half, c, half) in
// FP32 mode (4 floats each), so that a conformation block's panel rows are
// runs of 16 (32) consecutive 16-byte words.
template <int MODE>
__device__ inline void emit_operands(uint4* img, int kbl, int w, int c, const double (&ev)[3],
                                     double (&et)[3])
{
    constexpr int C = kFitConfs;
    uint16_t* img16 = reinterpret_cast<uint16_t*>(img);
    float* img32 = reinterpret_cast<float*>(img);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        if constexpr (MODE == kGramBF16x3) {
            // planes 2q (hi) and 2q + 1 (lo), lane-half w / 8, element w % 8
            const uint16_t hi = bf16_rne((float)ev[q]);
            const double dhi = bf16_to_double(hi);
            const uint16_t lo = bf16_rne((float)(ev[q] - dhi));
            et[q] = dhi + bf16_to_double(lo);
            const int h = w >> 3, e = w & 7;
            img16[((kbl * 12 + (2 * q) * 2 + h) * C + c) * 8 + e] = hi;
            img16[((kbl * 12 + (2 * q + 1) * 2 + h) * C + c) * 8 + e] = lo;
        } else {
            // plane q, lane-half w % 2, step st = w / 2: word half st / 4, float st % 4
            const float f = (float)ev[q];
            et[q] = (double)f;
            const int h = w & 1, st = w >> 1;
            img32[(((kbl * 3 + q) * 2 + h) * 2 * C + 2 * c + (st >> 2)) * 4 + (st & 3)] = f;
        }
    }
}

// The image's nkb atom blocks to their panel blocks (blk0: the first one's
// words, conformation rows crow0 .. crow0 + 15).
template <int MODE>
__device__ inline void store_image(const uint4* img, int nkb, uint4* __restrict__ blk0, int crow0)
{
    constexpr int C = kFitConfs;
    for (int u = threadIdx.x; u < nkb * 12 * C; u += 256) {
        const int kbl = u / (12 * C), rem = u % (12 * C);
        uint4* blk = blk0 + (int64_t)kbl * kPanelBlockU4;
        if constexpr (MODE == kGramBF16x3) {
            const int ph = rem / C, cc = rem % C;  // plane * 2 + lane-half, conformation
            blk[(ph >> 1) * 64 + crow0 + cc + 32 * (ph & 1)] = img[u];
        } else {
            const int ph = rem / (2 * C), r2 = rem % (2 * C);
            blk[(ph >> 1) * 128 + (crow0 + (r2 >> 1) + 32 * (ph & 1)) * 2 + (r2 & 1)] = img[u];
        }
    }
}

// Fit + pack of conformations i0 .. i0 + 15 of one ensemble (i0 a multiple
// of 16; rows i >= n are padding: zero panels and constants).  Pass 1 reduces
// sum x_k, sum |x_k|^2 and M = sum x_k r_k^T (raw coordinates: r is centred,
// so M is the centred covariance) and one thread per conformation turns M
// into U_i (Newton from E0 = (G_x + G_r) / 2).  Pass 2 (a second read of x)
// walks the atoms in chunks of kFitChunk atom blocks, one atom of each block
// per thread: it forms e = U_i (x - c_i) - r, splits it, accumulates
// g = |e~|^2 and C = sum r e~^T and drops each operand value into the LDS
// image of the chunk's panel words, which the work-group then writes out.
template <int MODE>
__device__ inline void fit_pack_block(const double* __restrict__ xyz, const double* __restrict__ ref,
                                      int n, int L, int n_kblk, int i0, double* __restrict__ rot,
                                      double* __restrict__ cent, uint4* __restrict__ panels,
                                      double* __restrict__ consts, int ld, FitPackLds& lds)
{
    const int tid = threadIdx.x, c = tid / kFitLanes, g = tid % kFitLanes;
    const int i = i0 + c;
    const bool live = i < n;
    const double* x = xyz + (int64_t)(live ? i : 0) * L * 3;
    {
        double v[13];
#pragma unroll
        for (int q = 0; q < 13; ++q) v[q] = 0.0;
        if (live) {
#pragma unroll 8
            for (int k = g; k < L; k += kFitLanes) {
                const double b0 = x[3 * k], b1 = x[3 * k + 1], b2 = x[3 * k + 2];
                const double a0 = ref[3 * k], a1 = ref[3 * k + 1], a2 = ref[3 * k + 2];
                v[0] += b0; v[1] += b1; v[2] += b2;
                v[3] += b0 * a0; v[4] += b0 * a1; v[5] += b0 * a2;
                v[6] += b1 * a0; v[7] += b1 * a1; v[8] += b1 * a2;
                v[9] += b2 * a0; v[10] += b2 * a1; v[11] += b2 * a2;
                v[12] += b0 * b0 + b1 * b1 + b2 * b2;
            }
        }
        fit_reduce(v, lds, c, g);
    }
    // pass 2's chunks: x and r of one atom per block, kFitChunk blocks
    double xb[kFitChunk][3], rb[kFitChunk][3], xn[kFitChunk][3], rn[kFitChunk][3];
    // Every call issues the same 6 kFitChunk loads (addresses clamped to atom
    // 0, values zeroed), so that the wait for one register set never has to
    // drain the loads of the other.
    auto load = [&](int kb0, double (&xd)[kFitChunk][3], double (&rd)[kFitChunk][3]) {
#pragma unroll
        for (int kbl = 0; kbl < kFitChunk; ++kbl) {
            const int k = (kb0 + kbl) * kAtomBlock + g;
            const bool in = live && kb0 + kbl < n_kblk && k < L;
            const int kk = in ? k : 0;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const double xv = x[3 * kk + q], rv = ref[3 * kk + q];
                xd[kbl][q] = in ? xv : 0.0;
                rd[kbl][q] = in ? rv : 0.0;
            }
        }
    };
    load(0, xb, rb);  // in flight while the rotations are solved
    __builtin_amdgcn_sched_barrier(0);
    const double* R0 = ref + 3 * L;  // xx yy zz xy xz yz
    const double trR0 = R0[0] + R0[1] + R0[2];
    if (tid < kFitConfs) {
        const int64_t ii = i0 + tid;
        double R[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, cc[3] = {0, 0, 0};
        if (ii < n) {
            const double* sm = lds.tot[tid];
            double m[9];
#pragma unroll
            for (int q = 0; q < 3; ++q) cc[q] = sm[q] / L;
#pragma unroll
            for (int q = 0; q < 9; ++q) m[q] = sm[3 + q];
            const double gx = sm[12] - L * (cc[0] * cc[0] + cc[1] * cc[1] + cc[2] * cc[2]);
            rotation_fast(m, R, 0.5 * (gx + trR0));
#pragma unroll
            for (int q = 0; q < 9; ++q) rot[ii * 9 + q] = R[q];
#pragma unroll
            for (int q = 0; q < 3; ++q) cent[ii * 3 + q] = cc[q];
        }
#pragma unroll
        for (int q = 0; q < 9; ++q) lds.rc[tid][q] = R[q];
#pragma unroll
        for (int q = 0; q < 3; ++q) lds.rc[tid][9 + q] = cc[q];
    }
    __syncthreads();
    double R[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) R[q] = lds.rc[c][q];
    const double c0 = lds.rc[c][9], c1 = lds.rc[c][10], c2 = lds.rc[c][11];

    const int cb = i0 / kConfBlock, crow0 = i0 % kConfBlock;
    uint4* blk0 = panels + (int64_t)cb * n_kblk * kPanelBlockU4;
    double gg = 0.0;
    double Cr[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // sum r e~^T
    auto chunk = [&](int kb0, const double (&xd)[kFitChunk][3], const double (&rd)[kFitChunk][3]) {
        const int nkb = min(kFitChunk, n_kblk - kb0);
#pragma unroll
        for (int kbl = 0; kbl < kFitChunk; ++kbl) {
            if (kbl >= nkb) break;  // uniform
            const int k = (kb0 + kbl) * kAtomBlock + g;
            double ev[3] = {0.0, 0.0, 0.0};
            if (live && k < L) {
                const double p0 = xd[kbl][0] - c0, p1 = xd[kbl][1] - c1, p2 = xd[kbl][2] - c2;
#pragma unroll
                for (int q = 0; q < 3; ++q)
                    ev[q] = (R[3 * q] * p0 + R[3 * q + 1] * p1 + R[3 * q + 2] * p2) - rd[kbl][q];
            }
            double et[3];
            emit_operands<MODE>(lds.img, kbl, g, c, ev, et);
            gg += et[0] * et[0] + et[1] * et[1] + et[2] * et[2];
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) Cr[3 * a + b] += rd[kbl][a] * et[b];
        }
        __syncthreads();
        store_image<MODE>(lds.img, nkb, blk0 + (int64_t)kb0 * kPanelBlockU4, crow0);
        __syncthreads();
    };
    // two chunks per trip, ping-ponging between the register sets: each
    // chunk's loads were issued one chunk earlier
    for (int kb0 = 0; kb0 < n_kblk; kb0 += 2 * kFitChunk) {
        load(kb0 + kFitChunk, xn, rn);
        __builtin_amdgcn_sched_barrier(0);  // keep the prefetch ahead of the chunk's work
        chunk(kb0, xb, rb);
        if (kb0 + kFitChunk >= n_kblk) break;
        load(kb0 + 2 * kFitChunk, xb, rb);
        __builtin_amdgcn_sched_barrier(0);
        chunk(kb0 + kFitChunk, xn, rn);
    }
    {
        double v[10] = {gg, Cr[0], Cr[1], Cr[2], Cr[3], Cr[4], Cr[5], Cr[6], Cr[7], Cr[8]};
        fit_reduce(v, lds, c, g);
    }
    if (tid < kFitConfs) {
        const int ii = i0 + tid;
        const double* sm = lds.tot[tid];
        const double g2 = sm[0];
        double Cs[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) Cs[q] = sm[1 + q];
        const double trC = Cs[0] + Cs[4] + Cs[8];
        double cst[kNumConsts];
        cst[0] = (R0[0] - trR0) + 2.0 * (Cs[0] - trC);
        cst[1] = (R0[1] - trR0) + 2.0 * (Cs[4] - trC);
        cst[2] = (R0[2] - trR0) + 2.0 * (Cs[8] - trC);
        cst[3] = R0[3] + (Cs[1] + Cs[3]);
        cst[4] = R0[4] + (Cs[2] + Cs[6]);
        cst[5] = R0[5] + (Cs[5] + Cs[7]);
        cst[6] = Cs[5] - Cs[7];
        cst[7] = Cs[6] - Cs[2];
        cst[8] = Cs[1] - Cs[3];
        cst[9] = 0.5 * g2;
        cst[10] = sqrt(g2);
        if (ii >= n)
#pragma unroll
            for (int q = 0; q < kNumConsts; ++q) cst[q] = 0.0;
#pragma unroll
        for (int q = 0; q < kNumConsts; ++q) consts[(int64_t)q * ld + ii] = cst[q];
    }
}

}  // namespace bs
