// kernels_superpose.hip -- optimal superposition of every conformation of an
// ensemble onto one reference, in FP64: rotation R and translation t with
// ref ~= R x + t (the Rototranslation convention of
// bioshell-pdb/src/calc/transformation.rs:22-37, applied as at 171-174), the
// RMSD of the fit and, optionally, the superposed coordinates R x + t.
//
// Three passes: one 256-thread block per conformation reduces the centroid and
// the covariance against the centred reference (the conformation, 24 L bytes,
// comes from HBM once; its second read hits L2); one thread per conformation
// turns that into Horn's quaternion (Jacobi on the 4x4 key matrix), R, t and
// the RMSD; one thread per atom writes R x + t.  HBM traffic per conformation:
// 24 L read for the fit, plus 24 L read and 24 L written for the coordinates.
#include <algorithm>

#include "kernels.hpp"
#include "prep_device.hpp"
#include "wave.hpp"

namespace bs {
namespace {

constexpr int kSupThreads = 256;
constexpr int kSupWaves = kSupThreads / 64;

// Sums v[0..N) over the block; every thread receives the totals.
template <int N>
__device__ inline void block_sum(double* v, double* lds)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < N; ++q) v[q] = wave_sum(v[q]);
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < N; ++q) lds[w * N + q] = v[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < N; ++q) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < kSupWaves; ++k) s += lds[k * N + q];
        v[q] = s;
    }
    __syncthreads();
}

// Reference: centroid (info[0..2]), centred copy, squared norm (info[3]).
__global__ __launch_bounds__(kSupThreads) void sup_ref_kernel(const double* __restrict__ ref, int L,
                                                              double* __restrict__ ref_c,
                                                              double* __restrict__ info)
{
    __shared__ double lds[kSupWaves * 3];
    double s[3] = {0.0, 0.0, 0.0};
    for (int k = threadIdx.x; k < L; k += kSupThreads) {
        s[0] += ref[3 * k];
        s[1] += ref[3 * k + 1];
        s[2] += ref[3 * k + 2];
    }
    block_sum<3>(s, lds);
    const double c0 = s[0] / L, c1 = s[1] / L, c2 = s[2] / L;
    double g[1] = {0.0};
    for (int k = threadIdx.x; k < L; k += kSupThreads) {
        const double d0 = ref[3 * k] - c0, d1 = ref[3 * k + 1] - c1, d2 = ref[3 * k + 2] - c2;
        ref_c[3 * k] = d0;
        ref_c[3 * k + 1] = d1;
        ref_c[3 * k + 2] = d2;
        g[0] += d0 * d0 + d1 * d1 + d2 * d2;
    }
    block_sum<1>(g, lds);
    if (threadIdx.x == 0) {
        info[0] = c0;
        info[1] = c1;
        info[2] = c2;
        info[3] = g[0];
    }
}

// Per conformation: centroid and the covariance of the centred conformation
// against the centred reference, cov[16 i + ...] = c(3), M(9, M[3p+q] =
// sum (x - c)_p r_q: moving onto the reference), G_x(1).
__global__ __launch_bounds__(kSupThreads) void sup_cov_kernel(const double* __restrict__ xyz, int L,
                                                              const double* __restrict__ ref_c,
                                                              double* __restrict__ cov)
{
    __shared__ double lds[kSupWaves * 10];
    const int64_t i = blockIdx.x;
    const double* x = xyz + i * 3 * (int64_t)L;
    double s[3] = {0.0, 0.0, 0.0};
    for (int k = threadIdx.x; k < L; k += kSupThreads) {
        s[0] += x[3 * k];
        s[1] += x[3 * k + 1];
        s[2] += x[3 * k + 2];
    }
    block_sum<3>(s, lds);
    const double c0 = s[0] / L, c1 = s[1] / L, c2 = s[2] / L;
    double m[10];
#pragma unroll
    for (int q = 0; q < 10; ++q) m[q] = 0.0;
    for (int k = threadIdx.x; k < L; k += kSupThreads) {
        const double p[3] = {x[3 * k] - c0, x[3 * k + 1] - c1, x[3 * k + 2] - c2};
        const double r[3] = {ref_c[3 * k], ref_c[3 * k + 1], ref_c[3 * k + 2]};
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) m[3 * a + b] += p[a] * r[b];
        m[9] += p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
    }
    block_sum<10>(m, lds);
    if (threadIdx.x < 16) {
        const int q = threadIdx.x;
        double v = 0.0;
        if (q < 3) v = q == 0 ? c0 : q == 1 ? c1 : c2;
#pragma unroll
        for (int k = 0; k < 10; ++k)
            if (q == 3 + k) v = m[k];
        cov[16 * i + q] = v;
    }
}

// One thread per conformation: Horn's quaternion (Jacobi on the key matrix),
// R, t = c_ref - R c, the fit's RMSD; rt[12 i ..] keeps R|t for the apply pass.
__global__ __launch_bounds__(64) void sup_rot_kernel(const double* __restrict__ cov, int n, int L,
                                                     const double* __restrict__ info,
                                                     double* __restrict__ rot, double* __restrict__ trans,
                                                     double* __restrict__ rmsd, double* __restrict__ rt)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* c = cov + 16 * i;
    double m[9];
    for (int q = 0; q < 9; ++q) m[q] = c[3 + q];
    double R[9], lam;
    rotation_from_cov(m, R, &lam);
    double t[3];
    for (int r = 0; r < 3; ++r) t[r] = info[r] - (R[3 * r] * c[0] + R[3 * r + 1] * c[1] + R[3 * r + 2] * c[2]);
    if (rot)
        for (int q = 0; q < 9; ++q) rot[9 * i + q] = R[q];
    if (trans)
        for (int r = 0; r < 3; ++r) trans[3 * i + r] = t[r];
    if (rmsd) {
        const double r2 = (c[12] + info[3] - 2.0 * lam) / L;
        rmsd[i] = sqrt(r2 > 0.0 ? r2 : 0.0);
    }
    if (rt) {
        for (int q = 0; q < 9; ++q) rt[12 * i + q] = R[q];
        for (int r = 0; r < 3; ++r) rt[12 * i + 9 + r] = t[r];
    }
}

// One thread per atom: out = R x + t of its conformation.
__global__ __launch_bounds__(256) void sup_apply_kernel(const double* __restrict__ xyz, int64_t n_atoms_total,
                                                        int L, const double* __restrict__ rt,
                                                        double* __restrict__ out)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n_atoms_total; k += stride) {
        const double* R = rt + 12 * (k / L);
        const double a = xyz[3 * k], b = xyz[3 * k + 1], c = xyz[3 * k + 2];
        out[3 * k] = R[0] * a + R[1] * b + R[2] * c + R[9];
        out[3 * k + 1] = R[3] * a + R[4] * b + R[5] * c + R[10];
        out[3 * k + 2] = R[6] * a + R[7] * b + R[8] * c + R[11];
    }
}

}  // namespace

size_t superpose_scratch_bytes(int n, int L) { return (size_t)L * 24 + 64 + (size_t)n * 28 * 8; }

void launch_superpose_ensemble(const double* xyz, int n, int L, const double* ref, double* scratch,
                               double* rot, double* trans, double* rmsd, double* out, hipStream_t s)
{
    double* ref_c = scratch;
    double* info = ref_c + 3 * (size_t)L;
    double* cov = info + 8;
    double* rt = cov + 16 * (size_t)n;
    sup_ref_kernel<<<1, kSupThreads, 0, s>>>(ref, L, ref_c, info);
    sup_cov_kernel<<<n, kSupThreads, 0, s>>>(xyz, L, ref_c, cov);
    sup_rot_kernel<<<(n + 63) / 64, 64, 0, s>>>(cov, n, L, info, rot, trans, rmsd, out ? rt : nullptr);
    if (out) {
        const int64_t total = (int64_t)n * L;
        const int64_t blocks = std::min<int64_t>((total + 255) / 256, 1 << 20);
        sup_apply_kernel<<<(unsigned)blocks, 256, 0, s>>>(xyz, total, L, rt, out);
    }
}

}  // namespace bs
