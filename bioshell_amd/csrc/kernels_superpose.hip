// kernels_superpose.hip -- optimal superposition of every conformation of an
// ensemble onto one reference, in FP64: rotation R and translation t with
// ref ~= R x + t (the Rototranslation convention of
// bioshell-pdb/src/calc/transformation.rs:22-37, applied as at 171-174), the
// RMSD of the fit and, optionally, the superposed coordinates R x + t.
//
// The reference is centred once (one block).  Then one launch does the rest:
// a work-group of 256 threads takes C = 256 / G conformations, G threads each
// (G = sup_lanes(L)), and every thread keeps its up to 8 atoms in registers,
// so a conformation is read from HBM once and its superposed copy written
// once (24 L bytes each way).  The sums (centroid, M = sum x r_c^T, |x|^2)
// are reduced through LDS, one thread per conformation solves the rotation
// (rotation_fast, Newton from E0 = (G_x + G_r) / 2) and the RMSD, and every
// thread writes R x + t for its atoms.  Chains longer than 2048 atoms take
// three passes (covariance per block, Jacobi per thread, apply per atom).
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

// ---------------------------------------------------------------- fused form
constexpr int kSupAtoms = 8;                 // atoms per thread held in registers
constexpr int kSupResidentMax = 256 * kSupAtoms;
constexpr int kSupValues = 13;               // sum d (3), sum d r_c^T (9), sum |d|^2; d = x - x_0

int sup_lanes(int L)
{
    int g = 16;
    while (g * kSupAtoms < L) g *= 2;
    return g;
}

struct SupLds {
    double red[256 * kSupValues + 256];  // G (13 C + 1) partials
    double part2[16 * kSupValues];
    double tot[16][kSupValues];
    double rt[16][12];  // R | t of the work-group's conformations
};

template <int G>
__global__ __launch_bounds__(256) void sup_fused_kernel(const double* __restrict__ xyz, int n, int L,
                                                        const double* __restrict__ ref_c,
                                                        const double* __restrict__ info,
                                                        double* __restrict__ rot, double* __restrict__ trans,
                                                        double* __restrict__ rmsd, double* __restrict__ out)
{
    constexpr int C = 256 / G, A = kSupAtoms, S = kSupValues * C + 1, PARTS = G / 16;
    constexpr int PAIRS = kSupValues * C;
    __shared__ SupLds lds;
    const int tid = threadIdx.x, c = tid / G, g = tid % G;
    const int64_t i = (int64_t)blockIdx.x * C + c;
    const bool live = i < n;
    const double* x = xyz + (live ? i : 0) * 3 * (int64_t)L;
    // sums of d = x - x_0 (atom 0 of the conformation): |d|^2 - L |c - x_0|^2
    // then cancels only to the extent of the chain, not its distance from the origin
    const double o0 = x[0], o1 = x[1], o2 = x[2];
    double xv[A][3];
    double v[kSupValues];
#pragma unroll
    for (int q = 0; q < kSupValues; ++q) v[q] = 0.0;
#pragma unroll
    for (int j = 0; j < A; ++j) {
        const int k = g + G * j;
        xv[j][0] = xv[j][1] = xv[j][2] = 0.0;
        if (live && k < L) {
            xv[j][0] = x[3 * k];
            xv[j][1] = x[3 * k + 1];
            xv[j][2] = x[3 * k + 2];
        }
    }
#pragma unroll
    for (int j = 0; j < A; ++j) {
        const int k = g + G * j;
        if (live && k < L) {
            const double b0 = xv[j][0] - o0, b1 = xv[j][1] - o1, b2 = xv[j][2] - o2;
            const double a0 = ref_c[3 * k], a1 = ref_c[3 * k + 1], a2 = ref_c[3 * k + 2];
            v[0] += b0; v[1] += b1; v[2] += b2;
            v[3] += b0 * a0; v[4] += b0 * a1; v[5] += b0 * a2;
            v[6] += b1 * a0; v[7] += b1 * a1; v[8] += b1 * a2;
            v[9] += b2 * a0; v[10] += b2 * a1; v[11] += b2 * a2;
            v[12] += b0 * b0 + b1 * b1 + b2 * b2;
        }
    }
    // transposed reduction: lane partials, then 16 at a time, then the G / 16 results
#pragma unroll
    for (int q = 0; q < kSupValues; ++q) lds.red[g * S + q * C + c] = v[q];
    __syncthreads();
    if (tid < PAIRS * PARTS) {
        const int pair = tid % PAIRS, part = tid / PAIRS;
        const double* r = lds.red + 16 * part * S + pair;
        double a = r[0];
#pragma unroll
        for (int u = 1; u < 16; ++u) a += r[u * S];
        if constexpr (PARTS == 1)
            lds.tot[pair % C][pair / C] = a;
        else
            lds.part2[tid] = a;
    }
    __syncthreads();
    if constexpr (PARTS > 1) {
        if (tid < PAIRS) {
            double a = lds.part2[tid];
#pragma unroll
            for (int p = 1; p < PARTS; ++p) a += lds.part2[p * PAIRS + tid];
            lds.tot[tid % C][tid / C] = a;
        }
        __syncthreads();
    }
    if (tid < C) {
        const int64_t m = (int64_t)blockIdx.x * C + tid;
        double R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, t[3] = {0, 0, 0};
        if (m < n) {
            const double* sm = lds.tot[tid];
            const double* xm = xyz + m * 3 * (int64_t)L;
            const double dc[3] = {sm[0] / L, sm[1] / L, sm[2] / L};  // c - x_0
            const double cc[3] = {xm[0] + dc[0], xm[1] + dc[1], xm[2] + dc[2]};
            double M[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) M[q] = sm[3 + q];  // r_c is centred: sum (x - c) r_c^T
            const double gx = sm[12] - L * (dc[0] * dc[0] + dc[1] * dc[1] + dc[2] * dc[2]);
            rotation_fast(M, R, 0.5 * (gx + info[3]));
#pragma unroll
            for (int r = 0; r < 3; ++r)
                t[r] = info[r] - (R[3 * r] * cc[0] + R[3 * r + 1] * cc[1] + R[3 * r + 2] * cc[2]);
            // sum r_c^T R (x - c) = sum_ab R_ab M_ba
            double lam = 0.0;
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) lam += R[3 * a + b] * M[3 * b + a];
            if (rot)
#pragma unroll
                for (int q = 0; q < 9; ++q) rot[9 * m + q] = R[q];
            if (trans)
#pragma unroll
                for (int r = 0; r < 3; ++r) trans[3 * m + r] = t[r];
            if (rmsd) {
                const double r2 = (gx + info[3] - 2.0 * lam) / L;
                rmsd[m] = sqrt(r2 > 0.0 ? r2 : 0.0);
            }
        }
#pragma unroll
        for (int q = 0; q < 9; ++q) lds.rt[tid][q] = R[q];
#pragma unroll
        for (int r = 0; r < 3; ++r) lds.rt[tid][9 + r] = t[r];
    }
    if (!out) return;  // uniform
    __syncthreads();
    const double* Rt = lds.rt[c];
#pragma unroll
    for (int j = 0; j < A; ++j) {
        const int k = g + G * j;
        if (live && k < L) {
            const double a = xv[j][0], b = xv[j][1], d = xv[j][2];
            double* o = out + (i * L + k) * 3;
            o[0] = Rt[0] * a + Rt[1] * b + Rt[2] * d + Rt[9];
            o[1] = Rt[3] * a + Rt[4] * b + Rt[5] * d + Rt[10];
            o[2] = Rt[6] * a + Rt[7] * b + Rt[8] * d + Rt[11];
        }
    }
}

template <int G>
void launch_fused(const double* xyz, int n, int L, const double* ref_c, const double* info, double* rot,
                  double* trans, double* rmsd, double* out, hipStream_t s)
{
    constexpr int C = 256 / G;
    sup_fused_kernel<G><<<(n + C - 1) / C, 256, 0, s>>>(xyz, n, L, ref_c, info, rot, trans, rmsd, out);
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
    if (L <= kSupResidentMax) {
        switch (sup_lanes(L)) {
        case 16: launch_fused<16>(xyz, n, L, ref_c, info, rot, trans, rmsd, out, s); break;
        case 32: launch_fused<32>(xyz, n, L, ref_c, info, rot, trans, rmsd, out, s); break;
        case 64: launch_fused<64>(xyz, n, L, ref_c, info, rot, trans, rmsd, out, s); break;
        case 128: launch_fused<128>(xyz, n, L, ref_c, info, rot, trans, rmsd, out, s); break;
        default: launch_fused<256>(xyz, n, L, ref_c, info, rot, trans, rmsd, out, s); break;
        }
        return;
    }
    sup_cov_kernel<<<n, kSupThreads, 0, s>>>(xyz, L, ref_c, cov);
    sup_rot_kernel<<<(n + 63) / 64, 64, 0, s>>>(cov, n, L, info, rot, trans, rmsd, out ? rt : nullptr);
    if (out) {
        const int64_t total = (int64_t)n * L;
        const int64_t blocks = std::min<int64_t>((total + 255) / 256, 1 << 20);
        sup_apply_kernel<<<(unsigned)blocks, 256, 0, s>>>(xyz, total, L, rt, out);
    }
}

}  // namespace bs
