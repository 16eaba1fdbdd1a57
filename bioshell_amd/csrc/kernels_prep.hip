// kernels_prep.hip -- per-conformation preparation for the Gram path, FP64.
//
//   centre   x_i = a_i - (sum a_i)/n                   (substructures.rs:27-42)
//   fit      rotation U_i that superposes x_i onto a reference r (Horn's
//            quaternion: top eigenvector of the 4x4 key matrix, Jacobi)
//   mean     r <- mean_i U_i x_i   (one Procrustes-mean step; r stays centred)
//   pack     e_i = U_i x_i - r split into MFMA operand panels, and the FP64
//            per-conformation constants of the shifted QCP frame (qcp.hpp)
//
// All of it is O(n L) and runs once per ensemble; none of it is per pair.
#include "kernels.hpp"
#include "qcp.hpp"

namespace bs {
namespace {

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// ------------------------------------------------------------------ centre
__global__ __launch_bounds__(256) void centre_kernel(const double* __restrict__ xyz, int n, int L,
                                                     double* __restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    const double* p = xyz + (int64_t)i * L * 3;
    double s0 = 0, s1 = 0, s2 = 0;
    for (int k = lane; k < L; k += 64) {
        s0 += p[3 * k];
        s1 += p[3 * k + 1];
        s2 += p[3 * k + 2];
    }
    s0 = wave_sum(s0) / L;
    s1 = wave_sum(s1) / L;
    s2 = wave_sum(s2) / L;
    double* q = out + (int64_t)i * L * 3;
    for (int k = lane; k < L; k += 64) {
        q[3 * k] = p[3 * k] - s0;
        q[3 * k + 1] = p[3 * k + 1] - s1;
        q[3 * k + 2] = p[3 * k + 2] - s2;
    }
}

// Rotation that moves b onto a, from M = sum b a^T (row-major).
__device__ void rotation_from_cov(const double* M, double* R, double* lambda)
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

// ------------------------------------------------------------------- fit
__global__ __launch_bounds__(256) void fit_kernel(const double* __restrict__ centred, int n, int L,
                                                  const double* __restrict__ ref,
                                                  double* __restrict__ rot)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    const double* x = centred + (int64_t)i * L * 3;
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
    if (lane == 0) {
        double R[9];
        rotation_from_cov(m, R, nullptr);
#pragma unroll
        for (int q = 0; q < 9; ++q) rot[(int64_t)i * 9 + q] = R[q];
    }
}

// ------------------------------------------------------------------ mean
constexpr int kMeanSlice = 64;

__global__ __launch_bounds__(256) void mean_partial_kernel(const double* __restrict__ centred,
                                                           const double* __restrict__ rot, int n,
                                                           int L, double* __restrict__ partial)
{
    __shared__ double sR[kMeanSlice][9];
    const int s = blockIdx.x;
    const int i0 = s * kMeanSlice;
    const int cnt = min(kMeanSlice, n - i0);
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

__global__ __launch_bounds__(256) void mean_final_kernel(const double* __restrict__ partial,
                                                         int n_slices, int n, int L,
                                                         double* __restrict__ ref)
{
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v >= 3 * L) return;
    double acc = 0.0;
    for (int s = 0; s < n_slices; ++s) acc += partial[(int64_t)s * 3 * L + v];
    ref[v] = acc / n;
}

// R0 = sum_k r_k r_k^T, stored as (xx, yy, zz, xy, xz, yz) after the 3L values.
__global__ __launch_bounds__(256) void ref_gram_kernel(double* __restrict__ ref, int L)
{
    __shared__ double part[4][6];
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

__global__ void copy_kernel(const double* __restrict__ src, double* __restrict__ dst, int count)
{
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v < count) dst[v] = src[v];
}

// ------------------------------------------------------------------ pack
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

template <int MODE>
__global__ __launch_bounds__(256) void pack_kernel(const double* __restrict__ centred,
                                                   const double* __restrict__ rot,
                                                   const double* __restrict__ ref, int n,
                                                   int n_pad, int L, int n_kblk,
                                                   uint4* __restrict__ panels,
                                                   double* __restrict__ consts, int ld)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n_pad) return;
    const bool live = i < n;
    double R[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) R[q] = live ? rot[(int64_t)i * 9 + q] : 0.0;
    const double* x = centred + (int64_t)(live ? i : 0) * L * 3;
    const int cb = i / kConfBlock, crow = i % kConfBlock;

    double g = 0.0;
    double C[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // sum r e~^T
    for (int grp = lane; grp < 2 * n_kblk; grp += 64) {
        const int kb = grp >> 1, h = grp & 1;
        uint32_t words[6][4];  // 6 uint4 outputs
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int k = kb * kAtomBlock + (MODE == kGramBF16x3 ? 8 * h + e : h + 2 * e);
            double ev[3] = {0.0, 0.0, 0.0};
            double rv[3] = {0.0, 0.0, 0.0};
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
                if (MODE == kGramBF16x3) {
                    const uint16_t hi = bf16_rne((float)ev[c]);
                    const double dhi = bf16_to_double(hi);
                    const uint16_t lo = bf16_rne((float)(ev[c] - dhi));
                    et[c] = dhi + bf16_to_double(lo);
                    // planes 2c (hi) and 2c+1 (lo); element e of the lane's 8
                    const int w = e >> 1, sh = (e & 1) * 16;
                    if ((e & 1) == 0) {
                        words[2 * c][w] = hi;
                        words[2 * c + 1][w] = lo;
                    } else {
                        words[2 * c][w] |= (uint32_t)hi << sh;
                        words[2 * c + 1][w] |= (uint32_t)lo << sh;
                    }
                } else {
                    const float f = (float)ev[c];
                    et[c] = (double)f;
                    // plane c holds 8 floats per lane = two uint4: words[2c], words[2c+1]
                    words[2 * c + (e >> 2)][e & 3] = __float_as_uint(f);
                }
            }
            g += et[0] * et[0] + et[1] * et[1] + et[2] * et[2];
#pragma unroll
            for (int p = 0; p < 3; ++p)
#pragma unroll
                for (int q = 0; q < 3; ++q) C[3 * p + q] += rv[p] * et[q];
        }
        const int plane_lane = crow + 32 * h;
        uint4* blk = panels + ((int64_t)cb * n_kblk + kb) * kPanelBlockU4;
#pragma unroll
        for (int t = 0; t < 6; ++t) {
            uint4 v = make_uint4(words[t][0], words[t][1], words[t][2], words[t][3]);
            if (MODE == kGramBF16x3)
                blk[t * 64 + plane_lane] = v;                            // plane t
            else
                blk[(t >> 1) * 128 + plane_lane * 2 + (t & 1)] = v;      // plane t/2, half t&1
        }
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

// ------------------------------------------------------ single superposition
__global__ __launch_bounds__(64) void superpose_kernel(const double* __restrict__ a,
                                                       const double* __restrict__ b, int L,
                                                       double* __restrict__ out)
{
    const int lane = threadIdx.x;
    double sa[3] = {0, 0, 0}, sb[3] = {0, 0, 0};
    for (int k = lane; k < L; k += 64)
        for (int c = 0; c < 3; ++c) {
            sa[c] += a[3 * k + c];
            sb[c] += b[3 * k + c];
        }
    for (int c = 0; c < 3; ++c) {
        sa[c] = wave_sum(sa[c]) / L;
        sb[c] = wave_sum(sb[c]) / L;
    }
    double m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double ga = 0, gb = 0;
    for (int k = lane; k < L; k += 64) {
        double pa[3], pb[3];
        for (int c = 0; c < 3; ++c) {
            pa[c] = a[3 * k + c] - sa[c];
            pb[c] = b[3 * k + c] - sb[c];
        }
        for (int p = 0; p < 3; ++p)
            for (int q = 0; q < 3; ++q) m[3 * p + q] += pb[p] * pa[q];
        ga += pa[0] * pa[0] + pa[1] * pa[1] + pa[2] * pa[2];
        gb += pb[0] * pb[0] + pb[1] * pb[1] + pb[2] * pb[2];
    }
    for (int q = 0; q < 9; ++q) m[q] = wave_sum(m[q]);
    ga = wave_sum(ga);
    gb = wave_sum(gb);
    if (lane == 0) {
        double R[9], lam;
        rotation_from_cov(m, R, &lam);
        for (int q = 0; q < 9; ++q) out[q] = R[q];
        for (int r = 0; r < 3; ++r)
            out[9 + r] = sa[r] - (R[3 * r] * sb[0] + R[3 * r + 1] * sb[1] + R[3 * r + 2] * sb[2]);
        const double r2 = 2.0 * (0.5 * (ga + gb) - lam) / L;
        out[12] = sqrt(r2 > 0.0 ? r2 : 0.0);
    }
}

}  // namespace

void launch_centre(const double* xyz, int n, int L, double* centred, hipStream_t s)
{
    centre_kernel<<<(n + 3) / 4, 256, 0, s>>>(xyz, n, L, centred);
}

void launch_fit(const double* centred, int n, int L, const double* ref, double* rot, hipStream_t s)
{
    fit_kernel<<<(n + 3) / 4, 256, 0, s>>>(centred, n, L, ref, rot);
}

void launch_mean(const double* centred, const double* rot, int n, int L, double* partial,
                 double* ref_out, hipStream_t s)
{
    const int slices = (n + kMeanSlice - 1) / kMeanSlice;
    mean_partial_kernel<<<slices, 256, 0, s>>>(centred, rot, n, L, partial);
    mean_final_kernel<<<(3 * L + 255) / 256, 256, 0, s>>>(partial, slices, n, L, ref_out);
    ref_gram_kernel<<<1, 256, 0, s>>>(ref_out, L);
}

void launch_ref_from_first(const double* centred, int L, double* ref, hipStream_t s)
{
    copy_kernel<<<(3 * L + 255) / 256, 256, 0, s>>>(centred, ref, 3 * L);
    ref_gram_kernel<<<1, 256, 0, s>>>(ref, L);
}

void launch_pack(const double* centred, const double* rot, const double* ref, int n, int n_pad,
                 int L, int n_kblk, GramMode mode, void* panels, double* consts, int ld,
                 hipStream_t s)
{
    const int blocks = (n_pad + 3) / 4;
    if (mode == kGramBF16x3)
        pack_kernel<kGramBF16x3><<<blocks, 256, 0, s>>>(centred, rot, ref, n, n_pad, L, n_kblk,
                                                        static_cast<uint4*>(panels), consts, ld);
    else
        pack_kernel<kGramFP32><<<blocks, 256, 0, s>>>(centred, rot, ref, n, n_pad, L, n_kblk,
                                                      static_cast<uint4*>(panels), consts, ld);
}

void launch_superpose(const double* a, const double* b, int L, double* out13, hipStream_t s)
{
    superpose_kernel<<<1, 64, 0, s>>>(a, b, L, out13);
}

}  // namespace bs
