// kernels_plain.hip -- the plain per-pair cRMSD kernel (FP64 and FP32).
//
// Each thread owns one pair and walks the textbook steps for it: the two
// centred conformations (centroid = sum / n, then subtract, as in
// bioshell-pdb/src/calc/substructures.rs:27-42), the 3x3 cross-covariance
// M = sum a b^T with the inner products Ga, Gb, the QCP quartic solved by
// Newton from E0, and RMSD = sqrt(2 (E0 - lambda) / n).  A work-group covers
// 16 x 16 pairs; the 32 conformations it touches are centred once per group
// and streamed through LDS in chunks of 32 atoms.  This is the baseline form:
// VALU-bound, 15 FMA per atom per pair.
#include "kernels.hpp"
#include "qcp.hpp"

namespace bs {
namespace {

constexpr int kTile = 16;
constexpr int kChunk = 32;
constexpr int kRowStride = kChunk * 3 + 1;  // +1 keeps the 16 rows in distinct banks

__device__ inline int64_t tri_index(int64_t n, int64_t i, int64_t j)
{
    return i * n - i * (i + 1) / 2 + (j - i - 1);
}

// Linear block id -> (bi, bj), bi <= bj, over an nt x nt upper triangle.
__device__ inline void tri_tile(int b, int nt, int& bi, int& bj)
{
    const double twon1 = 2.0 * nt + 1.0;
    int r = (int)floor((twon1 - sqrt(twon1 * twon1 - 8.0 * b)) * 0.5);
    auto start = [nt](int row) { return row * nt - row * (row - 1) / 2; };
    while (r > 0 && start(r) > b) --r;
    while (start(r + 1) <= b) ++r;
    bi = r;
    bj = r + (b - start(r));
}

template <class T>
__global__ __launch_bounds__(256) void plain_kernel(const double* __restrict__ rows_xyz,
                                                    const double* __restrict__ cols_xyz,
                                                    int n_rows, int n_cols, int L, int triangle,
                                                    int nt, int64_t out_ld, void* __restrict__ out,
                                                    int out_f64)
{
    __shared__ T cen[2 * kTile][3];
    __shared__ T sa[kTile][kRowStride];
    __shared__ T sb[kTile][kRowStride];

    int bi, bj;
    if (triangle) {
        tri_tile(blockIdx.x, nt, bi, bj);
    } else {
        bi = blockIdx.y;
        bj = blockIdx.x;
    }
    const int tid = threadIdx.x;

    // ---- step 1: centroids of the 16 row and 16 column conformations
    {
        const int c = tid >> 3, part = tid & 7;
        const bool is_row = c < kTile;
        const int conf = is_row ? bi * kTile + c : bj * kTile + (c - kTile);
        const int nconf = is_row ? n_rows : n_cols;
        const double* src = is_row ? rows_xyz : cols_xyz;
        T s0 = T(0), s1 = T(0), s2 = T(0);
        if (conf < nconf) {
            const double* p = src + (int64_t)conf * L * 3;
            for (int k = part; k < L; k += 8) {
                s0 += T(p[3 * k + 0]);
                s1 += T(p[3 * k + 1]);
                s2 += T(p[3 * k + 2]);
            }
        }
#pragma unroll
        for (int o = 1; o < 8; o <<= 1) {
            s0 += __shfl_xor(s0, o);
            s1 += __shfl_xor(s1, o);
            s2 += __shfl_xor(s2, o);
        }
        if (part == 0) {
            cen[c][0] = s0 / T(L);
            cen[c][1] = s1 / T(L);
            cen[c][2] = s2 / T(L);
        }
    }
    __syncthreads();

    // ---- step 2: covariance of the centred sets, chunk by chunk
    const int ti = tid / kTile, tj = tid % kTile;
    T m[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) m[q] = T(0);
    T ga = T(0), gb = T(0);

    for (int k0 = 0; k0 < L; k0 += kChunk) {
        const int nk = min(kChunk, L - k0);
        for (int v = tid; v < 2 * kTile * kChunk * 3; v += 256) {
            const int half = v / (kTile * kChunk * 3);
            const int w = v - half * (kTile * kChunk * 3);
            const int r = w / (kChunk * 3), rem = w - r * (kChunk * 3);
            const int conf = (half == 0 ? bi : bj) * kTile + r;
            const int nconf = half == 0 ? n_rows : n_cols;
            const double* src = half == 0 ? rows_xyz : cols_xyz;
            T val = T(0);
            if (conf < nconf && rem < nk * 3)
                val = T(src[((int64_t)conf * L + k0) * 3 + rem]) - cen[half * kTile + r][rem % 3];
            if (half == 0)
                sa[r][rem] = val;
            else
                sb[r][rem] = val;
        }
        __syncthreads();
        for (int kk = 0; kk < nk; ++kk) {
            const T a0 = sa[ti][3 * kk], a1 = sa[ti][3 * kk + 1], a2 = sa[ti][3 * kk + 2];
            const T b0 = sb[tj][3 * kk], b1 = sb[tj][3 * kk + 1], b2 = sb[tj][3 * kk + 2];
            m[0] += a0 * b0; m[1] += a0 * b1; m[2] += a0 * b2;
            m[3] += a1 * b0; m[4] += a1 * b1; m[5] += a1 * b2;
            m[6] += a2 * b0; m[7] += a2 * b1; m[8] += a2 * b2;
            ga += a0 * a0 + a1 * a1 + a2 * a2;
            gb += b0 * b0 + b1 * b1 + b2 * b2;
        }
        __syncthreads();
    }

    // ---- steps 3-4: QCP eigenvalue and RMSD
    const int i = bi * kTile + ti, j = bj * kTile + tj;
    if (i >= n_rows || j >= n_cols) return;
    if (triangle && i >= j) return;
    const T r = qcp_rmsd<T>(m, ga, gb, L);
    const int64_t o = triangle ? tri_index(n_rows, i, j) : (int64_t)i * out_ld + j;
    if (out_f64)
        static_cast<double*>(out)[o] = double(r);
    else
        static_cast<float*>(out)[o] = float(r);
}

}  // namespace

void launch_plain_avsa(const double* xyz, int n, int L, bool fp64, void* out, bool out_f64,
                       hipStream_t s)
{
    if (n < 2) return;
    const int nt = (n + kTile - 1) / kTile;
    const int blocks = nt * (nt + 1) / 2;
    if (fp64)
        plain_kernel<double><<<blocks, 256, 0, s>>>(xyz, xyz, n, n, L, 1, nt, 0, out, out_f64);
    else
        plain_kernel<float><<<blocks, 256, 0, s>>>(xyz, xyz, n, n, L, 1, nt, 0, out, out_f64);
}

void launch_plain_qvl(const double* q_xyz, int n_q, const double* lib_xyz, int n_lib, int L,
                      bool fp64, float* out, hipStream_t s)
{
    if (n_q < 1 || n_lib < 1) return;
    dim3 grid((n_lib + kTile - 1) / kTile, (n_q + kTile - 1) / kTile);
    if (fp64)
        plain_kernel<double><<<grid, 256, 0, s>>>(q_xyz, lib_xyz, n_q, n_lib, L, 0, 0, n_lib, out,
                                                  0);
    else
        plain_kernel<float><<<grid, 256, 0, s>>>(q_xyz, lib_xyz, n_q, n_lib, L, 0, 0, n_lib, out,
                                                 0);
}

}  // namespace bs
