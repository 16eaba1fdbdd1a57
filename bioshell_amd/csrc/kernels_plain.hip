// kernels_plain.hip -- the plain per-pair cRMSD kernel (FP64 and FP32).
//
// Every pair walks the textbook steps: the two centred conformations
// (centroid = sum / n, then subtract, as in
// bioshell-pdb/src/calc/substructures.rs:27-42), the 3x3 cross-covariance
// M = sum a b^T with the inner products Ga, Gb, the QCP quartic solved by
// Newton from E0, and RMSD = sqrt(2 (E0 - lambda) / n).
//
// Register-blocked for the VALU: a work-group of 256 threads covers a tile of
// 16R x 16R pairs and each thread owns R x R of them (R = 4 in FP32, 2 in
// FP64), rows ti R .. ti R + R-1 and columns tj R .. tj R + R-1 of the tile.
// The tile's 32R conformations are centred once, then streamed through LDS
// in chunks of 16 atoms, coordinate-major ([atom][coordinate][conformation]),
// so that a thread reads its R rows' and R columns' value of one coordinate
// as one R-wide vector: per atom 6 vector reads feed 9 R^2 multiply-adds
// (and 6 R for Ga, Gb).  The pairs' matrices then go through LDS in rounds,
// and one copy of the QCP solve serves every pair of a thread.
#include "kernels.hpp"
#include "qcp.hpp"

#include <cmath>

namespace bs {
namespace {

constexpr int kChunk = 16;  // atoms per LDS chunk
constexpr int kThreads = 256;

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

template <class T, int R>
struct PlainShape {
    static constexpr int kTile = 16 * R;               // conformations per tile side
    static constexpr int kConfs = 2 * kTile;           // rows then columns
    static constexpr int kLd = kConfs + 4;             // LDS row stride (floats or doubles)
    static constexpr int kPairs = R * R;               // pairs per thread
    static constexpr int kRound = R == 4 ? 4 : R == 2 ? 2 : 1;  // pairs per thread per stash round
    static constexpr int kStage = kChunk * 3 * kLd;    // staged coordinates (elements)
    static constexpr int kStash = kRound * 11 * kThreads;
    static constexpr int kLds = kStage > kStash ? kStage : kStash;
};

// Tile (bi, bj) of one ensemble.
template <class T, int R>
__device__ __forceinline__ void plain_tile(const double* __restrict__ xyz, int n, int L, int bi, int bj,
                                  void* __restrict__ out, int out_f64)
{
    using S = PlainShape<T, R>;
    __shared__ T cen[S::kConfs][3];
    __shared__ T buf[S::kLds];
    const int tid = threadIdx.x;

    // ---- step 1: centroids of the tile's conformations (sum, then divide)
    {
        constexpr int per = kThreads / S::kConfs;  // threads per conformation
        const int c = tid / per, part = tid % per;
        const int conf = c < S::kTile ? bi * S::kTile + c : bj * S::kTile + (c - S::kTile);
        T s0 = T(0), s1 = T(0), s2 = T(0);
        if (conf < n) {
            const double* p = xyz + (int64_t)conf * L * 3;
            for (int k = part; k < L; k += per) {
                s0 += T(p[3 * k + 0]);
                s1 += T(p[3 * k + 1]);
                s2 += T(p[3 * k + 2]);
            }
        }
#pragma unroll
        for (int o = 1; o < per; o <<= 1) {
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

    // ---- step 2: covariances of the thread's R x R pairs, chunk by chunk
    const int ti = tid / 16, tj = tid % 16;
    T m[R][R][9];
    T ga[R], gb[R];
#pragma unroll
    for (int u = 0; u < R; ++u) {
        ga[u] = gb[u] = T(0);
#pragma unroll
        for (int v = 0; v < R; ++v)
#pragma unroll
            for (int q = 0; q < 9; ++q) m[u][v][q] = T(0);
    }
    for (int k0 = 0; k0 < L; k0 += kChunk) {
        const int nk = min(kChunk, L - k0);
        // coalesced reads of each conformation's chunk (3 nk consecutive
        // doubles), centred, stored coordinate-major
        for (int v = tid; v < S::kConfs * kChunk * 3; v += kThreads) {
            const int c = v / (kChunk * 3), rem = v - c * (kChunk * 3);
            const int conf = c < S::kTile ? bi * S::kTile + c : bj * S::kTile + (c - S::kTile);
            T val = T(0);
            if (conf < n && rem < nk * 3)
                val = T(xyz[((int64_t)conf * L + k0) * 3 + rem]) - cen[c][rem % 3];
            buf[rem * S::kLd + c] = val;
        }
        __syncthreads();
        for (int kk = 0; kk < nk; ++kk) {
            T a[3][R], b[3][R];
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int u = 0; u < R; ++u) {
                    a[c][u] = buf[(3 * kk + c) * S::kLd + ti * R + u];
                    b[c][u] = buf[(3 * kk + c) * S::kLd + S::kTile + tj * R + u];
                }
#pragma unroll
            for (int u = 0; u < R; ++u) {
                ga[u] += a[0][u] * a[0][u] + a[1][u] * a[1][u] + a[2][u] * a[2][u];
                gb[u] += b[0][u] * b[0][u] + b[1][u] * b[1][u] + b[2][u] * b[2][u];
            }
#pragma unroll
            for (int u = 0; u < R; ++u)
#pragma unroll
                for (int v = 0; v < R; ++v)
#pragma unroll
                    for (int c = 0; c < 3; ++c)
#pragma unroll
                        for (int d = 0; d < 3; ++d) m[u][v][3 * c + d] += a[c][u] * b[d][v];
        }
        __syncthreads();
    }

    // ---- steps 3-4: QCP eigenvalue and RMSD, one pair at a time through LDS
    T* stash = buf;  // [pair slot][value][thread]
#pragma unroll
    for (int r0 = 0; r0 < S::kPairs; r0 += S::kRound) {
#pragma unroll
        for (int p = 0; p < S::kRound; ++p) {
            const int u = (r0 + p) / R, v = (r0 + p) % R;
#pragma unroll
            for (int q = 0; q < 9; ++q) stash[(p * 11 + q) * kThreads + tid] = m[u][v][q];
            stash[(p * 11 + 9) * kThreads + tid] = ga[u];
            stash[(p * 11 + 10) * kThreads + tid] = gb[v];
        }
#pragma unroll 1
        for (int p = 0; p < S::kRound; ++p) {  // each thread reads back its own slots
            const int u = (r0 + p) / R, v = (r0 + p) % R;
            const int i = bi * S::kTile + ti * R + u, j = bj * S::kTile + tj * R + v;
            T mm[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) mm[q] = stash[(p * 11 + q) * kThreads + tid];
            const T gi = stash[(p * 11 + 9) * kThreads + tid], gj = stash[(p * 11 + 10) * kThreads + tid];
            if (i < n && j < n && i < j) {
                const T r = qcp_rmsd<T>(mm, gi, gj, L);
                const int64_t o = tri_index(n, i, j);
                if (out_f64)
                    static_cast<double*>(out)[o] = double(r);
                else
                    static_cast<float*>(out)[o] = float(r);
            }
        }
    }
}

template <class T, int R>
__global__ __launch_bounds__(kThreads) void plain_kernel(const double* __restrict__ xyz, int n, int L,
                                                         int nt, void* __restrict__ out, int out_f64)
{
    int bi, bj;
    tri_tile(blockIdx.x, nt, bi, bj);
    plain_tile<T, R>(xyz, n, L, bi, bj, out, out_f64);
}

// Many ensembles in one launch: block b is tile b - prefix[e] of ensemble e,
// prefix[e] <= b < prefix[e + 1] (f32 output).
template <class T, int R>
__global__ __launch_bounds__(kThreads) void plain_batch_kernel(const PlainEns* __restrict__ ens,
                                                               const int* __restrict__ prefix, int n_ens)
{
    const int b = blockIdx.x;
    int lo = 0, hi = n_ens;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= b)
            lo = mid;
        else
            hi = mid;
    }
    const PlainEns E = ens[lo];
    int bi, bj;
    tri_tile(b - prefix[lo], E.nt, bi, bj);
    plain_tile<T, R>(E.xyz, E.n, E.L, bi, bj, E.out, 0);
}

template <class T, int R>
void launch(const double* xyz, int n, int L, void* out, bool out_f64, hipStream_t s)
{
    const int nt = plain_tiles_per_side(n, R);
    plain_kernel<T, R><<<nt * (nt + 1) / 2, kThreads, 0, s>>>(xyz, n, L, nt, out, out_f64 ? 1 : 0);
}

template <class T, int R>
void launch_batch(const PlainEns* ens, const int* prefix, int n_ens, int tiles, hipStream_t s)
{
    plain_batch_kernel<T, R><<<tiles, kThreads, 0, s>>>(ens, prefix, n_ens);
}

}  // namespace

int plain_block(int n, bool fp64, bool alone, int n_cu)
{
    // cost of one tile relative to R = 4 (fp32) or R = 2 (fp64), per pair slot:
    // smaller blocks re-read more LDS per multiply-add
    static const double cost32[3] = {3.0, 1.6, 1.0};  // R = 1, 2, 4
    static const double cost64[2] = {1.8, 1.0};       // R = 1, 2
    double best = 0.0;
    int best_r = 1;
    for (int r = 1; r <= (fp64 ? 2 : 4); r *= 2) {
        const double nt = plain_tiles_per_side(n, r);
        const double tiles = nt * (nt + 1) / 2;
        // a lone launch runs ceil(tiles / slots) rounds of 2 work-groups per CU
        const double rounds = alone ? std::ceil(tiles / (2.0 * n_cu)) * (2.0 * n_cu) : tiles;
        const double c = (fp64 ? cost64[r == 1 ? 0 : 1] : cost32[r == 1 ? 0 : r == 2 ? 1 : 2]) *
                         (16.0 * r) * (16.0 * r) * rounds;
        if (r == 1 || c < best) {
            best = c;
            best_r = r;
        }
    }
    return best_r;
}

void launch_plain_avsa(const double* xyz, int n, int L, bool fp64, void* out, bool out_f64,
                       int n_cu, hipStream_t s)
{
    if (n < 2) return;
    const int r = plain_block(n, fp64, true, n_cu);
    if (fp64) {
        if (r == 2)
            launch<double, 2>(xyz, n, L, out, out_f64, s);
        else
            launch<double, 1>(xyz, n, L, out, out_f64, s);
    } else {
        if (r == 4)
            launch<float, 4>(xyz, n, L, out, out_f64, s);
        else if (r == 2)
            launch<float, 2>(xyz, n, L, out, out_f64, s);
        else
            launch<float, 1>(xyz, n, L, out, out_f64, s);
    }
}

void launch_plain_batch(const PlainEns* ens, const int* prefix, int n_ens, int tiles, int r, bool fp64,
                        hipStream_t s)
{
    if (n_ens == 0 || tiles == 0) return;
    if (fp64) {
        if (r == 2)
            launch_batch<double, 2>(ens, prefix, n_ens, tiles, s);
        else
            launch_batch<double, 1>(ens, prefix, n_ens, tiles, s);
    } else {
        if (r == 4)
            launch_batch<float, 4>(ens, prefix, n_ens, tiles, s);
        else if (r == 2)
            launch_batch<float, 2>(ens, prefix, n_ens, tiles, s);
        else
            launch_batch<float, 1>(ens, prefix, n_ens, tiles, s);
    }
}

}  // namespace bs
