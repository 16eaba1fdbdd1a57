// kernels_cluster.hip -- reductions of an RMSD matrix that BioShell's
// clustering tools need, on the device: row minima and maxima (outliers,
// min-max medoid) and the agglomeration loop of hierarchical clustering.
//
// Row statistics read the packed upper triangle once, in 64 x 64 tiles: each
// value is the (i, j) entry of row i and the (j, i) entry of row j, so a tile
// reduces its rows and its columns and merges both into the global minima and
// maxima with integer atomics on the float bit patterns (distances are >= 0,
// whose IEEE patterns order like unsigned integers).
//
// The agglomeration follows bioshell-clustering/src/hierarchical/
// clustering_matrix.rs step for step on a dense f32 matrix in HBM (see the
// comments at each phase); the merge sequence it records is what the host
// rebuilds the tree from.
#include "kernels.hpp"

#include <algorithm>

namespace bs {
namespace {

__device__ inline unsigned fbits(float v) { return __float_as_uint(v); }

__device__ inline int64_t tri(int64_t n, int64_t i, int64_t j)  // i < j
{
    return i * n - i * (i + 1) / 2 + (j - i - 1);
}

// ------------------------------------------------------------------ row stats
__global__ void fill_u32_kernel(unsigned* p, unsigned v, int64_t n)
{
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (int64_t)gridDim.x * 256)
        p[k] = v;
}

// Block (tr, tc), tr <= tc: rows 64 tr.., columns 64 tc..; thread t takes row
// t / 4 and 16 columns from 16 (t % 4).
__global__ __launch_bounds__(256) void row_stats_kernel(const float* __restrict__ up, int n,
                                                        unsigned* __restrict__ rmin,
                                                        unsigned* __restrict__ rmax)
{
    const int tr = blockIdx.y, tc = blockIdx.x;
    if (tr > tc) return;
    __shared__ unsigned cmin[64], cmax[64], wmin[64], wmax[64];
    const int t = threadIdx.x;
    if (t < 64) {
        cmin[t] = 0x7F800000u;  // +inf
        cmax[t] = 0u;
        wmin[t] = 0x7F800000u;
        wmax[t] = 0u;
    }
    __syncthreads();
    const int r = t >> 2, c0 = (t & 3) * 16;
    const int i = tr * 64 + r;
    unsigned lo = 0x7F800000u, hi = 0u;
    if (i < n) {
        for (int c = c0; c < c0 + 16; ++c) {
            const int j = tc * 64 + c;
            if (j <= i || j >= n) continue;
            const unsigned v = fbits(up[tri(n, i, j)]);
            lo = min(lo, v);
            hi = max(hi, v);
            atomicMin(&cmin[c], v);
            atomicMax(&cmax[c], v);
        }
    }
    atomicMin(&wmin[r], lo);
    atomicMax(&wmax[r], hi);
    __syncthreads();
    if (t < 64) {
        const int ri = tr * 64 + t, cj = tc * 64 + t;
        if (ri < n && wmin[t] != 0x7F800000u) {
            atomicMin(rmin + ri, wmin[t]);
            atomicMax(rmax + ri, wmax[t]);
        }
        if (cj < n && cmin[t] != 0x7F800000u) {
            atomicMin(rmin + cj, cmin[t]);
            atomicMax(rmax + cj, cmax[t]);
        }
    }
}

// Member a of the list: the largest distance to the other members (gather).
__global__ __launch_bounds__(256) void members_row_max_kernel(const float* __restrict__ up, int n,
                                                              const unsigned* __restrict__ mem,
                                                              int m, float* __restrict__ out)
{
    __shared__ float red[256];
    const int a = blockIdx.x;
    const int ia = (int)mem[a];
    float hi = -3.402823466e+38f;  // -f32::MAX
    for (int b = threadIdx.x; b < m; b += 256) {
        const int ib = (int)mem[b];
        if (b == a) continue;
        const float d = ia == ib ? 0.0f : (ia < ib ? up[tri(n, ia, ib)] : up[tri(n, ib, ia)]);
        hi = fmaxf(hi, d);
    }
    red[threadIdx.x] = hi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[a] = red[0];
}

// ------------------------------------------------------------------ clustering
struct Cand {
    float v;
    int j, i;
};

// (v, j, i) lexicographic: the first strict minimum of the reference's scan
// (j ascending, then i ascending; clustering_matrix.rs:27-42).
__device__ inline bool better(const Cand& a, const Cand& b)
{
    if (a.v != b.v) return a.v < b.v;
    if (a.j != b.j) return a.j < b.j;
    return a.i < b.i;
}

__global__ void dense_from_upper_kernel(const float* __restrict__ up, int n, float* __restrict__ D,
                                        unsigned* __restrict__ sizes, int* __restrict__ state)
{
    const int64_t nn = (int64_t)n * n;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < nn; k += (int64_t)gridDim.x * 256) {
        const int r = (int)(k / n), c = (int)(k - (int64_t)r * n);
        D[k] = r == c ? 0.0f : (r < c ? up[tri(n, r, c)] : up[tri(n, c, r)]);
    }
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (int64_t)gridDim.x * 256)
        sizes[k] = 1u;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        state[0] = n;  // order
        state[1] = 0;  // step
    }
}

// Closest pair, stage 1: the matrix stays bitwise symmetric on the active
// part, so column j above the diagonal equals the row prefix D[j][0..j):
// wave w scans rows j = 1 + w, 1 + w + W, ...; the block keeps its best.
__global__ __launch_bounds__(256) void closest_partial_kernel(const float* __restrict__ D, int n,
                                                              const int* __restrict__ state,
                                                              Cand* __restrict__ part)
{
    __shared__ Cand wbest[4];
    const int order = state[0];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int W = gridDim.x * 4;
    Cand best{3.402823466e+38f, 0x7FFFFFFF, 0x7FFFFFFF};
    for (int j = 1 + blockIdx.x * 4 + w; j < order; j += W) {
        const float* row = D + (int64_t)j * n;
        float v = 3.402823466e+38f;
        int bi = 0x7FFFFFFF;
        for (int i = lane; i < j; i += 64) {
            const float x = row[i];
            if (x < v) {
                v = x;
                bi = i;
            }
        }
        // lanes: smallest value, then smallest i
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const float ov = __shfl_xor(v, o);
            const int oi = __shfl_xor(bi, o);
            if (ov < v || (ov == v && oi < bi)) {
                v = ov;
                bi = oi;
            }
        }
        const Cand c{v, j, bi};
        if (bi != 0x7FFFFFFF && better(c, best)) best = c;
    }
    if (lane == 0) wbest[w] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        Cand b = wbest[0];
        for (int k = 1; k < 4; ++k)
            if (better(wbest[k], b)) b = wbest[k];
        part[blockIdx.x] = b;
    }
}

// The merging rules of strategies/mod.rs in f32, evaluated left to right with
// no contraction into fused multiply-adds (every operation rounds, as in the
// reference); the division is the correctly rounded f32 division.
__device__ inline float merge_rule(int rule, unsigned si, unsigned sj, unsigned sk, float dij,
                                   float dik, float djk)
{
#pragma clang fp contract(off)
    switch (rule) {
    case 0: return fminf(dik, djk);
    case 1: return fmaxf(dik, djk);
    case 2: {
        const float d = 1.0f / (float)(si + sj);
        return d * (float)si * dik + d * (float)sj * djk;
    }
    case 3: return 0.5f * dik + 0.5f * djk - 0.25f * dij;
    case 4: {
        const float d = 1.0f / (float)(si + sj);
        return d * (float)si * dik + d * (float)sj * djk - (float)si * (float)sj * d * d * dij;
    }
    default: {
        const float d = 1.0f / (float)(si + sj + sk);
        return d * (float)(si + sk) * dik + d * (float)(sj + sk) * djk - (float)sk * d * dij;
    }
    }
}

// Closest pair, stage 2, and the merge -- one block:
//   the best of the partial candidates is the pair (i, j), i < j;
//   update_distances(i, j, rule, i): rule(size_i, size_j, sizes[j], d_ij,
//     D[i][k], D[j][k]) for every active k, computed before any is written,
//     then written to row and column i; D[i][i] = 0; sizes[i] = size_j + size_i
//     (clustering_matrix.rs:57-74, hierarchical.rs:66);
//   replace_with_last(j) if j is not the last active index: row j takes the
//     last row, then column j the last column (clustering_matrix.rs:47-55);
//     sizes[] is not moved, as in the reference; the order drops by one.
__global__ __launch_bounds__(1024) void merge_step_kernel(float* __restrict__ D, int n,
                                                          unsigned* __restrict__ sizes,
                                                          int* __restrict__ state,
                                                          const Cand* __restrict__ part, int n_part,
                                                          int rule, float* __restrict__ res,
                                                          unsigned* __restrict__ mi,
                                                          unsigned* __restrict__ mj,
                                                          float* __restrict__ md)
{
    __shared__ Cand red[1024];
    __shared__ int s_i, s_j;
    __shared__ float s_d;
    const int t = threadIdx.x;
    const int order = state[0];
    if (order < 2) return;  // every block sees the same order
    Cand b{3.402823466e+38f, 0x7FFFFFFF, 0x7FFFFFFF};
    for (int k = t; k < n_part; k += 1024)
        if (better(part[k], b)) b = part[k];
    red[t] = b;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (t < s && better(red[t + s], red[t])) red[t] = red[t + s];
        __syncthreads();
    }
    if (t == 0) {
        Cand c = red[0];
        if (c.j == 0x7FFFFFFF) {  // every value >= f32::MAX: the reference returns (0, 0)
            c.i = 0;
            c.j = 0;
        }
        s_i = c.i;
        s_j = c.j;
        s_d = D[(int64_t)c.i * n + c.j];
        const int step = state[1];
        mi[step] = (unsigned)c.i;
        mj[step] = (unsigned)c.j;
        md[step] = s_d;
    }
    __syncthreads();
    const int i = s_i, j = s_j;
    const float dij = s_d;
    const unsigned si = sizes[i], sj = sizes[j];
    const float* ri = D + (int64_t)i * n;
    const float* rj = D + (int64_t)j * n;
    for (int k = t; k < order; k += 1024) res[k] = merge_rule(rule, si, sj, sj, dij, ri[k], rj[k]);
    __syncthreads();
    for (int k = t; k < order; k += 1024) {
        D[(int64_t)i * n + k] = res[k];
        D[(int64_t)k * n + i] = res[k];
    }
    __syncthreads();
    if (t == 0) {
        D[(int64_t)i * n + i] = 0.0f;
        sizes[i] = sj + si;
    }
    __syncthreads();
    const int last = order - 1;
    if (j < last) {
        for (int c = t; c <= last; c += 1024) D[(int64_t)j * n + c] = D[(int64_t)last * n + c];
        __syncthreads();
        for (int r = t; r < last; r += 1024) D[(int64_t)r * n + j] = D[(int64_t)r * n + last];
        __syncthreads();
    }
    if (t == 0) {
        state[0] = last;
        state[1] += 1;
    }
}

}  // namespace

void launch_row_stats(const float* up, int n, float* rmin, float* rmax, hipStream_t s)
{
    const int64_t nb = std::max<int64_t>(1, std::min<int64_t>(1024, (n + 255) / 256));
    fill_u32_kernel<<<(int)nb, 256, 0, s>>>(reinterpret_cast<unsigned*>(rmin), 0x7F800000u, n);
    fill_u32_kernel<<<(int)nb, 256, 0, s>>>(reinterpret_cast<unsigned*>(rmax), 0u, n);
    const int T = (n + 63) / 64;
    if (n >= 2)
        row_stats_kernel<<<dim3(T, T), 256, 0, s>>>(up, n, reinterpret_cast<unsigned*>(rmin),
                                                   reinterpret_cast<unsigned*>(rmax));
}

void launch_members_row_max(const float* up, int n, const unsigned* members, int m, float* out,
                            hipStream_t s)
{
    members_row_max_kernel<<<m, 256, 0, s>>>(up, n, members, m, out);
}

void launch_cluster_init(const float* up, int n, float* D, unsigned* sizes, int* state,
                         hipStream_t s)
{
    const int64_t nb = std::max<int64_t>(1, std::min<int64_t>(4096, ((int64_t)n * n + 255) / 256));
    dense_from_upper_kernel<<<(int)nb, 256, 0, s>>>(up, n, D, sizes, state);
}

void launch_cluster_step(float* D, int n, unsigned* sizes, int* state, void* part, int n_part,
                         int rule, float* res, unsigned* mi, unsigned* mj, float* md, hipStream_t s)
{
    closest_partial_kernel<<<n_part, 256, 0, s>>>(D, n, state, static_cast<Cand*>(part));
    merge_step_kernel<<<1, 1024, 0, s>>>(D, n, sizes, state, static_cast<const Cand*>(part), n_part,
                                         rule, res, mi, mj, md);
}

size_t cluster_candidate_bytes() { return sizeof(Cand); }

}  // namespace bs
