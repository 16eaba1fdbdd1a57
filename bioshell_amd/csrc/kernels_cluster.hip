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
#include "wave.hpp"

#include <algorithm>
#include <cstdlib>

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

// Row statistics compare f32 values through order-preserving unsigned keys
// (fkey in kernels.hpp); NaN entries are skipped, as the reference's
// `d < min` / `d > max` tests skip them.

// First block of tile row r in the row-by-row order of the T x T upper
// triangle of tiles (tr <= tc).
__host__ __device__ inline int tile_row_first(int r, int T) { return r * T - r * (r - 1) / 2; }

// Block b covers tile base + b = (tr, tc), tr <= tc: rows 64 tr.., columns
// 64 tc.. of the packed triangle.  Wave w takes rows 64 tr + 16 w .. + 15; lane l holds
// column 64 tc + l, so every row's 64 values are one 256-byte read, and the
// wave issues its 16 reads before using any.  Columns fold in each lane's
// registers; rows fold through an LDS copy of the tile's keys, four threads
// per row.  The tile's results go to
// part[256 b ..]: row minima, row maxima, column minima, column maxima (64
// keys each; a row or column with no entry here holds the identity).  base
// is the first tile of the chunk of tile rows this launch covers.
__global__ __launch_bounds__(256) void row_stats_tile_kernel(const float* __restrict__ up, int n, int T,
                                                             int base, unsigned* __restrict__ part)
{
    const int b = base + (int)blockIdx.x;  // tile index in the whole triangle of tiles
    int tr = (int)floor(((2.0 * T + 1.0) - sqrt((2.0 * T + 1.0) * (2.0 * T + 1.0) - 8.0 * b)) * 0.5);
    while (tr > 0 && tile_row_first(tr, T) > b) --tr;
    while (tr + 1 < T && tile_row_first(tr + 1, T) <= b) ++tr;
    const int tc = tr + (b - tile_row_first(tr, T));
    __shared__ unsigned cmin[4][64], cmax[4][64];
    __shared__ unsigned tmin[64][65], tmax[64][65];  // the tile's keys, [row][column]
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int j = tc * 64 + lane;
    const int i0 = tr * 64 + w * 16;
    unsigned* out = part + (size_t)blockIdx.x * 256;
    float d[16];
#pragma unroll
    for (int rr = 0; rr < 16; ++rr) {
        const int i = i0 + rr;
        d[rr] = (i < n && j > i && j < n) ? up[tri(n, i, j)] : 0.0f;
    }
    // columns: each lane folds its 16 rows; rows: through LDS below
    unsigned clo = 0xFFFFFFFFu, chi = 0u;
#pragma unroll
    for (int rr = 0; rr < 16; ++rr) {
        const int i = i0 + rr;
        const bool ok = i < n && j > i && j < n && d[rr] == d[rr];
        const unsigned k = fkey(fbits(d[rr]));
        const unsigned klo = ok ? k : 0xFFFFFFFFu, khi = ok ? k : 0u;
        clo = min(clo, klo);
        chi = max(chi, khi);
        tmin[w * 16 + rr][lane] = klo;
        tmax[w * 16 + rr][lane] = khi;
    }
    cmin[w][lane] = clo;
    cmax[w][lane] = chi;
    __syncthreads();
    {
        // row r of the tile: four threads of 16 columns each, then two shuffles
        const int r = threadIdx.x >> 2, q = threadIdx.x & 3;
        unsigned lo = 0xFFFFFFFFu, hi = 0u;
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            lo = min(lo, tmin[r][q * 16 + c]);
            hi = max(hi, tmax[r][q * 16 + c]);
        }
        lo = min(lo, (unsigned)__shfl_xor((int)lo, 1));
        hi = max(hi, (unsigned)__shfl_xor((int)hi, 1));
        lo = min(lo, (unsigned)__shfl_xor((int)lo, 2));
        hi = max(hi, (unsigned)__shfl_xor((int)hi, 2));
        if (q == 0) {
            out[r] = lo;
            out[64 + r] = hi;
        }
    }
    if (w == 0) {
        out[128 + lane] = min(min(cmin[0][lane], cmin[1][lane]), min(cmin[2][lane], cmin[3][lane]));
        out[192 + lane] = max(max(cmax[0][lane], cmax[1][lane]), max(cmax[2][lane], cmax[3][lane]));
    }
}

// Block t: elements e = 64 t + lane, for the chunk of tile rows [t0, t1)
// whose tile results are in part (tile index minus base).  Element e's row
// part is in tiles (t, tc >= t), its column part in tiles (tr <= t, t); wave q
// takes every eighth of those in the chunk, and the eight waves are combined
// in LDS with the element's keys so far: the keys of f32::MAX and -f32::MAX
// for the first chunk, else the values rmin / rmax hold.  The last chunk
// decodes the keys to f32.
__global__ __launch_bounds__(512) void row_stats_reduce_kernel(const unsigned* __restrict__ part, int n, int T,
                                                               int t0, int t1, int base, int first, int last,
                                                               unsigned* __restrict__ rmin,
                                                               unsigned* __restrict__ rmax)
{
    __shared__ unsigned slo[8][64], shi[8][64];
    const int t = blockIdx.x, r = threadIdx.x & 63, q = threadIdx.x >> 6;
    unsigned lo = 0xFFFFFFFFu, hi = 0u;
    if (n >= 2 && t >= t0) {
        if (t < t1) {
            const int f = tile_row_first(t, T) - base;
            for (int tc = t + q; tc < T; tc += 8) {
                const unsigned* p = part + (size_t)(f + tc - t) * 256;
                lo = min(lo, p[r]);
                hi = max(hi, p[64 + r]);
            }
        }
        for (int tr = t0 + q; tr < t1 && tr <= t; tr += 8) {
            const unsigned* p = part + (size_t)(tile_row_first(tr, T) - base + t - tr) * 256;
            lo = min(lo, p[128 + r]);
            hi = max(hi, p[192 + r]);
        }
    }
    slo[q][r] = lo;
    shi[q][r] = hi;
    __syncthreads();
    const int e = t * 64 + r;
    if (q == 0 && e < n) {
        lo = first ? kKeyFltMax : rmin[e];
        hi = first ? kKeyNegFltMax : rmax[e];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            lo = min(lo, slo[k][r]);
            hi = max(hi, shi[k][r]);
        }
        rmin[e] = last ? fkey_inverse(lo) : lo;
        rmax[e] = last ? fkey_inverse(hi) : hi;
    }
}

// Keys of elements [0, n) to f32 values, in place.
__global__ void row_stats_decode_kernel(unsigned* __restrict__ kmin, unsigned* __restrict__ kmax, int n)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    kmin[e] = fkey_inverse(kmin[e]);
    kmax[e] = fkey_inverse(kmax[e]);
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
// A candidate pair (v = D[j][i], i < j) as one unsigned key whose order is (v, j, i)
// lexicographic, the first strict minimum of the reference's scan (j
// ascending, then i ascending; clustering_matrix.rs:27-42): the
// order-preserving bits of v (-0 taken as +0, which compares equal to it),
// then j, then i, each index in 16 bits (n <= 65536).  No candidate is the
// largest key, ~0.  NaN never becomes a candidate: a scan takes only values
// below f32::MAX.
__device__ inline uint64_t cand_key(float v, int j, int i)
{
    v = v == 0.0f ? 0.0f : v;
    return ((uint64_t)fkey(__float_as_uint(v)) << 32) | ((uint64_t)(unsigned)j << 16) | (unsigned)i;
}

__global__ void dense_from_upper_kernel(const float* __restrict__ up, int n, int ld, float* __restrict__ D,
                                        unsigned* __restrict__ sizes)
{
    const int64_t nn = (int64_t)n * ld;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < nn; k += (int64_t)gridDim.x * 256) {
        const int r = (int)(k / ld), c = (int)(k - (int64_t)r * ld);
        D[k] = c >= n ? 0.0f : (r == c ? 0.0f : (r < c ? up[tri(n, r, c)] : up[tri(n, c, r)]));
    }
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (int64_t)gridDim.x * 256)
        sizes[k] = 1u;
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

// The decision of one merge, written by the last work-group of a launch and
// applied by every work-group of the next.
struct MergeState {
    int s;             // the next merge to find (0 .. n - 2); n - 1 when done
    int i, j;          // the pair, i < j (or the reference's (0, 0))
    int order;         // active elements before the merge
    unsigned si, sj;   // sizes[i], sizes[j] before the merge
    float dij;
};

// Merges are applied and found by two launches per merge:
// cluster_decide_kernel (one work-group) takes the best of the work-groups'
// candidates for merge s and applies it to rows i and j; the next
// cluster_step_kernel (every work-group) applies it to every other row and
// finds each work-group's best candidate for merge s + 1.
//   update_distances(i, j, rule, i): res[k] = rule(size_i, size_j, sizes[j],
//     d_ij, D[i][k], D[j][k]), written to row and column i, D[i][i] = 0
//     (clustering_matrix.rs:57-74);
//   replace_with_last(j) when j is not the last active index: row and column
//     j take the last ones (clustering_matrix.rs:47-55); sizes[] not moved;
//   closest pair: the first strict minimum of the scan j ascending, i
//     ascending (clustering_matrix.rs:27-42), i.e. the smallest (value, j, i)
//     over values below f32::MAX, else (0, 0).
// Each wave of the step kernel owns rows r = W k + wave id.  In a row other
// than i and j the merge changes two entries only, D[r][i] and D[r][j], and by
// symmetry their new values need only the row's own old D[r][i], D[r][j] and
// D[r][last]; the wave scans the row's prefix D[r][0..r) (= column r above the
// diagonal) four floats per load with those two columns skipped, takes their
// new values as candidates, and stores them afterwards.  No wave reads
// another's rows, so the kernel needs no barrier between work-groups, and no
// store has to land before a scan.  Rows i and j, rewritten whole by the
// decide kernel, are scanned as they are.
__global__ __launch_bounds__(256) void cluster_step_kernel(float* __restrict__ D, int n, int ld,
                                                           const MergeState* __restrict__ st,
                                                           uint64_t* __restrict__ part, int rule)
{
    __shared__ uint64_t wbest[4];
    const int s = st->s;
    if (s >= n - 1) return;  // every merge found: a spare step of the graph
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int W = gridDim.x * 4, wid = blockIdx.x * 4 + w;
    const int order = n - s;  // active elements after the previous merge
    int mi_ = -1, mj_ = -1;
    bool moved = false;
    unsigned msi = 0, msj = 0;
    float mdij = 0.0f;
    if (s > 0) {
        mi_ = st->i;
        mj_ = st->j;
        moved = mj_ < order;  // j below the previous order's last index (= order)
        msi = st->si;
        msj = st->sj;
        mdij = st->dij;
    }
    const float kFltMax = 3.402823466e+38f, kInf = __builtin_inff();
    uint64_t best = ~0ull;  // cand_key order: (value, j = row, i = column)
    for (int r = wid; r < order; r += W) {
        float* row = D + (int64_t)r * ld;
        const bool patch = s > 0 && r != mi_ && !(moved && r == mj_);
        int pc0 = -1, pc1 = -1;  // patched columns
        float pv0 = 0.0f, pv1 = 0.0f;
        if (patch) {
            const float vi = row[mi_], vj = row[mj_], vl = row[order];
            pc0 = mi_;
            if (moved && mj_ == mi_) {  // the (0, 0) pair: the move wins
                pv0 = vl;
            } else {
                pv0 = merge_rule(rule, msi, msj, msj, mdij, vi, vj);
                if (moved) {
                    pc1 = mj_;
                    pv1 = vl;
                }
            }
        }
        if (r > 0) {
            // smallest value of the prefix, first index on ties: lane l takes the
            // 4-float groups l, l + 64, ... (rows are 16-byte aligned, ld % 4 == 0);
            // the patched columns read as +inf (never a candidate)
            auto mask = [&](float4 x, int g4) {
                if (pc0 >> 2 == g4) {
                    const int k = pc0 & 3;
                    x.x = k == 0 ? kInf : x.x; x.y = k == 1 ? kInf : x.y;
                    x.z = k == 2 ? kInf : x.z; x.w = k == 3 ? kInf : x.w;
                }
                if (pc1 >> 2 == g4) {
                    const int k = pc1 & 3;
                    x.x = k == 0 ? kInf : x.x; x.y = k == 1 ? kInf : x.y;
                    x.z = k == 2 ? kInf : x.z; x.w = k == 3 ? kInf : x.w;
                }
                return x;
            };
            float v = kFltMax;
            int bi = 0x7FFFFFFF;
            const int r4 = r >> 2;
            const float4* row4 = reinterpret_cast<const float4*>(row);
            int g = lane;
            for (; g + 192 < r4; g += 256) {
                const float4 x0 = row4[g], x1 = row4[g + 64], x2 = row4[g + 128], x3 = row4[g + 192];
                const float4 xs[4] = {mask(x0, g), mask(x1, g + 64), mask(x2, g + 128), mask(x3, g + 192)};
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int c0 = 4 * (g + 64 * u);
                    if (xs[u].x < v) { v = xs[u].x; bi = c0; }
                    if (xs[u].y < v) { v = xs[u].y; bi = c0 + 1; }
                    if (xs[u].z < v) { v = xs[u].z; bi = c0 + 2; }
                    if (xs[u].w < v) { v = xs[u].w; bi = c0 + 3; }
                }
            }
            for (; g < r4; g += 64) {
                const float4 x = mask(row4[g], g);
                const int c0 = 4 * g;
                if (x.x < v) { v = x.x; bi = c0; }
                if (x.y < v) { v = x.y; bi = c0 + 1; }
                if (x.z < v) { v = x.z; bi = c0 + 2; }
                if (x.w < v) { v = x.w; bi = c0 + 3; }
            }
            const int c_tail = 4 * r4 + lane;
            if (c_tail < r && c_tail != pc0 && c_tail != pc1) {
                const float x = row[c_tail];
                if (x < v) { v = x; bi = c_tail; }
            }
            // the lane's candidate of this row joins its running minimum; the
            // wave's lanes are combined once, after the last row
            if (bi != 0x7FFFFFFF) best = min(best, cand_key(v, r, bi));
            // the patched prefix entries, under the scan's rule: a value below
            // f32::MAX (the same in every lane)
            if (pc0 >= 0 && pc0 < r && pv0 < kFltMax) best = min(best, cand_key(pv0, r, pc0));
            if (pc1 >= 0 && pc1 < r && pv1 < kFltMax) best = min(best, cand_key(pv1, r, pc1));
        }
        if (patch && lane == 0) {
            row[pc0] = pv0;
            if (pc1 >= 0) row[pc1] = pv1;
        }
    }
    best = wave_min_u64(best);
    if (lane == 0) wbest[w] = best;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = min(min(wbest[0], wbest[1]), min(wbest[2], wbest[3]));
}

// The closest pair over the work-groups' candidates (merge s): records it,
// sets sizes[i] = size_j + size_i (hierarchical.rs:66), and rewrites rows i
// and j for the active elements c < order - 1 that remain.  Thread c reads
// column c of the old rows i, j and last and writes column c of the new rows
// i and j, so the threads never touch each other's data.
__global__ __launch_bounds__(1024) void cluster_decide_kernel(float* __restrict__ D, int n, int ld,
                                                              MergeState* __restrict__ st,
                                                              const uint64_t* __restrict__ part, int n_part,
                                                              unsigned* __restrict__ sizes, int rule,
                                                              unsigned* __restrict__ mi,
                                                              unsigned* __restrict__ mj,
                                                              float* __restrict__ md)
{
    __shared__ uint64_t wkey[16];
    const int s = st->s;  // read by every thread before thread 0 advances it below
    if (s >= n - 1) return;
    // the minimum key over the candidates: one wave reduction each, then the
    // 16 wave minima from LDS in every thread
    uint64_t key = ~0ull;
    for (int k = threadIdx.x; k < n_part; k += 1024) key = min(key, part[k]);
    key = wave_min_u64(key);
    if ((threadIdx.x & 63) == 0) wkey[threadIdx.x >> 6] = key;
    __syncthreads();
    key = wkey[0];
#pragma unroll
    for (int w = 1; w < 16; ++w) key = min(key, wkey[w]);
    // every thread loads, in one round, the old values it needs: the pair's
    // sizes and distance, both rows' entries at the last active index and its
    // first groups of rows i, j and last.  No candidate (every value >=
    // f32::MAX): the reference keeps (0, 0).
    const int i = key == ~0ull ? 0 : (int)(key & 0xFFFFu);
    const int j = key == ~0ull ? 0 : (int)((key >> 16) & 0xFFFFu);
    const int order = n - s - 1;  // active elements after this merge; index `order` is the last
    const bool moved = j < order;
    float* Ri = D + (int64_t)i * ld;
    float* Rj = D + (int64_t)j * ld;
    const float* Rl = D + (int64_t)order * ld;
    const unsigned si = sizes[i], sj = sizes[j];
    const float dij = Ri[j];
    const float ri_last = Ri[order], rj_last = Rj[order];
    // four columns per thread and group, two groups per trip with every load
    // issued before any store (rows are 16-byte aligned; columns at or past
    // `order` keep their values)
    float4* Ri4 = reinterpret_cast<float4*>(Ri);
    float4* Rj4 = reinterpret_cast<float4*>(Rj);
    const float4* Rl4 = reinterpret_cast<const float4*>(Rl);
    const int g_end = (order + 3) >> 2;
    float4 a[2], b4[2], l4[2];
    auto load_trip = [&](int g0) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int g = g0 + 1024 * u;
            if (g < g_end) {
                a[u] = Ri4[g];
                b4[u] = Rj4[g];
                l4[u] = Rl4[g];
            }
        }
    };
    load_trip(threadIdx.x);
    // every read of a value that this kernel overwrites (D[i][j] is in row i,
    // sizes[i]) is done before any store
    __syncthreads();
    if (threadIdx.x == 0) {
        st->i = i;
        st->j = j;
        st->si = si;
        st->sj = sj;
        st->dij = dij;
        sizes[i] = sj + si;
        mi[s] = (unsigned)i;
        mj[s] = (unsigned)j;
        md[s] = dij;
        st->s = s + 1;
    }
    // res[last]: the merged cluster to the element that moves into j
    const float rl = moved ? merge_rule(rule, si, sj, sj, dij, ri_last, rj_last) : 0.0f;
    auto store_trip = [&](int g0) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int g = g0 + 1024 * u;
            if (g >= g_end) continue;
            const float oi[4] = {a[u].x, a[u].y, a[u].z, a[u].w};
            const float oj[4] = {b4[u].x, b4[u].y, b4[u].z, b4[u].w};
            const float ol[4] = {l4[u].x, l4[u].y, l4[u].z, l4[u].w};
            float ni[4], nj[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int col = 4 * g + k;
                if (col >= order) {
                    ni[k] = oi[k];
                    nj[k] = oj[k];
                    continue;
                }
                if (col == i)
                    ni[k] = 0.0f;
                else if (moved && col == j)
                    ni[k] = rl;
                else
                    ni[k] = merge_rule(rule, si, sj, sj, dij, oi[k], oj[k]);
                nj[k] = col == j ? 0.0f : (col == i ? rl : ol[k]);
            }
            Ri4[g] = make_float4(ni[0], ni[1], ni[2], ni[3]);
            if (moved) Rj4[g] = make_float4(nj[0], nj[1], nj[2], nj[3]);  // with the (0, 0) pair the move wins
        }
    };
    store_trip(threadIdx.x);
    for (int g0 = threadIdx.x + 2048; g0 < g_end; g0 += 2048) {
        load_trip(g0);
        store_trip(g0);
    }
}

}  // namespace

// Tiles per launch: the scratch holds the results (1 KiB each) of at most
// this many tiles, or of one tile row if that is longer, so it stays bounded
// for any n.  BIOSHELL_ROW_STATS_CHUNK_TILES lowers it (tests of the chunking).
static int row_stats_chunk_tiles()
{
    int c = 1 << 16;
    if (const char* e = std::getenv("BIOSHELL_ROW_STATS_CHUNK_TILES")) c = std::max(1, std::atoi(e));
    return c;
}

size_t row_stats_scratch_bytes(int n)
{
    const size_t T = (size_t)(n + 63) / 64;
    const size_t cap = std::max<size_t>((size_t)row_stats_chunk_tiles(), T);
    return std::max<size_t>(1, std::min<size_t>(T * (T + 1) / 2, cap)) * 256 * 4;
}

void launch_row_stats(const float* up, int n, float* rmin, float* rmax, void* scratch, hipStream_t s)
{
    const int T = (n + 63) / 64;
    auto* part = static_cast<unsigned*>(scratch);
    auto* a = reinterpret_cast<unsigned*>(rmin);
    auto* b = reinterpret_cast<unsigned*>(rmax);
    if (n < 2) {
        row_stats_reduce_kernel<<<T, 512, 0, s>>>(part, n, T, 0, 0, 0, 1, 1, a, b);
        return;
    }
    // chunks of whole tile rows, each within the chunk size (or one tile row)
    const int chunk = row_stats_chunk_tiles();
    int t0 = 0;
    while (t0 < T) {
        const int base = t0 * T - t0 * (t0 - 1) / 2;
        int t1 = t0 + 1;
        while (t1 < T && (t1 + 1) * T - (t1 + 1) * t1 / 2 - base <= chunk) ++t1;
        const int tiles = t1 * T - t1 * (t1 - 1) / 2 - base;
        row_stats_tile_kernel<<<tiles, 256, 0, s>>>(up, n, T, base, part);
        row_stats_reduce_kernel<<<T, 512, 0, s>>>(part, n, T, t0, t1, base, t0 == 0, t1 == T, a, b);
        t0 = t1;
    }
}

void launch_row_stats_combine(const unsigned* part, int n, unsigned* kmin, unsigned* kmax, hipStream_t s)
{
    const int T = (n + 63) / 64;
    row_stats_reduce_kernel<<<T, 512, 0, s>>>(part, n, T, 0, T, 0, 1, 0, kmin, kmax);
}

void launch_row_stats_decode(unsigned* kmin, unsigned* kmax, int n, hipStream_t s)
{
    row_stats_decode_kernel<<<(n + 255) / 256, 256, 0, s>>>(kmin, kmax, n);
}

void launch_members_row_max(const float* up, int n, const unsigned* members, int m, float* out,
                            hipStream_t s)
{
    members_row_max_kernel<<<m, 256, 0, s>>>(up, n, members, m, out);
}

void launch_cluster_init(const float* up, int n, int ld, float* D, unsigned* sizes, void* state, hipStream_t s)
{
    const int64_t nb = std::max<int64_t>(1, std::min<int64_t>(4096, ((int64_t)n * ld + 255) / 256));
    dense_from_upper_kernel<<<(int)nb, 256, 0, s>>>(up, n, ld, D, sizes);
    fill_u32_kernel<<<1, 256, 0, s>>>(static_cast<unsigned*>(state), 0u,
                                      (int64_t)(sizeof(MergeState) / sizeof(unsigned)));
}

void launch_cluster_step(float* D, int n, int ld, void* state, void* part, int n_part, unsigned* sizes,
                         int rule, unsigned* mi, unsigned* mj, float* md, hipStream_t s)
{
    auto* st = static_cast<MergeState*>(state);
    auto* cand = static_cast<uint64_t*>(part);
    cluster_step_kernel<<<n_part, 256, 0, s>>>(D, n, ld, st, cand, rule);
    cluster_decide_kernel<<<1, 1024, 0, s>>>(D, n, ld, st, cand, n_part, sizes, rule, mi, mj, md);
}

size_t cluster_state_bytes() { return sizeof(MergeState); }
size_t cluster_candidate_bytes() { return sizeof(uint64_t); }

}  // namespace bs
