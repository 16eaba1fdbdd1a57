/*
 * cluster_oracle.c -- CPU restatement of BioShell's agglomerative clustering
 * loop and of the two row reductions its tools use.
 *
 * TEST INFRASTRUCTURE ONLY: the checker the device clustering is compared
 * against.  Nothing in the product library links or loads it.
 *
 * Follows bioshell-clustering/src/hierarchical/ exactly, including the
 * behaviour a faithful port must keep:
 *   - the matrix is filled from d(i, j) for i > j and mirrored
 *     (clustering_matrix.rs:11-21);
 *   - the closest pair is the first strict minimum scanning j = 1.., then
 *     i = 0..j-1 over the active order (clustering_matrix.rs:27-42);
 *   - merging (i, j) rewrites row and column i with
 *     rule(size_i, size_j, sizes[j], d_ij, d_ik, d_jk) for every active k --
 *     the third argument is sizes[j], not sizes[k] (clustering_matrix.rs:57-74) --
 *     then sets d_ii = 0 and sizes[i] = size_i + size_j;
 *   - the last active element then replaces j: its row is swapped into j and
 *     its column copied over column j; sizes[] is not moved
 *     (clustering_matrix.rs:47-55, hierarchical.rs:62-72);
 *   - the merging rules are the f32 expressions of strategies/mod.rs,
 *     evaluated left to right without fused multiply-adds (this file is
 *     compiled with -ffp-contract=off).
 * The merge sequence (i, j, d_ij) per step is what the tree is rebuilt from.
 *
 * Pinned by tests/test_clustering_oracle.py: the reference's clustering tests,
 * MST heights (single), scipy heights (single, complete; average, centroid and
 * median on a chain where no cluster moves), the textbook first merge of
 * Ward's rule, and hand-traced six-element cases for average, centroid and
 * Ward linkage in which a merged cluster is moved and then merged again (the
 * sizes[j] argument and the sizes left in place decide those heights).
 * or_hclust_variant() plants the two plausible mistakes -- sizes moved with
 * the row, sizes[k] as the third argument -- so that the test can show the
 * hand-traced cases reject them.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { OR_SINGLE = 0, OR_COMPLETE = 1, OR_AVERAGE = 2, OR_MEDIAN = 3, OR_CENTROID = 4, OR_WARDS = 5 };

/* strategies/mod.rs: single_link, complete_link, average_link, median_link,
 * centroid_link, wards_method */
static float or_rule(int rule, long si, long sj, long sk, float dij, float dik, float djk)
{
    switch (rule) {
    case OR_SINGLE: return fminf(dik, djk);
    case OR_COMPLETE: return fmaxf(dik, djk);
    case OR_AVERAGE: {
        const float d = 1.0f / (float)(si + sj);
        return d * (float)si * dik + d * (float)sj * djk;
    }
    case OR_MEDIAN: return 0.5f * dik + 0.5f * djk - 0.25f * dij;
    case OR_CENTROID: {
        const float d = 1.0f / (float)(si + sj);
        return d * (float)si * dik + d * (float)sj * djk - (float)si * (float)sj * d * d * dij;
    }
    default: {
        const float d = 1.0f / (float)(si + sj + sk);
        return d * (float)(si + sk) * dik + d * (float)(sj + sk) * djk - (float)sk * d * dij;
    }
    }
}

/* upper: packed upper triangle (i < j, row-major) of n elements.  Writes the
 * n - 1 merges.  Returns 0, or -1 on allocation failure.
 * variant 0 is the reference's loop; 1 moves sizes[last] to j with the row,
 * 2 passes sizes[k] as the rule's third argument (planted mistakes). */
int or_hclust_variant(const float* upper, long n, int rule, int variant, int* mi, int* mj,
                      float* md)
{
    if (n < 2) return 0;
    float* m = (float*)malloc(sizeof(float) * (size_t)n * (size_t)n);
    long* sizes = (long*)malloc(sizeof(long) * (size_t)n);
    float* res = (float*)malloc(sizeof(float) * (size_t)n);
    if (!m || !sizes || !res) {
        free(m); free(sizes); free(res);
        return -1;
    }
#define M(r, c) m[(size_t)(r) * (size_t)n + (size_t)(c)]
    for (long i = 0; i < n; ++i) {
        M(i, i) = 0.0f;
        sizes[i] = 1;
    }
    for (long i = 1; i < n; ++i)
        for (long j = 0; j < i; ++j) {
            const float v = upper[j * n - j * (j + 1) / 2 + (i - j - 1)]; /* d(i, j) = d(j, i) */
            M(i, j) = v;
            M(j, i) = v;
        }
    long order = n;
    for (long step = 0; step < n - 1; ++step) {
        /* closest_elements */
        float best = FLT_MAX;
        long bi = 0, bj = 0;
        for (long j = 1; j < order; ++j)
            for (long i = 0; i < j; ++i)
                if (M(i, j) < best) {
                    best = M(i, j);
                    bi = i;
                    bj = j;
                }
        const float dij = M(bi, bj);
        mi[step] = (int)bi;
        mj[step] = (int)bj;
        md[step] = dij;
        /* update_distances(i, j, rule, new_index = i) */
        const long si = sizes[bi], sj = sizes[bj];
        for (long k = 0; k < order; ++k)
            res[k] = or_rule(rule, si, sj, variant == 2 ? sizes[k] : sizes[bj], dij, M(bi, k),
                             M(bj, k));
        for (long k = 0; k < order; ++k) {
            M(bi, k) = res[k];
            M(k, bi) = res[k];
        }
        M(bi, bi) = 0.0f;
        sizes[bi] = sj + si;
        /* replace_with_last(j) when j is not the last element */
        const long last = order - 1;
        if (bj < last) {
            order -= 1;
            for (long c = 0; c < n; ++c) { /* swap rows j and last */
                const float t = M(bj, c);
                M(bj, c) = M(order, c);
                M(order, c) = t;
            }
            for (long r = 0; r < order; ++r) M(r, bj) = M(r, order);
            if (variant == 1) sizes[bj] = sizes[order];
        } else {
            order -= 1;
        }
    }
#undef M
    free(m);
    free(sizes);
    free(res);
    return 0;
}

int or_hclust(const float* upper, long n, int rule, int* mi, int* mj, float* md)
{
    return or_hclust_variant(upper, n, rule, 0, mi, mj, md);
}

/* Row minimum and maximum over j != i (retrieve_outliers scans for the minimum
 * from f32::MAX with <, hierarchical.rs:198-217; medoid_by_min_max for the
 * maximum from -f32::MAX with >, hierarchical.rs:105-131). */
void or_row_stats(const float* upper, long n, float* rmin, float* rmax)
{
    for (long i = 0; i < n; ++i) {
        float lo = FLT_MAX, hi = -FLT_MAX;
        for (long j = 0; j < n; ++j) {
            if (i == j) continue;
            const long a = i < j ? i : j, b = i < j ? j : i;
            const float d = upper[a * n - a * (a + 1) / 2 + (b - a - 1)];
            if (d < lo) lo = d;
            if (d > hi) hi = d;
        }
        rmin[i] = lo;
        rmax[i] = hi;
    }
}
