/*
 * crmsd_oracle.c -- FP64 CPU oracle for optimal-superposition cRMSD.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the GPU path is
 * compared against (and the CPU baseline that is timed beside it).  Nothing in
 * the product library links, loads or calls it.
 *
 * The reference snapshot (dgront/BioShell v4.0.0-dev) has no crmsd function;
 * the only trace is the dependency note at bioshell-pdb/Cargo.toml:27
 * ("nalgebra ... for helical axis and crmsd").  This is therefore a
 * restatement of the textbook algorithm in the reference's own numerical
 * conventions:
 *   - coordinates are f64 x,y,z triples, one Vec3 per atom
 *     (bioshell-core/src/vec3.rs:43-51), stored AoS: xyz[3*k + c];
 *   - 3x3 matrices are row-major [f64; 9] (bioshell-core/src/matrix.rs:7-12),
 *     det as in matrix.rs:339-345;
 *   - centring is "sum positions, divide by n, subtract"
 *     (bioshell-pdb/src/calc/substructures.rs:27-42);
 *   - the covariance has the contraction shape of substructures.rs:42-45
 *     (3xL times Lx3) but is the cross-covariance M = sum_k a_k b_k^T, no 1/n;
 *   - the largest eigenvalue of Horn's 4x4 key matrix is found by the QCP
 *     method (Theobald, Acta Cryst. A61:478, 2005; Liu, Agrafiotis & Theobald,
 *     J. Comput. Chem. 31:1561, 2010): Newton on the characteristic quartic
 *     from E0 = (Ga+Gb)/2;
 *   - RMSD = sqrt(max(0, 2 (E0 - lambda) / L)).
 * Independent cross-checks in the same file: a Jacobi-SVD Kabsch path and a
 * 4x4 Jacobi eigen-solver on Horn's matrix.  The superposition output follows
 * the Rototranslation convention x' = R x + t
 * (bioshell-pdb/src/calc/transformation.rs:22-37, apply_mut at 171-174).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- basics */

/* substructures.rs:27-42: sum, divide by n, subtract. */
void or_centre(const double* xyz, long n_atoms, double* out, double centroid[3])
{
    double c[3] = {0.0, 0.0, 0.0};
    for (long k = 0; k < n_atoms; ++k) {
        c[0] += xyz[3 * k + 0];
        c[1] += xyz[3 * k + 1];
        c[2] += xyz[3 * k + 2];
    }
    c[0] /= (double)n_atoms;
    c[1] /= (double)n_atoms;
    c[2] /= (double)n_atoms;
    for (long k = 0; k < n_atoms; ++k) {
        out[3 * k + 0] = xyz[3 * k + 0] - c[0];
        out[3 * k + 1] = xyz[3 * k + 1] - c[1];
        out[3 * k + 2] = xyz[3 * k + 2] - c[2];
    }
    if (centroid) { centroid[0] = c[0]; centroid[1] = c[1]; centroid[2] = c[2]; }
}

/* Cross-covariance M[3p+q] = sum_k a_kp b_kq of already-centred sets, plus
 * the inner products Ga = sum |a_k|^2 and Gb = sum |b_k|^2. */
void or_covariance(const double* a, const double* b, long n_atoms, double M[9],
                   double* Ga, double* Gb)
{
    double m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double ga = 0.0, gb = 0.0;
    for (long k = 0; k < n_atoms; ++k) {
        const double* p = a + 3 * k;
        const double* q = b + 3 * k;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) m[3 * r + c] += p[r] * q[c];
        ga += p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
        gb += q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
    }
    memcpy(M, m, sizeof(m));
    if (Ga) *Ga = ga;
    if (Gb) *Gb = gb;
}

/* matrix.rs:339-345 (row-major determinant). */
static double det3(const double m[9])
{
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
           m[2] * (m[3] * m[7] - m[4] * m[6]);
}

/* Horn's symmetric, traceless 4x4 key matrix of M (row-major 16). */
void or_key_matrix(const double M[9], double K[16])
{
    const double xx = M[0], xy = M[1], xz = M[2];
    const double yx = M[3], yy = M[4], yz = M[5];
    const double zx = M[6], zy = M[7], zz = M[8];
    const double k[16] = {
        xx + yy + zz, yz - zy,       zx - xz,       xy - yx,
        yz - zy,      xx - yy - zz,  xy + yx,       zx + xz,
        zx - xz,      xy + yx,       -xx + yy - zz, yz + zy,
        xy - yx,      zx + xz,       yz + zy,       -xx - yy + zz};
    memcpy(K, k, sizeof(k));
}

/* Determinant of a 4x4 by cofactor expansion along row 0. */
static double det4(const double K[16])
{
    double d = 0.0;
    for (int c = 0; c < 4; ++c) {
        double sub[9];
        int t = 0;
        for (int r = 1; r < 4; ++r)
            for (int cc = 0; cc < 4; ++cc)
                if (cc != c) sub[t++] = K[4 * r + cc];
        const double cof = det3(sub);
        d += ((c & 1) ? -1.0 : 1.0) * K[c] * cof;
    }
    return d;
}

/* ------------------------------------------------------------------- QCP */

double or_lambda_jacobi4(const double M[9]);

/* Largest root of P(l) = l^4 + c2 l^2 + c1 l + c0, the characteristic
 * polynomial of Horn's matrix, with c2 = -2 |M|_F^2, c1 = -8 det M and
 * c0 = det K (Theobald 2005).  Newton from the upper bound E0; the iteration
 * stops when the step is below 1e-11 relative (or after 50 steps).
 * Near-degenerate polish: Newton converges only linearly at a double root
 * (e.g. collinear sets), where the polynomial fixes the root to ~sqrt(eps)
 * only; when more than 8 steps were needed, or P'(lambda) is below 1e-4 of
 * the size of its terms, the eigenvalue is taken from a Jacobi solve of K
 * instead, which is accurate to eps |K| regardless. */
double or_qcp_lambda(const double M[9], double E0, int* iterations)
{
    double K[16];
    or_key_matrix(M, K);
    double f2 = 0.0;
    for (int i = 0; i < 9; ++i) f2 += M[i] * M[i];
    const double c2 = -2.0 * f2;
    const double c1 = -8.0 * det3(M);
    const double c0 = det4(K);

    double l = E0, dp = 0.0;
    int it = 0;
    for (; it < 50; ++it) {
        const double l2 = l * l;
        const double p = ((l2 + c2) * l + c1) * l + c0;
        dp = (4.0 * l2 + 2.0 * c2) * l + c1;
        if (dp == 0.0) break;
        const double step = p / dp;
        l -= step;
        if (fabs(step) <= 1e-11 * fabs(l)) { ++it; break; }
    }
    if (iterations) *iterations = it;
    /* P'(lambda) is the product of the gaps to the other three eigenvalues: a
     * small P' relative to its terms means a near-double root */
    const double scale = fabs(4.0 * l * l * l) + fabs(2.0 * c2 * l) + fabs(c1);
    if (it > 8 || fabs(dp) <= 1e-4 * scale) return or_lambda_jacobi4(M);
    return l;
}

static double rmsd_from(double E0, double lambda, long n_atoms)
{
    double r2 = 2.0 * (E0 - lambda) / (double)n_atoms;
    return sqrt(r2 > 0.0 ? r2 : 0.0);
}

/* cRMSD of two conformations after optimal superposition (QCP).
 * Returns NAN when n_atoms < 1 (the precondition the reference asserts on at
 * substructures.rs:19). */
double or_crmsd(const double* a_xyz, const double* b_xyz, long n_atoms)
{
    if (n_atoms < 1) return NAN;
    double* a = (double*)malloc(sizeof(double) * 3 * n_atoms);
    double* b = (double*)malloc(sizeof(double) * 3 * n_atoms);
    or_centre(a_xyz, n_atoms, a, NULL);
    or_centre(b_xyz, n_atoms, b, NULL);
    double M[9], Ga, Gb;
    or_covariance(a, b, n_atoms, M, &Ga, &Gb);
    free(a);
    free(b);
    const double E0 = 0.5 * (Ga + Gb);
    return rmsd_from(E0, or_qcp_lambda(M, E0, NULL), n_atoms);
}

/* Same quantity on already-centred inputs (used by the all-vs-all driver). */
static double crmsd_centred(const double* a, const double* b, long n_atoms)
{
    double M[9], Ga, Gb;
    or_covariance(a, b, n_atoms, M, &Ga, &Gb);
    const double E0 = 0.5 * (Ga + Gb);
    return rmsd_from(E0, or_qcp_lambda(M, E0, NULL), n_atoms);
}

/* ------------------------------------------------------ Jacobi solvers */

/* Cyclic Jacobi eigen-decomposition of a symmetric n x n matrix (n <= 4),
 * row-major A (destroyed).  Eigenvalues to w[], eigenvectors as columns of V. */
static void jacobi_sym(double* A, int n, double* w, double* V)
{
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) V[n * i + j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                tot += A[n * i + j] * A[n * i + j];
                if (i != j) off += A[n * i + j] * A[n * i + j];
            }
        if (off <= 1e-60 * (tot > 0 ? tot : 1.0) || off == 0.0) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[n * p + q];
                if (apq == 0.0) continue;
                const double app = A[n * p + p], aqq = A[n * q + q];
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) /
                                 (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = A[n * k + p], akq = A[n * k + q];
                    A[n * k + p] = c * akp - s * akq;
                    A[n * k + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = A[n * p + k], aqk = A[n * q + k];
                    A[n * p + k] = c * apk - s * aqk;
                    A[n * q + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = V[n * k + p], vkq = V[n * k + q];
                    V[n * k + p] = c * vkp - s * vkq;
                    V[n * k + q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < n; ++i) w[i] = A[n * i + i];
}

/* Largest eigenvalue of Horn's matrix by Jacobi (independent of QCP). */
double or_lambda_jacobi4(const double M[9])
{
    double K[16], w[4], V[16];
    or_key_matrix(M, K);
    jacobi_sym(K, 4, w, V);
    double best = w[0];
    for (int i = 1; i < 4; ++i)
        if (w[i] > best) best = w[i];
    return best;
}

/* Kabsch: lambda = s1 + s2 + sign(det M) s3 from the singular values of M,
 * obtained as square roots of the eigenvalues of M^T M (Jacobi). */
double or_lambda_kabsch(const double M[9])
{
    double MtM[9], w[3], V[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += M[3 * k + i] * M[3 * k + j];
            MtM[3 * i + j] = s;
        }
    jacobi_sym(MtM, 3, w, V);
    double s[3];
    for (int i = 0; i < 3; ++i) s[i] = sqrt(w[i] > 0 ? w[i] : 0.0);
    /* sort descending */
    for (int i = 0; i < 3; ++i)
        for (int j = i + 1; j < 3; ++j)
            if (s[j] > s[i]) { double t = s[i]; s[i] = s[j]; s[j] = t; }
    const double sg = det3(M) < 0.0 ? -1.0 : 1.0;
    return s[0] + s[1] + sg * s[2];
}

double or_crmsd_kabsch(const double* a_xyz, const double* b_xyz, long n_atoms)
{
    if (n_atoms < 1) return NAN;
    double* a = (double*)malloc(sizeof(double) * 3 * n_atoms);
    double* b = (double*)malloc(sizeof(double) * 3 * n_atoms);
    or_centre(a_xyz, n_atoms, a, NULL);
    or_centre(b_xyz, n_atoms, b, NULL);
    double M[9], Ga, Gb;
    or_covariance(a, b, n_atoms, M, &Ga, &Gb);
    free(a);
    free(b);
    return rmsd_from(0.5 * (Ga + Gb), or_lambda_kabsch(M), n_atoms);
}

/* ------------------------------------------------------- superposition */

/* Rotation R and translation t with a ~= R b + t (the Rototranslation
 * convention of transformation.rs:171-174 applied to b).  R comes from the
 * unit quaternion that is the eigenvector of Horn's matrix for its largest
 * eigenvalue (Jacobi).  Returns the RMSD of the fit. */
double or_superpose(const double* a_xyz, const double* b_xyz, long n_atoms, double R[9],
                    double t[3])
{
    if (n_atoms < 1) return NAN;
    double* a = (double*)malloc(sizeof(double) * 3 * n_atoms);
    double* b = (double*)malloc(sizeof(double) * 3 * n_atoms);
    double ca[3], cb[3];
    or_centre(a_xyz, n_atoms, a, ca);
    or_centre(b_xyz, n_atoms, b, cb);
    double M[9], Ga, Gb;
    /* Horn's matrix built from sum b a^T: its top eigenvector rotates b onto a */
    or_covariance(b, a, n_atoms, M, &Gb, &Ga);
    free(a);
    free(b);
    double K[16], w[4], V[16];
    or_key_matrix(M, K);
    jacobi_sym(K, 4, w, V);
    int best = 0;
    for (int i = 1; i < 4; ++i)
        if (w[i] > w[best]) best = i;
    double q0 = V[4 * 0 + best], q1 = V[4 * 1 + best], q2 = V[4 * 2 + best], q3 = V[4 * 3 + best];
    const double nq = sqrt(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3);
    q0 /= nq; q1 /= nq; q2 /= nq; q3 /= nq;
    R[0] = q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3;
    R[1] = 2.0 * (q1 * q2 - q0 * q3);
    R[2] = 2.0 * (q1 * q3 + q0 * q2);
    R[3] = 2.0 * (q1 * q2 + q0 * q3);
    R[4] = q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3;
    R[5] = 2.0 * (q2 * q3 - q0 * q1);
    R[6] = 2.0 * (q1 * q3 - q0 * q2);
    R[7] = 2.0 * (q2 * q3 + q0 * q1);
    R[8] = q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3;
    for (int r = 0; r < 3; ++r)
        t[r] = ca[r] - (R[3 * r + 0] * cb[0] + R[3 * r + 1] * cb[1] + R[3 * r + 2] * cb[2]);
    return rmsd_from(0.5 * (Ga + Gb), w[best], n_atoms);
}

/* ----------------------------------------------------- all-vs-all driver */

/* Packed upper triangle index of (i, j), i < j, row-major over i. */
static inline long tri_index(long n, long i, long j)
{
    return i * n - i * (i + 1) / 2 + (j - i - 1);
}

typedef struct {
    const double* centred;
    long n, L, row0, row1;
    int tid, nthreads;
    double* out;   /* packed upper triangle (full) or NULL */
    double* rows;  /* rows [row0,row1) x n dense, or NULL */
} avsa_job;

static void* avsa_worker(void* p)
{
    avsa_job* j = (avsa_job*)p;
    /* rows dealt round-robin: row i holds n-1-i pairs, so interleaving balances */
    for (long i = j->row0 + j->tid; i < j->row1; i += j->nthreads) {
        const double* a = j->centred + 3 * j->L * i;
        for (long k = i + 1; k < j->n; ++k) {
            const double r = crmsd_centred(a, j->centred + 3 * j->L * k, j->L);
            if (j->out) j->out[tri_index(j->n, i, k)] = r;
            if (j->rows) j->rows[(i - j->row0) * j->n + k] = r;
        }
    }
    return NULL;
}

static int run_avsa(const double* xyz, long n, long L, long row0, long row1, double* out,
                    double* rows, int nthreads)
{
    if (n < 1 || L < 1 || row0 < 0 || row1 > n || row0 > row1) return -1;
    if (nthreads < 1) nthreads = 1;
    double* centred = (double*)malloc(sizeof(double) * 3 * L * n);
    if (!centred) return -2;
    for (long i = 0; i < n; ++i) or_centre(xyz + 3 * L * i, L, centred + 3 * L * i, NULL);
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
    avsa_job* jobs = (avsa_job*)malloc(sizeof(avsa_job) * nthreads);
    for (int t = 0; t < nthreads; ++t) {
        jobs[t] = (avsa_job){centred, n, L, row0, row1, t, nthreads, out, rows};
        pthread_create(&th[t], NULL, avsa_worker, &jobs[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th);
    free(jobs);
    free(centred);
    return 0;
}

/* All-vs-all cRMSD: xyz is n_conf x n_atoms x 3 (AoS, row-major), out_upper
 * receives the n(n-1)/2 values of the packed upper triangle, row-major. */
int or_all_vs_all(const double* xyz, long n_conf, long n_atoms, double* out_upper, int nthreads)
{
    return run_avsa(xyz, n_conf, n_atoms, 0, n_conf, out_upper, NULL, nthreads);
}

/* Rows [row0, row1) of the all-vs-all matrix (only j > i is written) into a
 * dense (row1-row0) x n_conf buffer; used to time a bounded row sample. */
int or_all_vs_all_rows(const double* xyz, long n_conf, long n_atoms, long row0, long row1,
                       double* out_rows, int nthreads)
{
    return run_avsa(xyz, n_conf, n_atoms, row0, row1, NULL, out_rows, nthreads);
}

/* Query-vs-library cRMSD: out[q * n_lib + l]. */
int or_query_vs_library(const double* q_xyz, long n_q, const double* lib_xyz, long n_lib,
                        long n_atoms, double* out)
{
    if (n_atoms < 1) return -1;
    double* a = (double*)malloc(sizeof(double) * 3 * n_atoms);
    double* b = (double*)malloc(sizeof(double) * 3 * n_atoms);
    for (long q = 0; q < n_q; ++q) {
        or_centre(q_xyz + 3 * n_atoms * q, n_atoms, a, NULL);
        for (long l = 0; l < n_lib; ++l) {
            or_centre(lib_xyz + 3 * n_atoms * l, n_atoms, b, NULL);
            out[q * n_lib + l] = crmsd_centred(a, b, n_atoms);
        }
    }
    free(a);
    free(b);
    return 0;
}
