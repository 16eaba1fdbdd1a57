// qcp.hpp -- quaternion characteristic polynomial (QCP) solves, host+device.
//
// Two forms of the same eigenvalue problem:
//
//  qcp_rmsd<T>()   the textbook form (Theobald 2005): the largest eigenvalue
//                  lambda of Horn's 4x4 key matrix K(M) is the largest root of
//                  P(l) = l^4 + c2 l^2 + c1 l + c0 with c2 = -2|M|_F^2,
//                  c1 = -8 det M, c0 = det K, found by Newton from the upper
//                  bound E0 = (Ga+Gb)/2; RMSD = sqrt(max(0, 2(E0-lambda)/L)).
//                  Used by the per-pair kernels (T = double or float) and by
//                  the FP64 re-evaluation of flagged pairs.
//
//  qcp_shifted()   the same root, solved in a frame where no large numbers
//                  cancel.  Every conformation x_i of an ensemble is first
//                  superposed onto a common reference r and written
//                  x_i = r + e_i.  Then, exactly,
//                      M_ij = R0 + C_i^T + C_j + E_ij,   R0 = sum r r^T,
//                      C_i = sum r e_i^T,                E_ij = sum e_i e_j^T,
//                  and with K' = K - tr(M) I = [[0, a^T], [a, B]]:
//                      a = w(C_j) - w(C_i) + w(E)     (antisymmetric parts,
//                          w(X) = (X_yz-X_zy, X_zx-X_xz, X_xy-X_yx); R0 drops out)
//                      B = Bh_i + Bh_j + B(E),  B(X) = X + X^T - 2 tr(X) I,
//                          Bh_i = B(R0)/2 + B(C_i)      (per conformation)
//                      H = E0 - tr M = (g_i + g_j)/2 - tr E,  g_i = sum |e_i|^2.
//                  The largest root mu of det(mu I - K') is lambda - tr M, lies
//                  in [0, H], and E0 - lambda = H - mu.  Only E_ij (residual
//                  magnitudes) comes from the MFMA Gram contraction; H - mu is
//                  formed without subtracting two numbers of size |x|^2.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace bs {

#define BS_HD __host__ __device__ __forceinline__

template <class T>
BS_HD T det3(const T* m)
{
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
           m[2] * (m[3] * m[7] - m[4] * m[6]);
}

// det of Horn's key matrix K(M), expanded along row 0.
template <class T>
BS_HD T det_key(const T* M)
{
    const T xx = M[0], xy = M[1], xz = M[2];
    const T yx = M[3], yy = M[4], yz = M[5];
    const T zx = M[6], zy = M[7], zz = M[8];
    const T k00 = xx + yy + zz, k01 = yz - zy, k02 = zx - xz, k03 = xy - yx;
    const T k11 = xx - yy - zz, k12 = xy + yx, k13 = zx + xz;
    const T k22 = -xx + yy - zz, k23 = yz + zy;
    const T k33 = -xx - yy + zz;
    // 2x2 minors of rows 2,3
    const T m23_01 = k02 * k13 - k12 * k03;  // cols (0,1) of rows (2,3): [k02 k12; k03 k13]
    const T m23_02 = k02 * k23 - k22 * k03;
    const T m23_03 = k02 * k33 - k23 * k03;
    const T m23_12 = k12 * k23 - k22 * k13;
    const T m23_13 = k12 * k33 - k23 * k13;
    const T m23_23 = k22 * k33 - k23 * k23;
    // cofactors of row 0 using rows 1..3
    const T c0 = k11 * m23_23 - k12 * m23_13 + k13 * m23_12;
    const T c1 = k01 * m23_23 - k12 * m23_03 + k13 * m23_02;
    const T c2 = k01 * m23_13 - k11 * m23_03 + k13 * m23_01;
    const T c3 = k01 * m23_12 - k11 * m23_02 + k12 * m23_01;
    return k00 * c0 - k01 * c1 + k02 * c2 - k03 * c3;
}

// Cyclic Jacobi on a symmetric 4x4 (destroyed).  The eigenvalues of a
// symmetric matrix are perfectly conditioned, so this stays accurate where the
// characteristic polynomial has a (near-)double root.
// Returns the unit eigenvector of the largest eigenvalue of symmetric A.
BS_HD void top_eigvec4(double A[4][4], double q[4], double* lambda)
{
    double V[4][4];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) V[r][c] = (r == c) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 30; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) {
                if (r == c)
                    diag += A[r][c] * A[r][c];
                else
                    off += A[r][c] * A[r][c];
            }
        if (off <= 1e-30 * diag || off == 0.0) break;
        for (int p = 0; p < 3; ++p)
            for (int r = p + 1; r < 4; ++r) {
                const double apr = A[p][r];
                if (apr == 0.0) continue;
                const double th = (A[r][r] - A[p][p]) / (2.0 * apr);
                const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 4; ++k) {
                    const double akp = A[k][p], akr = A[k][r];
                    A[k][p] = c * akp - s * akr;
                    A[k][r] = s * akp + c * akr;
                }
                for (int k = 0; k < 4; ++k) {
                    const double apk = A[p][k], ark = A[r][k];
                    A[p][k] = c * apk - s * ark;
                    A[r][k] = s * apk + c * ark;
                }
                for (int k = 0; k < 4; ++k) {
                    const double vkp = V[k][p], vkr = V[k][r];
                    V[k][p] = c * vkp - s * vkr;
                    V[k][r] = s * vkp + c * vkr;
                }
            }
    }
    int best = 0;
    for (int r = 1; r < 4; ++r)
        if (A[r][r] > A[best][best]) best = r;
    double nq = 0.0;
    for (int r = 0; r < 4; ++r) nq += V[r][best] * V[r][best];
    nq = sqrt(nq);
    for (int r = 0; r < 4; ++r) q[r] = V[r][best] / nq;
    if (lambda) *lambda = A[best][best];
}

// Horn's key matrix of M as a 4x4 array.
BS_HD void key_matrix(const double* M, double K[4][4])
{
    const double xx = M[0], xy = M[1], xz = M[2], yx = M[3], yy = M[4], yz = M[5], zx = M[6],
                 zy = M[7], zz = M[8];
    K[0][0] = xx + yy + zz; K[0][1] = yz - zy;       K[0][2] = zx - xz;       K[0][3] = xy - yx;
    K[1][0] = yz - zy;      K[1][1] = xx - yy - zz;  K[1][2] = xy + yx;       K[1][3] = zx + xz;
    K[2][0] = zx - xz;      K[2][1] = xy + yx;       K[2][2] = -xx + yy - zz; K[2][3] = yz + zy;
    K[3][0] = xy - yx;      K[3][1] = zx + xz;       K[3][2] = yz + zy;       K[3][3] = -xx - yy + zz;
}

// Textbook QCP; returns RMSD.  M row-major (M[3p+q] = sum a_p b_q) of centred
// sets, Ga, Gb their inner products, n the atom count.
template <class T>
BS_HD T qcp_rmsd(const T* M, T Ga, T Gb, int n, int max_iter = 50)
{
    T f2 = T(0);
#pragma unroll
    for (int i = 0; i < 9; ++i) f2 += M[i] * M[i];
    const T c2 = T(-2) * f2;
    const T c1 = T(-8) * det3(M);
    const T c0 = det_key(M);
    const T E0 = T(0.5) * (Ga + Gb);
    T l = E0;
    const T tol = sizeof(T) == 8 ? T(1e-11) : T(1e-6);
    int it = 0;
    for (; it < max_iter; ++it) {
        const T l2 = l * l;
        const T p = ((l2 + c2) * l + c1) * l + c0;
        const T dp = (T(4) * l2 + T(2) * c2) * l + c1;
        if (dp == T(0)) break;
        const T step = p / dp;
        l -= step;
        if (fabs(step) <= tol * fabs(l)) {
            ++it;
            break;
        }
    }
    // Near-degenerate polish (FP64 only): linear convergence means a (near-)
    // double root, which the polynomial pins to ~sqrt(eps) relative only.
    if constexpr (sizeof(T) == 8) {
      if (it > 8) {
        double K[4][4], q[4], lam;
        key_matrix(reinterpret_cast<const double*>(M), K);
        top_eigvec4(K, q, &lam);
        l = T(lam);
      }
    }
    T r2 = T(2) * (E0 - l) / T(n);
    return sqrt(r2 > T(0) ? r2 : T(0));
}

// Per-conformation constants of the shifted frame, stored SoA (stride = ld):
//   0..5  Bh = B(R0)/2 + B(C_i) as (b00, b11, b22, b01, b02, b12)
//   6..8  w(C_i)
//   9     h_i = g_i / 2
//   10    sqrt(g_i)
constexpr int kNumConsts = 11;

struct ShiftedOut {
    double rmsd2_L;  // 2 (H - mu), i.e. L * RMSD^2
    double H;
    int iterations;
    bool converged;
};

// Solves for the largest root of det(mu I - K') given the pair's residual
// Gram E (row-major, E[3p+q] = sum_k e_i,kp e_j,kq) and both conformations'
// constants.
BS_HD ShiftedOut qcp_shifted(const double* E, const double* ci, const double* cj,
                             int max_iter = 40)
{
    const double trE = E[0] + E[4] + E[8];
    // B = Bh_i + Bh_j + (E + E^T - 2 tr E I)
    const double b00 = ci[0] + cj[0] + 2.0 * (E[0] - trE);
    const double b11 = ci[1] + cj[1] + 2.0 * (E[4] - trE);
    const double b22 = ci[2] + cj[2] + 2.0 * (E[8] - trE);
    const double b01 = ci[3] + cj[3] + (E[1] + E[3]);
    const double b02 = ci[4] + cj[4] + (E[2] + E[6]);
    const double b12 = ci[5] + cj[5] + (E[5] + E[7]);
    // a = w(C_j) - w(C_i) + w(E)
    const double a0 = (cj[6] - ci[6]) + (E[5] - E[7]);
    const double a1 = (cj[7] - ci[7]) + (E[6] - E[2]);
    const double a2 = (cj[8] - ci[8]) + (E[1] - E[3]);
    const double H = (ci[9] + cj[9]) - trE;

    // invariants of B and its adjugate
    const double A00 = b11 * b22 - b12 * b12;
    const double A11 = b00 * b22 - b02 * b02;
    const double A22 = b00 * b11 - b01 * b01;
    const double A01 = b02 * b12 - b01 * b22;
    const double A02 = b01 * b12 - b02 * b11;
    const double A12 = b01 * b02 - b00 * b12;
    const double t1 = b00 + b11 + b22;
    const double t2 = A00 + A11 + A22;
    const double t3 = b00 * A00 + b01 * A01 + b02 * A02;
    const double aa = a0 * a0 + a1 * a1 + a2 * a2;
    const double aBa = b00 * a0 * a0 + b11 * a1 * a1 + b22 * a2 * a2 +
                       2.0 * (b01 * a0 * a1 + b02 * a0 * a2 + b12 * a1 * a2);
    const double aAa = A00 * a0 * a0 + A11 * a1 * a1 + A22 * a2 * a2 +
                       2.0 * (A01 * a0 * a1 + A02 * a0 * a2 + A12 * a1 * a2);
    // det(mu I - K') = mu^4 + p3 mu^3 + p2 mu^2 + p1 mu + p0
    const double p3 = -t1;
    const double p2 = t2 - aa;
    const double p1 = t1 * aa - aBa - t3;
    const double p0 = -aAa;

    double mu = H > 0.0 ? H : 0.0;
    const double floor_tol = 1e-15 * (H > 0.0 ? H : 0.0);
    int it = 0;
    bool conv = false;
    for (; it < max_iter; ++it) {
        const double p = (((mu + p3) * mu + p2) * mu + p1) * mu + p0;
        const double dp = ((4.0 * mu + 3.0 * p3) * mu + 2.0 * p2) * mu + p1;
        if (!(dp > 0.0)) {  // at or below a stationary point: stop, flag
            break;
        }
        const double step = p / dp;
        mu -= step;
        if (fabs(step) <= 1e-13 * fabs(mu) + floor_tol) {
            conv = true;
            ++it;
            break;
        }
    }
    if (aa == 0.0 && !(mu > 0.0)) {  // exactly aligned pair: mu = max(0, lmax(B))
        conv = true;
    }
    if (it > 8) conv = false;  // linear convergence: (near-)double root, re-evaluate in FP64
    ShiftedOut o;
    const double d = H - mu;
    o.rmsd2_L = 2.0 * (d > 0.0 ? d : 0.0);
    o.H = H;
    o.iterations = it;
    o.converged = conv;
    return o;
}

}  // namespace bs
