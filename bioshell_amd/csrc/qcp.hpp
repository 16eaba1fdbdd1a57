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
//  qcp_secular()   the same root, solved in a frame where no large numbers
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
//                  Used by the Gram epilogue.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace bs {

#define BS_HD __host__ __device__ __forceinline__

// a / b for the Newton steps: on the device v_rcp_f64 refined by two
// Newton-Raphson steps (relative error ~1e-16 after the second), instead of the
// IEEE division sequence; a Newton step tolerates any error far below its size.
BS_HD double newton_div(double a, double b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(b);
    double e = fma(-b, r, 1.0);
    r = fma(r, e, r);
    e = fma(-b, r, 1.0);
    r = fma(r, e, r);
    return a * r;
#else
    return a / b;
#endif
}

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
template <class T>
BS_HD void top_eigvec4(T A[4][4], T q[4], T* lambda)
{
    T V[4][4];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) V[r][c] = (r == c) ? T(1) : T(0);
    for (int sweep = 0; sweep < 30; ++sweep) {
        T off = T(0), diag = T(0);
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) {
                if (r == c)
                    diag += A[r][c] * A[r][c];
                else
                    off += A[r][c] * A[r][c];
            }
        // off-diagonal mass below 1e-24 (FP64) of the diagonal: eigenvalues are
        // then exact to ~1e-24 relative and eigenvectors to ~1e-12
        if (off <= (sizeof(T) == 8 ? T(1e-24) : T(1e-12)) * diag || off == T(0)) break;
        for (int p = 0; p < 3; ++p)
            for (int r = p + 1; r < 4; ++r) {
                const T apr = A[p][r];
                if (apr == T(0)) continue;
                const T th = (A[r][r] - A[p][p]) / (T(2) * apr);
                const T t = (th >= T(0) ? T(1) : T(-1)) / (fabs(th) + sqrt(th * th + T(1)));
                const T c = T(1) / sqrt(t * t + T(1)), s = t * c;
                for (int k = 0; k < 4; ++k) {
                    const T akp = A[k][p], akr = A[k][r];
                    A[k][p] = c * akp - s * akr;
                    A[k][r] = s * akp + c * akr;
                }
                for (int k = 0; k < 4; ++k) {
                    const T apk = A[p][k], ark = A[r][k];
                    A[p][k] = c * apk - s * ark;
                    A[r][k] = s * apk + c * ark;
                }
                for (int k = 0; k < 4; ++k) {
                    const T vkp = V[k][p], vkr = V[k][r];
                    V[k][p] = c * vkp - s * vkr;
                    V[k][r] = s * vkp + c * vkr;
                }
            }
    }
    int best = 0;
    for (int r = 1; r < 4; ++r)
        if (A[r][r] > A[best][best]) best = r;
    T nq = T(0);
    for (int r = 0; r < 4; ++r) nq += V[r][best] * V[r][best];
    nq = sqrt(nq);
    for (int r = 0; r < 4; ++r) q[r] = V[r][best] / nq;
    if (lambda) *lambda = A[best][best];
}

// Horn's key matrix of M as a 4x4 array.
template <class T>
BS_HD void key_matrix(const T* M, T K[4][4])
{
    const T xx = M[0], xy = M[1], xz = M[2], yx = M[3], yy = M[4], yz = M[5], zx = M[6],
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
    T l = E0, dp = T(0);
    const T tol = sizeof(T) == 8 ? T(1e-11) : T(1e-6);
    int it = 0;
    for (; it < max_iter; ++it) {
        const T l2 = l * l;
        const T p = ((l2 + c2) * l + c1) * l + c0;
        dp = (T(4) * l2 + T(2) * c2) * l + c1;
        if (dp == T(0)) break;
        const T step = p / dp;
        l -= step;
        if (fabs(step) <= tol * fabs(l)) {
            ++it;
            break;
        }
    }
    // Near-degenerate polish: linear convergence means a (near-) double root,
    // which the polynomial pins to ~sqrt(eps) relative only.
    // In FP32 the polynomial is too noisy near a double root for the step test
    // to be trusted, and sqrt() amplifies any error in a small E0 - lambda.
    // P'(lambda) is the product of the gaps to the other eigenvalues: small
    // relative to its terms means a near-double root.
    const T scale = fabs(T(4) * l * l * l) + fabs(T(2) * c2 * l) + fabs(c1);
    bool polish = it > 8 || fabs(dp) <= T(1e-4) * scale;
    if constexpr (sizeof(T) == 4) polish = polish || (E0 - l) < T(1e-3) * E0;
    if (polish) {
        T K[4][4], q[4], lam;
        key_matrix(M, K);
        top_eigvec4(K, q, &lam);
        l = lam;
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

// A pair's solve: 2 (H - mu) = L RMSD^2, H, and whether the root is verified.
struct PairSolve {
    double rmsd2_L;
    double H;
    bool converged;
};

// ---------------------------------------------------------------------------
// The same largest root through the Schur complement of K' = [[0, a^T], [a, B]]:
//     det(mu I - K') = det(mu I - B) * f(mu),  f(mu) = mu - a^T (mu I - B)^{-1} a.
// When S = -B is positive definite -- the usual case, since B is dominated by
// the reference's B(R0) = 2 R0 - 2 tr(R0) I, whose eigenvalues are minus twice
// the sums of pairs of R0's eigenvalues -- every eigenvalue of B is negative,
// the eigenvalues of K' interlace B's, and the largest root mu* (>= 0, as K'
// has a zero diagonal entry) is the only root of f in [0, inf).  There f is
// increasing and concave (its second derivative is -2 a^T (S + mu I)^{-3} a),
// so Newton from mu = 0 approaches mu* from below, and
//     g0(mu) = a^T (S + mu I)^{-1} a >= mu*   for every mu <= mu*
// bounds it from above: after a step from mu, mu* lies in [mu', g0(mu)].  The
// step solves (S + mu I) v = a by the adjugate (v = adj(S + mu I) a / det),
// f'(mu) = 1 + v.v, so the step and the bracket cost one 3x3 solve.  With
// residuals e small against the reference, mu* is itself a small correction
// to H (it is the gain of the residual rotation between the two
// conformations) and one step leaves a bracket of relative width
// |v|^2 ~ mu*/|S|, far below the tolerance for chains of a few hundred atoms;
// short chains take a second step.  A pair whose S is not positive definite,
// or whose bracket stays too wide, is left unconverged: the Gram path then
// re-evaluates it with the textbook form above from the coordinates.
struct SecularSystem {
    double b00, b11, b22, b01, b02, b12;  // B
    double a0, a1, a2;                    // a
    double H;
};

// B, a and H of one pair from its residual Gram E (row-major,
// E[3p+q] = sum_k e_i,kp e_j,kq) and both conformations' constants.
BS_HD SecularSystem secular_system(const double* E, const double* ci, const double* cj)
{
    SecularSystem y;
    const double trE = E[0] + E[4] + E[8];
    y.b00 = ci[0] + cj[0] + 2.0 * (E[0] - trE);
    y.b11 = ci[1] + cj[1] + 2.0 * (E[4] - trE);
    y.b22 = ci[2] + cj[2] + 2.0 * (E[8] - trE);
    y.b01 = ci[3] + cj[3] + (E[1] + E[3]);
    y.b02 = ci[4] + cj[4] + (E[2] + E[6]);
    y.b12 = ci[5] + cj[5] + (E[5] + E[7]);
    y.a0 = (cj[6] - ci[6]) + (E[5] - E[7]);
    y.a1 = (cj[7] - ci[7]) + (E[6] - E[2]);
    y.a2 = (cj[8] - ci[8]) + (E[1] - E[3]);
    y.H = (ci[9] + cj[9]) - trE;
    return y;
}

// mu* in [lo, hi], and whether S was found positive definite on the way.
struct SecularBracket {
    double H, lo, hi;
    bool pd;
};

// One Newton step of f from m (<= mu*): the new lower bound and g0(m) as the
// upper bound; pd reports S + m I positive definite (else the bracket is void).
BS_HD SecularBracket secular_step(const SecularSystem& y, double m)
{
    // S + m I = -(B - m I); adj(-X) = adj(X) for 3x3, det(-X) = -det(X)
    const double c00 = y.b00 - m, c11 = y.b11 - m, c22 = y.b22 - m;
    const double A00 = c11 * c22 - y.b12 * y.b12;
    const double A11 = c00 * c22 - y.b02 * y.b02;
    const double A22 = c00 * c11 - y.b01 * y.b01;
    const double A01 = y.b02 * y.b12 - y.b01 * c22;
    const double A02 = y.b01 * y.b12 - y.b02 * c11;
    const double A12 = y.b01 * y.b02 - c00 * y.b12;
    const double det = -(c00 * A00 + y.b01 * A01 + y.b02 * A02);  // det(S + m I)
    const double w0 = A00 * y.a0 + A01 * y.a1 + A02 * y.a2;
    const double w1 = A01 * y.a0 + A11 * y.a1 + A12 * y.a2;
    const double w2 = A02 * y.a0 + A12 * y.a1 + A22 * y.a2;
    const double aw = y.a0 * w0 + y.a1 * w1 + y.a2 * w2;  // det * g0(m)
    const double ww = w0 * w0 + w1 * w1 + w2 * w2;        // det^2 * |v|^2
    SecularBracket b;
    b.H = y.H;
    // Sylvester: leading minors of S + m I are -c00, A22, det
    b.pd = (c00 < 0.0) & (A22 > 0.0) & (det > 0.0);
    // new = m - (m - g0) / (1 + |v|^2) = (m ww + aw det) / (det^2 + ww), and
    // g0(m) = aw / det: both quotients from one reciprocal of det (det^2 + ww)
    const double den = det * det + ww;
    const double r = newton_div(1.0, det * den);
    b.lo = (m * ww + aw * det) * (det * r);
    b.hi = aw * (den * r);
    return b;
}

// The result of a bracket: mu* is taken at its midpoint; converged when S is
// positive definite and the half-width h, an error bound of H - mu, moves the
// RMSD by at most tol: h <= tol L RMSD, i.e. h^2 <= tol2L * (L RMSD^2) with
// tol2L = tol^2 L and L RMSD^2 = 2 (H - mu).  A NaN or negative h fails.
BS_HD PairSolve secular_out(const SecularBracket& b, double tol2L)
{
    PairSolve o;
    const double mu = 0.5 * (b.lo + b.hi), half = 0.5 * (b.hi - b.lo);
    const double d = b.H - mu;
    o.rmsd2_L = 2.0 * (d > 0.0 ? d : 0.0);
    o.H = b.H;
    o.converged = b.pd & (half >= 0.0) & (half * half <= tol2L * o.rmsd2_L);
    return o;
}

// The sequence the Gram epilogue runs for a pair: one secular step from 0,
// a second from the new lower bound while S is positive definite but the
// bracket too wide; a pair that still fails is left unconverged (the Gram
// path re-evaluates it in FP64).  `steps` receives the steps taken.
BS_HD PairSolve qcp_secular(const double* E, const double* ci, const double* cj, double tol2L,
                            int* steps = nullptr)
{
    const SecularSystem y = secular_system(E, ci, cj);
    SecularBracket b = secular_step(y, 0.0);
    PairSolve o = secular_out(b, tol2L);
    int n = 1;
    if (!o.converged && b.pd) {
        b = secular_step(y, b.lo);
        o = secular_out(b, tol2L);
        n = 2;
    }
    if (steps) *steps = n;
    return o;
}

}  // namespace bs
