"""Pins the FP64 CPU oracle (oracle/crmsd_oracle.c) before anything is compared to it.

The reference has no crmsd to run (bioshell-pdb/Cargo.toml:27 only announces
it), so each step of the oracle is fixed by something other than its own
formula: closed forms, an independent eigen-solver, an independent SVD
(numpy), brute force over rotations on tiny inputs, exact integer arithmetic,
and golden vectors computed by numpy from the reference's own 2jqb fixture.
"""
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O


def rand_rot(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([
        [a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
        [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
        [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d],
    ])


def numpy_kabsch(a, b):
    a = a - a.mean(0)
    b = b - b.mean(0)
    M = a.T @ b
    s = np.linalg.svd(M, compute_uv=False)
    d = np.sign(np.linalg.det(M)) or 1.0
    lam = s[0] + s[1] + d * s[2]
    return np.sqrt(max(0.0, ((a * a).sum() + (b * b).sum() - 2 * lam) / len(a)))


# ------------------------------------------------------------ eigenvalue step
@pytest.mark.parametrize("seed", range(20))
def test_qcp_lambda_matches_jacobi_and_svd(seed):
    rng = np.random.default_rng(seed)
    M = rng.normal(size=(3, 3)) * rng.uniform(0.1, 1000)
    if seed % 3 == 0:
        M[:, 0] *= -1  # reflections: det M < 0 exercises the sign term
    if seed % 5 == 0:
        M[2] = M[0] + M[1]  # rank-deficient
    E0 = np.sqrt((M * M).sum()) * 3 + 1.0  # any upper bound of lambda_max works
    lam, _ = O.qcp_lambda(M, E0)
    K = O.key_matrix(M)
    assert np.allclose(K, K.T)
    assert abs(np.trace(K)) < 1e-9 * np.abs(M).max()
    ev = np.linalg.eigvalsh(K)  # numpy, independent of both
    s = np.linalg.svd(M, compute_uv=False)
    d = np.sign(np.linalg.det(M)) or 1.0
    tol = 1e-9 * np.abs(M).sum()
    assert lam == pytest.approx(ev[-1], abs=tol)
    assert O.lambda_jacobi4(M) == pytest.approx(ev[-1], abs=tol)
    assert O.lambda_kabsch(M) == pytest.approx(s[0] + s[1] + d * s[2], abs=tol)
    assert lam == pytest.approx(s[0] + s[1] + d * s[2], abs=tol)


def test_key_matrix_frobenius_identity():
    # |K|_F^2 = 4 |M|_F^2: the c2 = -2|M|_F^2 coefficient of the quartic rests on it
    rng = np.random.default_rng(1)
    for _ in range(10):
        M = rng.normal(size=(3, 3))
        K = O.key_matrix(M)
        assert (K * K).sum() == pytest.approx(4 * (M * M).sum(), rel=1e-12)


# ------------------------------------------------------------ centring / cov
def test_centre_sum_then_divide():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(77, 3)) * 10 + np.array([50.0, -50.0, 25.0])
    c, cen = O.centre(x)
    assert np.allclose(cen, x.sum(0) / 77, rtol=0, atol=1e-12)
    assert np.abs(c.sum(0)).max() <= 1e-12 * 77 * np.abs(x).max()


def test_covariance_exact_on_integers():
    rng = np.random.default_rng(3)
    a = rng.integers(-40, 40, size=(31, 3)).astype(float)
    b = rng.integers(-40, 40, size=(31, 3)).astype(float)
    M, ga, gb = O.covariance(a, b)
    for p in range(3):
        for q in range(3):
            exact = sum(Fraction(int(a[k, p])) * Fraction(int(b[k, q])) for k in range(31))
            assert M[p, q] == float(exact)
    assert ga == float(sum(int(v) ** 2 for v in a.ravel()))
    assert gb == float(sum(int(v) ** 2 for v in b.ravel()))
    # row-major convention: M[p, q] = sum a_p b_q, not its transpose
    assert np.array_equal(M, a.T @ b)


# ------------------------------------------------------------ closed forms
@pytest.mark.parametrize("L", [3, 10, 76, 300])
def test_rigid_copy_has_zero_rmsd(L):
    rng = np.random.default_rng(L)
    a = rng.normal(size=(L, 3)) * 8
    b = a @ rand_rot(rng).T + rng.uniform(-50, 50, 3)
    # RMSD = sqrt(2 (E0 - lambda) / L): an FP64 rounding of ~1e-16 lambda in
    # lambda becomes ~1e-6 A after the square root, so "zero" means < 1e-5 here
    assert O.crmsd(a, b) < 1e-5
    assert O.crmsd(a, a) < 1e-5


def test_symmetry_and_rigid_invariance():
    rng = np.random.default_rng(5)
    for _ in range(10):
        a = rng.normal(size=(40, 3)) * 6
        b = a + rng.normal(size=a.shape)
        r = O.crmsd(a, b)
        assert O.crmsd(b, a) == pytest.approx(r, abs=1e-10)
        a2 = a @ rand_rot(rng).T + rng.uniform(-50, 50, 3)
        b2 = b @ rand_rot(rng).T + rng.uniform(-50, 50, 3)
        assert O.crmsd(a2, b2) == pytest.approx(r, abs=1e-9)


def test_two_atoms_closed_form():
    rng = np.random.default_rng(6)
    for _ in range(20):
        a = rng.normal(size=(2, 3)) * 5
        b = rng.normal(size=(2, 3)) * 5
        da = np.linalg.norm(a[0] - a[1])
        db = np.linalg.norm(b[0] - b[1])
        assert O.crmsd(a, b) == pytest.approx(abs(da - db) / 2, abs=1e-9)


def test_one_and_zero_atoms():
    assert O.crmsd(np.ones((1, 3)), np.zeros((1, 3))) == 0.0
    assert np.isnan(O.crmsd(np.zeros((0, 3)), np.zeros((0, 3))))


def test_planar_sets_closed_form():
    """Planar sets: proper 2-D rotation or a mirror (a 3-D rotation by pi about an in-plane axis)."""
    rng = np.random.default_rng(7)
    for _ in range(20):
        n = rng.integers(3, 30)
        a = np.c_[rng.normal(size=(n, 2)) * 5, np.zeros(n)]
        b = np.c_[rng.normal(size=(n, 2)) * 5, np.zeros(n)]
        ac, bc = a - a.mean(0), b - b.mean(0)
        P = (ac[:, 0] * bc[:, 0] + ac[:, 1] * bc[:, 1]).sum()
        Q = (ac[:, 0] * bc[:, 1] - ac[:, 1] * bc[:, 0]).sum()
        Pm = (ac[:, 0] * bc[:, 0] - ac[:, 1] * bc[:, 1]).sum()
        Qm = (-ac[:, 0] * bc[:, 1] - ac[:, 1] * bc[:, 0]).sum()
        lam = max(np.hypot(P, Q), np.hypot(Pm, Qm))
        g = (ac ** 2).sum() + (bc ** 2).sum()
        expect = np.sqrt(max(0.0, (g - 2 * lam) / n))
        assert O.crmsd(a, b) == pytest.approx(expect, abs=1e-9)


def test_uncorrelated_bound():
    rng = np.random.default_rng(8)
    for _ in range(10):
        a = rng.normal(size=(50, 3)) * 7
        b = rng.normal(size=(50, 3)) * 3
        ac, bc = a - a.mean(0), b - b.mean(0)
        assert O.crmsd(a, b) <= np.sqrt(((ac ** 2).sum() + (bc ** 2).sum()) / 50) + 1e-12


def test_brute_force_tiny_sets():
    """Minimum over rotations by dense sampling + local refinement equals QCP."""
    from scipy.optimize import minimize

    rng = np.random.default_rng(9)
    for L in (3, 4, 5, 6):
        a = rng.normal(size=(L, 3)) * 3
        b = rng.normal(size=(L, 3)) * 3
        ac, bc = a - a.mean(0), b - b.mean(0)

        def f(q):
            q = q / np.linalg.norm(q)
            w, x, y, z = q
            R = np.array([
                [w * w + x * x - y * y - z * z, 2 * (x * y - w * z), 2 * (x * z + w * y)],
                [2 * (x * y + w * z), w * w - x * x + y * y - z * z, 2 * (y * z - w * x)],
                [2 * (x * z - w * y), 2 * (y * z + w * x), w * w - x * x - y * y + z * z],
            ])
            return ((ac - bc @ R.T) ** 2).sum() / L

        qs = rng.normal(size=(4000, 4))
        vals = np.array([f(q) for q in qs])
        best = min((minimize(f, qs[k], method="Nelder-Mead", options={"xatol": 1e-12, "fatol": 1e-15,
                                                                      "maxiter": 20000}).fun
                    for k in np.argsort(vals)[:5]))
        r = O.crmsd(a, b)
        assert r * r <= vals.min() + 1e-12  # no sampled rotation beats it
        assert np.sqrt(max(best, 0)) == pytest.approx(r, abs=1e-6)


# ------------------------------------------------------------ superposition
def test_superpose_convention_and_fit():
    rng = np.random.default_rng(10)
    for _ in range(10):
        a = rng.normal(size=(30, 3)) * 5
        b = (a + rng.normal(size=a.shape) * 0.3) @ rand_rot(rng).T + rng.uniform(-50, 50, 3)
        R, t, r = O.superpose(a, b)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-12)
        assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-12)
        fit = b @ R.T + t  # x' = R x + t, applied to b
        assert np.sqrt(((a - fit) ** 2).sum(1).mean()) == pytest.approx(r, abs=1e-9)
        assert r == pytest.approx(O.crmsd(a, b), abs=1e-9)


# ------------------------------------------------------------ golden vectors
def test_golden_2jqb(golden_2jqb):
    for key in ("ca", "all_atom"):
        x = golden_2jqb[key]
        got = O.all_vs_all(x, nthreads=2)
        assert np.abs(got - golden_2jqb[key + "_rmsd"]).max() < 1e-9


def test_golden_config0(golden_config0):
    import bioshell_amd as B

    x = B.generate_ensemble(0, 200, 76, sigma=1.5)
    got = O.all_vs_all(x, nthreads=4)
    want = golden_config0["rmsd"]
    assert got.shape == want.shape == (19900,)
    assert np.abs(got - want).max() < 1e-9


def test_rows_driver_matches_full():
    rng = np.random.default_rng(11)
    x = rng.normal(size=(23, 9, 3)) * 4
    full = O.all_vs_all(x, nthreads=3)
    rows = O.all_vs_all_rows(x, 5, 12, nthreads=2)
    for i in range(5, 12):
        for j in range(i + 1, 23):
            assert rows[i - 5, j] == full[O.tri_index(23, i, j)]
    q = O.query_vs_library(x[:4], x)
    for i in range(4):
        for j in range(i + 1, 23):
            assert q[i, j] == pytest.approx(full[O.tri_index(23, i, j)], abs=1e-12)


# ------------------------------------------------------------ sensitivity
def test_checks_reject_planted_mistakes():
    """The checks above fail for the plausible mistakes they are meant to catch."""
    rng = np.random.default_rng(12)
    M = rng.normal(size=(3, 3))
    M[:, 0] *= -1 if np.linalg.det(M) > 0 else 1  # make det M < 0
    s = np.linalg.svd(M, compute_uv=False)
    good = s[0] + s[1] - s[2]
    dropped_sign = s[0] + s[1] + s[2]  # a dropped sign(det) term
    assert abs(good - dropped_sign) > 1e-3
    assert O.lambda_kabsch(M) == pytest.approx(good, abs=1e-10)
    # a transposed covariance gives the same lambda (K(M^T) is similar to K(M))
    # but a transposed operand in the superposition would not fit:
    a = rng.normal(size=(20, 3)) * 4
    b = a @ rand_rot(rng).T + 3.0
    R, t, _ = O.superpose(a, b)
    assert np.abs(b @ R.T + t - a).max() < 1e-9
    assert np.abs(b @ R + t - a).max() > 1e-3
