"""CPU-side checks of the C ABI: symbols, errors, the generator, host validation.

No compute call here touches a GPU.
"""
import ctypes
import hashlib
import os
import re

import numpy as np
import pytest

import bioshell_amd as B
from bioshell_amd import _native
from tests.conftest import ROOT


def header_symbols():
    text = open(os.path.join(ROOT, "include", "bioshell_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bs_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    syms = header_symbols()
    assert len(syms) >= 30
    lib = ctypes.CDLL(_native.LIB_PATH)
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # and the Python mirror declares exactly these
    assert sorted(_native.SIGNATURES) == syms


@pytest.mark.skipif(B.device_count() > 0, reason="checks the no-device error path")
def test_default_context_without_device_is_an_error():
    """A null context selects the process-wide default context; without a GPU
    every call that needs it reports why it could not be created."""
    L = _native.lib()
    a = np.zeros((4, 3))
    out = ctypes.c_double()
    for _ in range(2):  # the failed creation is remembered, not retried
        st = L.bs_crmsd_pair_f64(None, a.ctypes.data_as(_native.c_dp), a.ctypes.data_as(_native.c_dp), 4,
                                 ctypes.byref(out))
        assert st == -2 and b"default context" in L.bs_last_error()
    h = ctypes.c_void_p()
    assert L.bs_default_context(ctypes.byref(h)) == -2 and not h.value
    x = np.zeros((1, 5, 3))
    up = np.zeros(0, np.float32)
    # nothing to compute: returns before the default context is needed
    assert L.bs_crmsd_all_vs_all(None, x.ctypes.data_as(_native.c_dp), 1, 5, 3,
                                 up.ctypes.data_as(ctypes.c_void_p), 0) == 0


def test_version_and_device_count():
    assert "gfx950" in B.version()
    assert B.device_count() >= 0


@pytest.mark.skipif(B.device_count() > 0, reason="checks the no-device error path")
def test_context_without_device_is_an_error():
    with pytest.raises(B.BioshellGpuError) as e:
        B.Context(0)
    assert e.value.status == -2
    assert "no HIP device" in str(e.value)


def test_null_and_invalid_arguments_return_status():
    L = _native.lib()
    assert L.bs_context_create(0, 0, None) == -1
    assert b"null" in L.bs_last_error()
    assert L.bs_crmsd_pair_f64(None, None, None, 3, None) == -1
    assert b"null pointer" in L.bs_last_error()
    assert L.bs_context_synchronize(None) == -1  # the asynchronous side has no default
    assert b"context is null" in L.bs_last_error()
    assert L.bs_context_last_stats(None, None) == -1
    out = np.empty((2, 3, 3))
    assert L.bs_generate_ensemble(1, 2, 3, 2.0, 1.0, 0.0, 0.0, out.ctypes.data_as(_native.c_dp)) == -1
    assert b"sigma" in L.bs_last_error()
    # error state is per call: a good call clears it
    assert L.bs_generate_ensemble(1, 2, 3, 1.0, 1.0, 0.0, 0.0, out.ctypes.data_as(_native.c_dp)) == 0
    assert L.bs_last_error() == b""


def test_python_shape_validation():
    with pytest.raises(ValueError):
        B.crmsd(np.zeros((4, 3)), np.zeros((5, 3)))
    with pytest.raises(ValueError):
        B.crmsd_all_vs_all(np.zeros((4, 3)))


# ------------------------------------------------------------ generator
def test_generator_deterministic_and_subset_stable():
    a = B.generate_ensemble(42, 30, 50, sigma=(0.5, 4.0), near_fraction=0.2)
    b = B.generate_ensemble(42, 30, 50, sigma=(0.5, 4.0), near_fraction=0.2)
    assert np.array_equal(a, b)
    c = B.generate_ensemble(42, 10, 50, sigma=(0.5, 4.0), near_fraction=0.2)
    assert np.array_equal(a[:10], c)  # conformation i depends only on (seed, i)
    d = B.generate_ensemble(43, 10, 50, sigma=(0.5, 4.0), near_fraction=0.2)
    assert not np.array_equal(c, d)


def test_generator_chain_geometry():
    """sigma = 0: every conformation is a rigid copy of the NeRF chain."""
    x = B.generate_ensemble(5, 3, 200, sigma=0.0)
    for c in x:
        bonds = np.linalg.norm(np.diff(c, axis=0), axis=1)
        assert np.allclose(bonds, 3.8, atol=1e-9)
        u = c[:-2] - c[1:-1]
        v = c[2:] - c[1:-1]
        ang = np.degrees(np.arccos((u * v).sum(1) / np.linalg.norm(u, axis=1) / np.linalg.norm(v, axis=1)))
        assert ang.min() >= 85 - 1e-6 and ang.max() <= 150 + 1e-6
    # rigid copies: identical inter-atomic distances
    d0 = np.linalg.norm(x[0][:, None] - x[0][None], axis=2)
    d1 = np.linalg.norm(x[1][:, None] - x[1][None], axis=2)
    assert np.abs(d0 - d1).max() < 1e-9
    # translations within the +-50 A box (centroid of a rotated centred chain + t)
    cen = x.mean(axis=1)
    assert np.abs(cen).max() < 50 + np.abs(x[0] - x[0].mean(0)).max()


def test_generator_noise_levels():
    x = B.generate_ensemble(9, 400, 100, sigma=(2.0, 2.0), near_fraction=0.25, near_sigma=0.01)
    base = B.generate_ensemble(9, 1, 100, sigma=0.0, near_fraction=0.0)[0]
    from oracle import oracle as O

    r = np.array([O.crmsd(c, base) for c in x])
    near = r < 0.1
    assert 0.15 < near.mean() < 0.35
    assert np.all(r[near] < 0.05)
    # iid N(0, s^2) per coordinate: RMSD to base ~ s sqrt(3) (slightly less after fitting)
    assert np.median(r[~near]) == pytest.approx(2.0 * np.sqrt(3), rel=0.1)


def test_generator_golden(golden_config0):
    x = B.generate_ensemble(0, 200, 76, sigma=1.5)
    assert np.array_equal(x[:2, :4], golden_config0["first_atoms"])
    assert hashlib.sha256(x.tobytes()).digest() == golden_config0["sha256"].tobytes()


# ------------------------------------------------------------ host QCP builds
def _hook(name, res, args):
    fn = getattr(_native.diag_lib(), name)
    fn.restype = res
    fn.argtypes = args
    return fn


def test_host_build_of_device_qcp_matches_oracle():
    from oracle import oracle as O

    dp = ctypes.POINTER(ctypes.c_double)
    f64 = _hook("bsi_qcp_rmsd_f64", ctypes.c_double, [dp, ctypes.c_double, ctypes.c_double, ctypes.c_int])
    rng = np.random.default_rng(0)
    for L in (2, 3, 7, 76):
        for _ in range(5):
            a = rng.normal(size=(L, 3)) * 5
            b = a + rng.normal(size=(L, 3))
            ac, bc = O.centre(a)[0], O.centre(b)[0]
            M, ga, gb = O.covariance(ac, bc)
            M = np.ascontiguousarray(M.ravel())
            assert f64(M.ctypes.data_as(dp), ga, gb, L) == pytest.approx(O.crmsd(a, b), abs=1e-9)


def _kabsch(P, Q):
    """Rotation R minimising |P - Q R^T| for centred P, Q (numpy SVD)."""
    U, _, Vt = np.linalg.svd(Q.T @ P)
    d = np.sign(np.linalg.det(Vt.T @ U.T))
    return Vt.T @ np.diag([1.0, 1.0, d]) @ U.T


def _shift_consts(e, r):
    C, R0, g = r.T @ e, r.T @ r, (e * e).sum()
    t0, tc = np.trace(R0), np.trace(C)
    return np.array([
        R0[0, 0] - t0 + 2 * (C[0, 0] - tc), R0[1, 1] - t0 + 2 * (C[1, 1] - tc),
        R0[2, 2] - t0 + 2 * (C[2, 2] - tc), R0[0, 1] + C[0, 1] + C[1, 0],
        R0[0, 2] + C[0, 2] + C[2, 0], R0[1, 2] + C[1, 2] + C[2, 1],
        C[1, 2] - C[2, 1], C[2, 0] - C[0, 2], C[0, 1] - C[1, 0], g / 2, np.sqrt(g)])


@pytest.mark.parametrize("L,sigma,aligned", [(300, (0.5, 4.0), True), (76, (1.5, 1.5), True), (30, (0.5, 3.0), True),
                                             (12, (4.0, 8.0), True), (3, (0.5, 3.0), True), (50, (0.5, 4.0), False)])
def test_host_build_of_secular_qcp_matches_oracle(L, sigma, aligned):
    """The Gram epilogue's solve: Newton on the secular form of the QCP
    polynomial from mu = 0, a second step when the bracket is wide; a pair it
    cannot verify (S = -B not positive definite, or the bracket still wide) is
    left unconverged for the FP64 re-evaluation.  Exact residual Grams (FP64)
    against the oracle's cRMSD; with conformations fitted onto their mean (the
    device's frame) long chains converge in one step, short ones in two, and
    with any reference every converged pair is right."""
    from oracle import oracle as O

    dp = ctypes.POINTER(ctypes.c_double)
    fn = _hook("bsi_qcp_secular", None, [dp, dp, dp, ctypes.c_double, dp])
    n = 40
    x = B.generate_ensemble(17 + L, n, L, sigma=sigma, near_fraction=0.1)
    xc = np.array([a - a.mean(0) for a in x])
    if aligned:
        fit0 = np.array([c @ _kabsch(xc[0], c).T for c in xc])
        ref = fit0.mean(0)
        e = np.array([c @ _kabsch(ref, c).T for c in xc]) - ref
    else:  # any reference is exact, only slower: noisy and unaligned
        ref = xc[0] + np.random.default_rng(1).normal(size=xc[0].shape)
        ref -= ref.mean(0)
        e = xc - ref
    cs = [np.ascontiguousarray(_shift_consts(ei, ref)) for ei in e]
    tol2L = (2e-5 / 4) ** 2 * L
    out = np.empty(4)
    steps, conv = [], []
    for i in range(n):
        for j in range(i + 1, n):
            E = np.ascontiguousarray((e[i].T @ e[j]).ravel())
            fn(E.ctypes.data_as(dp), cs[i].ctypes.data_as(dp), cs[j].ctypes.data_as(dp), tol2L,
               out.ctypes.data_as(dp))
            steps.append(out[2])
            want = O.crmsd(x[i], x[j])
            if out[3] == 1.0:
                assert np.sqrt(out[0] / L) == pytest.approx(want, abs=2e-5 / 4 + 1e-9), (i, j, out)
            conv.append(out[3] == 1.0)
    steps, conv = np.array(steps), np.array(conv)
    print(f"L={L}: one step {np.mean(steps == 1):.3f}, two {np.mean(steps == 2):.3f}, "
          f"unconverged {np.mean(~conv):.3f}")
    if aligned and L >= 76:
        assert np.all(steps == 1) and conv.all()
    if aligned and L >= 30:
        assert conv.all()


def test_host_build_of_fast_rotation():
    """The preparation's rotation (Newton on the key quartic + adjugate column)
    is a proper rotation, optimal for well-posed sets (its fit RMSD equals the
    oracle's cRMSD), and still a proper rotation for degenerate sets."""
    from oracle import oracle as O

    dp = ctypes.POINTER(ctypes.c_double)
    fn = _hook("bsi_rotation_fast", None, [dp, dp])
    rng = np.random.default_rng(5)

    def rand_rot(g):
        q = g.normal(size=4)
        q /= np.linalg.norm(q)
        a, b, c, d = q
        return np.array([[a*a+b*b-c*c-d*d, 2*(b*c-a*d), 2*(b*d+a*c)],
                         [2*(b*c+a*d), a*a-b*b+c*c-d*d, 2*(c*d-a*b)],
                         [2*(b*d-a*c), 2*(c*d+a*b), a*a-b*b-c*c+d*d]])

    def rot_of(a, b):
        ac, bc = a - a.mean(0), b - b.mean(0)
        M = np.ascontiguousarray(bc.T @ ac)  # sum b a^T
        R = np.empty(9)
        fn(M.ravel().ctypes.data_as(dp), R.ctypes.data_as(dp))
        R = R.reshape(3, 3)
        assert np.abs(R @ R.T - np.eye(3)).max() < 1e-12
        assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-12)
        return R, ac, bc

    for L in (3, 4, 10, 76, 300):
        for _ in range(20):
            a = rng.normal(size=(L, 3)) * 5
            b = a @ rand_rot(rng).T + rng.normal(size=(L, 3)) * rng.uniform(0.01, 3.0)
            R, ac, bc = rot_of(a, b)
            fit = np.sqrt(((bc @ R.T - ac) ** 2).sum() / L)
            assert fit == pytest.approx(O.crmsd(a, b), abs=1e-9)
    # Newton started from E0 = (G_a + G_b) / 2 (the preparation kernels' bound),
    # and from E0 rounded 1e-13 below lambda_max: the same optimal rotation
    fn_from = _hook("bsi_rotation_fast_from", None, [dp, ctypes.c_double, dp])
    for L in (3, 10, 76, 300):
        for _ in range(10):
            a = rng.normal(size=(L, 3)) * 5 + rng.normal(size=3) * 100
            b = a @ rand_rot(rng).T + rng.normal(size=(L, 3)) * rng.uniform(0.01, 3.0)
            R0, ac, bc = rot_of(a, b)
            M = np.ascontiguousarray(bc.T @ ac)
            e0 = 0.5 * ((ac ** 2).sum() + (bc ** 2).sum())
            lam = e0 - 0.5 * L * O.crmsd(a, b) ** 2
            for upper in (e0, lam * (1 - 1e-13)):
                R = np.empty(9)
                fn_from(M.ravel().ctypes.data_as(dp), upper, R.ctypes.data_as(dp))
                R = R.reshape(3, 3)
                assert np.abs(R - R0).max() < 1e-9
                fit = np.sqrt(((bc @ R.T - ac) ** 2).sum() / L)
                assert fit == pytest.approx(O.crmsd(a, b), abs=1e-9)
    # degenerate: M = 0 (one atom), collinear pairs, coplanar sets, identical sets
    for a, b in [(np.zeros((1, 3)), np.ones((1, 3))),
                 (np.array([[0.0, 0, 0], [3.8, 0, 0]]), np.array([[1.0, 2, 3], [1.0, 5.8, 3]])),
                 (np.c_[rng.normal(size=(8, 2)), np.zeros(8)], np.c_[rng.normal(size=(8, 2)), np.zeros(8)]),
                 (np.arange(12.0).reshape(4, 3) ** 1.5, np.arange(12.0).reshape(4, 3) ** 1.5)]:
        R, ac, bc = rot_of(a, b)
        fit = np.sqrt(((bc @ R.T - ac) ** 2).sum() / len(a))
        assert fit <= O.crmsd(a, b) + 1e-6 or len(a) < 3
