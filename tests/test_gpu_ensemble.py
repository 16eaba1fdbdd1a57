"""Ensemble superposition on the GPU (bs_superpose_ensemble) against the
oracle's FP64 superposition, and ingestion -> cRMSD -> superposition -> PDB
end to end.

Tolerances (FP64 throughout): rotations and coordinates to 1e-9 (relative to
the coordinate scale) where the fit is unique; squared RMSD to 1e-11 of the
mean squared radius (the RMSD of a near-perfect fit is the square root of a
difference of two sums, so its absolute error grows as it approaches zero).
"""
import os

import numpy as np
import pytest

import bioshell_amd as B
from bioshell_amd.crmsd import BioshellGpuError
from oracle import oracle as O
from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu


def check_fit(fit, x, ref, unique=True):
    n, L, _ = x.shape
    scale = max(1.0, float(np.abs(x).max()), float(np.abs(ref).max()))
    rc = ref - ref.mean(0)
    for m in range(n):
        R0, t0, r0 = O.superpose(ref, x[m])
        xc = x[m] - x[m].mean(0)
        ms = ((rc ** 2).sum() + (xc ** 2).sum()) / (2 * L)
        assert abs(fit.rmsd[m] ** 2 - r0 ** 2) <= 1e-11 * max(1.0, ms), (m, fit.rmsd[m], r0)
        R = fit.rotation[m]
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-12) and np.linalg.det(R) == pytest.approx(1.0, abs=1e-12)
        if unique:
            assert np.allclose(R, R0, atol=1e-9)
            assert np.allclose(fit.translation[m], t0, atol=1e-9 * scale)
        if fit.superposed is not None:
            assert np.allclose(fit.superposed[m], x[m] @ R.T + fit.translation[m], atol=1e-12 * scale, rtol=0)
            resid2 = ((fit.superposed[m] - ref) ** 2).sum(1).mean()
            assert abs(resid2 - r0 ** 2) <= 1e-11 * max(1.0, ms) + 1e-16 * scale ** 2


@pytest.mark.parametrize("n,L,seed", [(5, 3, 1), (17, 5, 2), (200, 76, 3), (64, 300, 4), (9, 1100, 5),
                                      (33, 129, 6), (21, 256, 7), (10, 600, 8), (5, 2048, 9), (3, 2600, 10)])
def test_matches_oracle(gpu_ctx, n, L, seed):
    x = B.generate_ensemble(seed, n, L, sigma=(0.5, 3.0))
    fit = B.superpose_ensemble(x, ref_index=n // 2, ctx=gpu_ctx)
    check_fit(fit, x, x[n // 2])
    assert fit.rmsd[n // 2] < 1e-5 and np.allclose(fit.rotation[n // 2], np.eye(3), atol=1e-9)
    ref = B.generate_ensemble(seed + 100, 1, L, sigma=(1.0, 1.0))[0]
    fit = B.superpose_ensemble(x, reference=ref, ctx=gpu_ctx)
    check_fit(fit, x, ref)


def test_rigid_copies_land_on_the_reference(gpu_ctx):
    rng = np.random.default_rng(11)
    ref = rng.normal(size=(123, 3)) * 8
    Q = [np.linalg.qr(rng.normal(size=(3, 3)))[0] for _ in range(40)]
    Q = [q if np.linalg.det(q) > 0 else -q for q in Q]
    x = np.stack([ref @ q.T + rng.uniform(-50, 50, 3) for q in Q])
    fit = B.superpose_ensemble(x, reference=ref, ctx=gpu_ctx)
    assert np.abs(fit.superposed - ref).max() < 1e-10 * 60
    assert fit.rmsd.max() < 1e-5
    for m, q in enumerate(Q):
        assert np.allclose(fit.rotation[m], q.T, atol=1e-12)


def test_small_and_degenerate_shapes(gpu_ctx):
    # one atom: any rotation is optimal, the fit is exact
    x = np.random.default_rng(2).normal(size=(6, 1, 3)) * 10
    fit = B.superpose_ensemble(x, ref_index=0, ctx=gpu_ctx)
    assert np.allclose(fit.superposed, x[0], atol=1e-12) and fit.rmsd.max() == 0.0
    # two atoms: the rotation about the pair's axis is free; RMSD = |d_a - d_b| / 2
    x = np.random.default_rng(3).normal(size=(8, 2, 3)) * 4
    fit = B.superpose_ensemble(x, ref_index=0, ctx=gpu_ctx)
    check_fit(fit, x, x[0], unique=False)
    d = np.linalg.norm(x[:, 0] - x[:, 1], axis=1)
    assert np.allclose(fit.rmsd, np.abs(d - d[0]) / 2, atol=1e-9)
    # identical conformations and a collinear set
    line = np.outer(np.arange(10.0), [1.0, 2.0, 3.0])
    x = np.stack([line, line + 5.0, line * 1.0])
    fit = B.superpose_ensemble(x, ref_index=0, ctx=gpu_ctx)
    check_fit(fit, x, x[0], unique=False)
    assert fit.rmsd.max() < 1e-6


def test_outputs_are_optional_and_errors(gpu_ctx):
    x = B.generate_ensemble(4, 10, 30)
    fit = B.superpose_ensemble(x, ref_index=3, coordinates=False, ctx=gpu_ctx)
    assert fit.superposed is None
    full = B.superpose_ensemble(x, ref_index=3, ctx=gpu_ctx)
    assert np.array_equal(fit.rotation, full.rotation) and np.array_equal(fit.rmsd, full.rmsd)
    with pytest.raises(BioshellGpuError):
        B.superpose_ensemble(x, ref_index=10, ctx=gpu_ctx)
    with pytest.raises(ValueError):
        B.superpose_ensemble(x, reference=np.zeros((29, 3)), ctx=gpu_ctx)


def test_device_entry_matches_host(gpu_ctx):
    import torch

    x = B.generate_ensemble(8, 300, 150, sigma=(0.5, 3.0))
    host = B.superpose_ensemble(x, ref_index=7, ctx=gpu_ctx)
    d_x = torch.from_numpy(x).cuda()
    d_rot = torch.empty((300, 9), dtype=torch.float64, device="cuda")
    d_tr = torch.empty((300, 3), dtype=torch.float64, device="cuda")
    d_rm = torch.empty(300, dtype=torch.float64, device="cuda")
    d_sp = torch.empty_like(d_x)
    B.superpose_ensemble_device(d_x.data_ptr(), 300, 150, None, 7, d_rot.data_ptr(), d_tr.data_ptr(),
                                d_rm.data_ptr(), d_sp.data_ptr(), ctx=gpu_ctx)
    gpu_ctx.synchronize()
    assert np.array_equal(d_rot.cpu().numpy().reshape(300, 3, 3), host.rotation)
    assert np.array_equal(d_tr.cpu().numpy(), host.translation)
    assert np.array_equal(d_rm.cpu().numpy(), host.rmsd)
    assert np.array_equal(d_sp.cpu().numpy(), host.superposed)


def test_file_to_superposed_pdb(gpu_ctx, golden_2jqb, tmp_path):
    """2jqb read natively, cRMSD against the numpy-SVD golden values, every
    model fitted onto model 0, written as PDB and read back."""
    path = os.path.join(GOLDEN, "structures", "2jqb.cif.gz")
    for select, key in (("ca", "ca"), ("all", "all_atom")):
        e = B.load_ensemble(path, select=select)
        assert np.array_equal(e.xyz, golden_2jqb[key])
        m = B.crmsd_all_vs_all(e.xyz, B.Method.FP64_PLAIN, out_f64=True, ctx=gpu_ctx)
        assert np.abs(m.upper - golden_2jqb[key + "_rmsd"]).max() < 1e-9
        fit = B.superpose_ensemble(e.xyz, ref_index=0, ctx=gpu_ctx)
        check_fit(fit, e.xyz, e.xyz[0])
        assert np.allclose(fit.rmsd[1:], m.upper[:19], atol=1e-9)
        # a rigid motion per model leaves every pairwise cRMSD unchanged
        m2 = B.crmsd_all_vs_all(fit.superposed, B.Method.FP64_PLAIN, out_f64=True, ctx=gpu_ctx)
        assert np.abs(m2.upper - m.upper).max() < 1e-9
        out = tmp_path / f"2jqb_{select}.pdb"
        e.write_pdb(str(out), fit.superposed)
        back = B.load_ensemble(str(out), select="all")
        assert back.count_models() == 20 and back.n_atoms == e.n_atoms
        assert np.abs(back.xyz - fit.superposed).max() <= 0.0005 + 1e-9  # three decimals in the file


def test_upload(gpu_ctx):
    import torch

    e = B.load_ensemble(os.path.join(GOLDEN, "structures", "2jqb.cif.gz"), select="ca")
    d = torch.empty((20, 15, 3), dtype=torch.float64, device="cuda")
    e.upload(d.data_ptr(), ctx=gpu_ctx)
    assert np.array_equal(d.cpu().numpy(), e.xyz)
