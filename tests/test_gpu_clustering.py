"""Device clustering and row reductions against the oracle's restatement of
BioShell's clustering (bit-identical merge sequences and reductions)."""
import ctypes
import os
import time

import numpy as np
import pytest

import bioshell_amd as B
from bioshell_amd import _native
from bioshell_amd import clustering as C
from oracle import oracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def device_merges(ctx, up, n, linkage):
    m = max(n - 1, 1)
    mi = np.zeros(m, np.uint32)
    mj = np.zeros(m, np.uint32)
    md = np.zeros(m, np.float32)
    B.check(_native.lib().bs_hierarchical_clustering(
        ctx.handle, np.ascontiguousarray(up, np.float32).ctypes.data_as(_native.c_fp), n,
        C.LINKAGES.index(linkage), mi.ctypes.data_as(_native.c_u32p), mj.ctypes.data_as(_native.c_u32p),
        md.ctypes.data_as(_native.c_fp)))
    return mi[:n - 1], mj[:n - 1], md[:n - 1]


def random_upper(rng, n, kind):
    k = n * (n - 1) // 2
    if kind == "uniform":
        return rng.uniform(0.1, 9.0, k).astype(np.float32)
    if kind == "ties":  # few distinct values: ties decide the order
        return rng.integers(1, 5, k).astype(np.float32)
    pts = rng.normal(size=(n, 3)) * 4
    D = np.sqrt(((pts[:, None] - pts[None]) ** 2).sum(-1)).astype(np.float32)
    return np.array([D[i, j] for i in range(n) for j in range(i + 1, n)], dtype=np.float32)


@pytest.mark.parametrize("linkage", C.LINKAGES)
@pytest.mark.parametrize("n,kind", [(2, "uniform"), (3, "ties"), (17, "uniform"), (64, "ties"),
                                    (65, "euclid"), (203, "uniform"), (150, "ties"),
                                    (1500, "ties"), (1100, "euclid")])
def test_merges_bit_identical_to_oracle(gpu_ctx, linkage, n, kind):
    rng = np.random.default_rng(n * 7 + len(kind))
    up = random_upper(rng, n, kind)
    got = device_merges(gpu_ctx, up, n, linkage)
    want = O.hclust(up, n, linkage)
    assert np.array_equal(got[0], want[0].astype(np.uint32)), "merge i"
    assert np.array_equal(got[1], want[1].astype(np.uint32)), "merge j"
    assert np.array_equal(got[2].view(np.uint32), want[2].view(np.uint32)), "merge distance bits"


@pytest.mark.parametrize("linkage", ["average", "centroid", "wards"])
def test_moved_cluster_hand_trace_on_device(gpu_ctx, linkage):
    """The hand-traced six-element case (a merged cluster moved, then merged
    again with the size left at its new index) gives the traced heights on the
    device too."""
    from tests.test_clustering_oracle import _average, _centroid, _hand_trace, _six, _wards

    rule = {"average": _average, "centroid": _centroid, "wards": _wards}[linkage]
    want = _hand_trace(rule)
    D = _six()
    mi, mj, md = device_merges(gpu_ctx, D[np.triu_indices(6, 1)], 6, linkage)
    assert list(zip(mi.tolist(), mj.tolist())) == [(i, j) for i, j, _ in want]
    assert np.array_equal(md.view(np.uint32), np.array([h for _, _, h in want], np.float32).view(np.uint32))


@pytest.mark.parametrize("linkage", ["single", "average", "wards"])
def test_rmsd_matrix_clustering_and_device_entry(gpu_ctx, linkage):
    """An all-vs-all RMSD matrix: host triangle and HBM triangle give the same
    merges as the oracle; the tree has every element once."""
    import torch

    n, L = 300, 60
    x = B.generate_ensemble(77, n, L, sigma=(0.3, 3.0), near_fraction=0.1)
    m = B.crmsd_all_vs_all(x, B.Method.BF16X3_GRAM, ctx=gpu_ctx)
    root = C.hierarchical_clustering(m, linkage=linkage, ctx=gpu_ctx)
    mi, mj, md = O.hclust(m.upper, n, linkage)
    ref = C.tree_from_merges(n, mi, mj, md)
    assert C.retrieve_data_id(root) == C.retrieve_data_id(ref)
    assert sorted(C.retrieve_data_id(root)) == list(range(n))
    d_x = torch.from_numpy(x).cuda()
    d_up = torch.empty(n * (n - 1) // 2, dtype=torch.float32, device="cuda")
    B.crmsd_all_vs_all_device(d_x.data_ptr(), n, L, d_up.data_ptr(), B.Method.BF16X3_GRAM, ctx=gpu_ctx)
    gpu_ctx.synchronize()
    root2 = C.hierarchical_clustering(n, linkage=linkage, ctx=gpu_ctx, device_upper=d_up.data_ptr())
    assert C.retrieve_data_id(root2) == C.retrieve_data_id(root)


def test_reference_tests_through_the_device(gpu_ctx):
    """test_hierarchical.rs cluster_numbers / cluster_letters / d5.tsv"""
    data = np.array([6.0, 2.0, 2.5, 5.9], dtype=np.float32)
    d = lambda a, b: float(np.abs(data[a] - data[b]))
    root = C.hierarchical_clustering(4, d, "single", ctx=gpu_ctx)
    assert [float(v) for v in C.retrieve_data(root, data)] == [6.0, np.float32(5.9), 2.0, 2.5]
    C.balance_clustering_tree(root, d)
    assert [float(v) for v in C.retrieve_data(root, data)] == [6.0, np.float32(5.9), 2.5, 2.0]
    letters = ['A', 'Z', 'D', 'Y', 'C', 'X', 'B', 'W', 'F', 'V']
    dl = lambda a, b: float(abs(ord(letters[a]) - ord(letters[b])))
    root = C.hierarchical_clustering(len(letters), dl, "single", ctx=gpu_ctx)
    assert C.retrieve_data(root, letters) == ['A', 'B', 'D', 'C', 'F', 'Z', 'Y', 'X', 'W', 'V']
    C.balance_clustering_tree(root, dl)
    assert C.retrieve_data(root, letters) == ['A', 'B', 'C', 'D', 'F', 'V', 'W', 'X', 'Y', 'Z']


def test_row_stats_outliers_medoid(gpu_ctx):
    n, L = 700, 45
    x = B.generate_ensemble(5, n, L, sigma=(0.5, 3.0), near_fraction=0.05)
    m = B.crmsd_all_vs_all(x, B.Method.BF16X3_GRAM, ctx=gpu_ctx)
    lo, hi = C.row_stats(m, ctx=gpu_ctx)
    wlo, whi = O.row_stats(m.upper, n)
    assert np.array_equal(lo, wlo) and np.array_equal(hi, whi)
    cut = float(np.median(wlo))
    assert C.retrieve_outliers(m, cutoff=cut, ctx=gpu_ctx) == [i for i in range(n) if wlo[i] > np.float32(cut)]
    # medoid of a cluster: members' row maxima on the device, first minimum
    members = list(range(0, n, 3))
    got = C.members_row_max(m, members, ctx=gpu_ctx)
    D = np.zeros((n, n), np.float32)
    iu = np.triu_indices(n, 1)
    D[iu] = m.upper
    D = D + D.T
    want = np.array([max(D[a, b] for b in members if b != a) for a in members], np.float32)
    assert np.array_equal(got, want)
    lo1, hi1 = C.row_stats(B.RmsdMatrix(1, np.zeros(0, np.float32)), ctx=gpu_ctx)
    assert lo1[0] == np.float32(np.finfo(np.float32).max) and hi1[0] == -np.float32(np.finfo(np.float32).max)


@pytest.mark.parametrize("n", [2, 3, 63, 64, 65, 130, 257])
def test_row_stats_shapes_ties_and_inf(gpu_ctx, n):
    """Row minima / maxima over tile edges, with tied values, negative values,
    +-inf and NaN entries, against the oracle (the reference's `d < min`,
    `d > max` scans from f32::MAX and -f32::MAX: NaN never counts, an
    all-+inf row keeps f32::MAX as its minimum)."""
    rng = np.random.default_rng(n)
    up = rng.integers(-4, 6, n * (n - 1) // 2).astype(np.float32) * np.float32(0.25)
    up[up == 0] = np.float32(0.5)  # no -0.0 / +0.0 pairs: equal values, different bits
    r = rng.random(up.size)
    up[r < 0.1] = np.float32(np.inf)
    up[(r >= 0.1) & (r < 0.13)] = np.float32(np.nan)
    up[(r >= 0.13) & (r < 0.14)] = -np.float32(np.inf)
    m = B.RmsdMatrix(n, up)
    lo, hi = C.row_stats(m, ctx=gpu_ctx)
    wlo, whi = O.row_stats(up, n)
    assert np.array_equal(lo, wlo) and np.array_equal(hi, whi)


@pytest.mark.parametrize("chunk", ["1", "7", "40"])
def test_row_stats_in_chunks_of_tile_rows(gpu_ctx, monkeypatch, chunk):
    """Row statistics computed over chunks of tile rows (the bounded-scratch
    path large n takes) equal the oracle's."""
    monkeypatch.setenv("BIOSHELL_ROW_STATS_CHUNK_TILES", chunk)
    n = 700
    rng = np.random.default_rng(int(chunk))
    up = rng.integers(-3, 40, n * (n - 1) // 2).astype(np.float32) * np.float32(0.125)
    up[rng.random(up.size) < 0.05] = np.float32(np.inf)
    lo, hi = C.row_stats(B.RmsdMatrix(n, up), ctx=gpu_ctx)
    wlo, whi = O.row_stats(up, n)
    assert np.array_equal(lo, wlo) and np.array_equal(hi, whi)


def fused_matrix_and_stats(ctx, x, method):
    """bs_crmsd_all_vs_all_row_stats_device on torch device buffers."""
    import torch
    n, L, _ = x.shape
    dx = torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")
    out = torch.full((n * (n - 1) // 2,), -1.0, dtype=torch.float32, device="cuda:0")
    lo = torch.full((n,), -2.0, dtype=torch.float32, device="cuda:0")
    hi = torch.full((n,), -2.0, dtype=torch.float32, device="cuda:0")
    torch.cuda.synchronize()
    B.crmsd_all_vs_all_row_stats_device(dx.data_ptr(), n, L, out.data_ptr(), lo.data_ptr(), hi.data_ptr(),
                                        method, ctx=ctx)
    ctx.synchronize()
    return out.cpu().numpy(), lo.cpu().numpy(), hi.cpu().numpy()


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("n,L", [(2, 30), (3, 12), (63, 40), (64, 300), (65, 40), (130, 77), (777, 45)])
@pytest.mark.parametrize("method", [B.Method.BF16X3_GRAM, B.Method.FP32_GRAM, B.Method.FP32_PLAIN])
def test_row_stats_fused_into_all_vs_all(gpu_ctx, n, L, method):
    """The matrix of the fused call equals bs_crmsd_all_vs_all_device's, and
    its row statistics equal bs_rmsd_row_stats' (and the oracle's) on that
    matrix, bit for bit.  5 % near-identical conformations put flagged pairs,
    whose final values only the re-evaluation gives, among the row minima."""
    x = B.generate_ensemble(700 + n, n, L, sigma=(0.5, 3.0), near_fraction=0.05)
    up, lo, hi = fused_matrix_and_stats(gpu_ctx, x, method)
    flagged = gpu_ctx.last_stats()["flagged_pairs"]
    want_up = B.crmsd_all_vs_all(x, method, ctx=gpu_ctx).upper
    assert np.array_equal(bits(up), bits(want_up))
    slo, shi = C.row_stats(B.RmsdMatrix(n, up), ctx=gpu_ctx)
    wlo, whi = O.row_stats(up, n)
    assert np.array_equal(bits(lo), bits(slo)) and np.array_equal(bits(hi), bits(shi))
    assert np.array_equal(bits(lo), bits(wlo)) and np.array_equal(bits(hi), bits(whi))
    print(method.name, n, L, "flagged", flagged)


def test_row_stats_fused_with_flagged_minima(gpu_ctx):
    """Near-duplicate pairs are flagged and re-evaluated after the epilogue:
    here every element's row minimum comes from such a pair (each has a copy
    0.01 A away), so the fused minima are right only if the re-evaluated
    values reach them."""
    n2, L = 150, 60
    base = B.generate_ensemble(41, n2, L, sigma=(1.0, 3.0))
    rng = np.random.default_rng(41)
    x = np.concatenate([base, base + rng.normal(scale=0.01, size=base.shape)])
    up, lo, hi = fused_matrix_and_stats(gpu_ctx, x, B.Method.BF16X3_GRAM)
    assert gpu_ctx.last_stats()["flagged_pairs"] >= n2
    wlo, whi = O.row_stats(up, 2 * n2)
    assert np.array_equal(bits(lo), bits(wlo)) and np.array_equal(bits(hi), bits(whi))
    want = O.all_vs_all(x)
    assert np.abs(up.astype(np.float64) - want).max() <= 1e-4


def test_row_stats_fused_through_bitmap_scan(gpu_ctx):
    """More flagged pairs than the re-evaluation queue holds (rigid copies of
    two chains, 12 atoms): the bitmap scan's pairs reach the statistics too."""
    x = np.concatenate([B.generate_ensemble(1, 1200, 12, sigma=0.0),
                        B.generate_ensemble(2, 1200, 12, sigma=0.0)])
    up, lo, hi = fused_matrix_and_stats(gpu_ctx, x, B.Method.BF16X3_GRAM)
    assert gpu_ctx.last_stats()["flagged_pairs"] > (1 << 20)
    slo, shi = C.row_stats(B.RmsdMatrix(x.shape[0], up), ctx=gpu_ctx)
    assert np.array_equal(bits(lo), bits(slo)) and np.array_equal(bits(hi), bits(shi))


@pytest.mark.parametrize("limit", ["0", "10"])
def test_row_stats_fused_above_the_tile_limit(gpu_ctx, monkeypatch, limit):
    """Triangles beyond the epilogue's tile limit (lowered here) take the
    row-statistics kernels over the written matrix: the same results."""
    n, L = 777, 45
    x = B.generate_ensemble(777, n, L, sigma=(0.5, 3.0), near_fraction=0.05)
    up0, lo0, hi0 = fused_matrix_and_stats(gpu_ctx, x, B.Method.BF16X3_GRAM)
    monkeypatch.setenv("BIOSHELL_FUSED_STATS_MAX_TILES", limit)
    up, lo, hi = fused_matrix_and_stats(gpu_ctx, x, B.Method.BF16X3_GRAM)
    assert np.array_equal(bits(up), bits(up0))
    assert np.array_equal(bits(lo), bits(lo0)) and np.array_equal(bits(hi), bits(hi0))


def test_row_stats_fused_argument_errors(gpu_ctx):
    import torch
    d = torch.zeros(64, dtype=torch.float64, device="cuda:0")
    f = torch.zeros(64, dtype=torch.float32, device="cuda:0")
    with pytest.raises(B.BioshellGpuError):
        B.crmsd_all_vs_all_row_stats_device(d.data_ptr(), 1, 4, f.data_ptr(), f.data_ptr(), f.data_ptr(),
                                            ctx=gpu_ctx)
    with pytest.raises(B.BioshellGpuError):
        B.crmsd_all_vs_all_row_stats_device(d.data_ptr(), 2, 4, f.data_ptr(), 0, f.data_ptr(), ctx=gpu_ctx)


def test_clustering_2000_timing(gpu_ctx):
    n, L = 2000, 100
    x = B.generate_ensemble(9, n, L, sigma=(0.5, 3.0))
    m = B.crmsd_all_vs_all(x, B.Method.BF16X3_GRAM, ctx=gpu_ctx)
    t0 = time.perf_counter()
    got = device_merges(gpu_ctx, m.upper, n, "average")
    dt = time.perf_counter() - t0
    want = O.hclust(m.upper, n, "average")
    print(f"device clustering of {n}: {dt * 1e3:.1f} ms")
    assert np.array_equal(got[2].view(np.uint32), want[2].view(np.uint32))
    assert np.array_equal(got[0], want[0].astype(np.uint32))
