"""GPU parity of the all-vs-all cRMSD path against the FP64 oracle, through the C ABI.

Tolerances (|dRMSD| in A):
  FP64 plain      1e-5   (the FP64 gate)
  FP32 plain      5e-2   (reported, not a gate: textbook QCP in FP32 cancels
                          E0 - lambda for close pairs)
  FP32 Gram       1e-4
  BF16x3 Gram     1e-4   (the split-BF16 target)
"""
import numpy as np
import pytest

import bioshell_amd as B
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = {
    B.Method.FP64_PLAIN: 1e-5,
    B.Method.FP32_PLAIN: 5e-2,
    B.Method.FP32_GRAM: 1e-4,
    B.Method.BF16X3_GRAM: 1e-4,
}
ALL = list(TOL)


def run(ctx, x, method, out_f64=None):
    if out_f64 is None:
        out_f64 = method == B.Method.FP64_PLAIN
    return B.crmsd_all_vs_all(x, method, out_f64=out_f64, ctx=ctx).upper.astype(np.float64)


def report(name, err):
    print(f"{name}: max |dRMSD| = {err.max():.3e}  mean = {err.mean():.3e}", flush=True)


# ---------------------------------------------------------------- single pair
def test_pair_fp64(gpu_ctx):
    rng = np.random.default_rng(0)
    for L in (1, 2, 3, 17, 76, 300):
        a = rng.normal(size=(L, 3)) * 6
        b = a + rng.normal(size=(L, 3))
        assert B.crmsd(a, b, ctx=gpu_ctx) == pytest.approx(O.crmsd(a, b), abs=1e-9)


def test_superpose_fp64(gpu_ctx):
    rng = np.random.default_rng(1)
    a = rng.normal(size=(40, 3)) * 6
    b = (a + rng.normal(size=a.shape) * 0.5) @ np.linalg.qr(rng.normal(size=(3, 3)))[0].T + 20.0
    if np.linalg.det(np.linalg.qr(rng.normal(size=(3, 3)))[0]) < 0:
        pass
    T, r = B.superpose(a, b, ctx=gpu_ctx)
    R, t, r0 = O.superpose(a, b)
    assert r == pytest.approx(r0, abs=1e-9)
    assert np.allclose(T.rotation, R, atol=1e-9)
    assert np.allclose(T.translation, t, atol=1e-8)
    assert np.sqrt(((a - T.apply(b)) ** 2).sum(1).mean()) == pytest.approx(r0, abs=1e-9)


# ---------------------------------------------------------------- golden
@pytest.mark.parametrize("method", ALL)
def test_golden_2jqb(gpu_ctx, golden_2jqb, method):
    for key in ("ca", "all_atom"):
        got = run(gpu_ctx, golden_2jqb[key], method)
        err = np.abs(got - golden_2jqb[key + "_rmsd"])
        report(f"2jqb {key} {method.name}", err)
        assert err.max() <= TOL[method]


@pytest.mark.parametrize("method", ALL)
def test_config0_all_vs_all(gpu_ctx, golden_config0, method):
    """200 conformations x 76 CA, sigma 1.5 A: 19,900 pairs vs the committed golden values."""
    x = B.generate_ensemble(0, 200, 76, sigma=1.5)
    got = run(gpu_ctx, x, method)
    err = np.abs(got - golden_config0["rmsd"])
    report(f"config0 {method.name}", err)
    assert err.max() <= TOL[method]
    st = gpu_ctx.last_stats()
    assert st["pairs"] == 19900


# ---------------------------------------------------------------- ragged shapes
@pytest.mark.parametrize("n", [2, 3, 31, 32, 33, 63, 64, 65, 130])
@pytest.mark.parametrize("L", [1, 2, 3, 15, 16, 17, 33])
def test_ragged_shapes(gpu_ctx, n, L):
    x = B.generate_ensemble(100 + n * 7 + L, n, L, sigma=(0.5, 3.0), near_fraction=0.1)
    want = O.all_vs_all(x, nthreads=4)
    for method in ALL:
        got = run(gpu_ctx, x, method)
        assert got.shape == want.shape
        err = np.abs(got - want)
        tol = TOL[method]
        if L == 2 and method == B.Method.FP32_PLAIN:
            # two atoms are collinear: Horn's matrix has an exact double top
            # eigenvalue and FP32 resolves E0 - lambda only to ~1e-5 E0, which
            # the square root turns into a few 1e-2 A (reported, not a gate)
            tol = 0.1
        assert err.max() <= tol, (method.name, err.max())


@pytest.mark.parametrize("seed", range(8))
def test_random_shapes_and_scales(gpu_ctx, seed):
    """Seeded random ensembles: n up to 400, L up to 700, residual noise from
    0.01 to 8 A, coordinates scaled by up to 20 (large residuals go to FP64
    re-evaluation); every method against the oracle."""
    rng = np.random.default_rng(1000 + seed)
    n, L = int(rng.integers(2, 401)), int(rng.integers(1, 701))
    lo = float(rng.uniform(0.01, 2.0))
    x = B.generate_ensemble(2000 + seed, n, L, sigma=(lo, lo + float(rng.uniform(0.0, 6.0))),
                            near_fraction=float(rng.uniform(0.0, 0.2)))
    x = x * float(rng.choice([1.0, 1.0, 5.0, 20.0]))
    want = O.all_vs_all(x, nthreads=8)
    print(f"seed {seed}: n={n} L={L}", flush=True)
    for method in ALL:
        err = np.abs(run(gpu_ctx, x, method) - want)
        report(f"random {method.name}", err)
        # FP32 plain: its error grows with the coordinate magnitude
        tol = TOL[method] * (float(np.abs(x).max()) / 50.0 if method == B.Method.FP32_PLAIN else 1.0)
        assert err.max() <= max(tol, TOL[method]), (method, n, L)


@pytest.mark.parametrize("L", [128, 129, 256, 257, 512, 513, 1024, 1025, 2048, 2049, 2600])
def test_chain_lengths_across_fit_shapes(gpu_ctx, L):
    """The preparation fits each conformation with 16 .. 256 threads holding up
    to 8 atoms each (lane count doubling at 128, 256, 512, 1024 atoms) and
    streams chains beyond 2048 atoms; every shape and both Gram modes against
    the oracle, n = 45 (padding rows in the last work-group)."""
    x = B.generate_ensemble(4000 + L, 45, L, sigma=(0.5, 3.0), near_fraction=0.1)
    want = O.all_vs_all(x, nthreads=8)
    for method in (B.Method.BF16X3_GRAM, B.Method.FP32_GRAM, B.Method.FP64_PLAIN):
        err = np.abs(run(gpu_ctx, x, method) - want)
        assert err.max() <= TOL[method], (method.name, L, err.max())


def test_results_independent_of_earlier_calls(gpu_ctx):
    """The same ensemble gives the same triangle bit for bit whatever ran on the
    device before (other shapes, streamed blocks, batches): no kernel reads
    memory or LDS it has not written."""
    x = B.generate_ensemble(4200, 257, 300, sigma=(0.5, 4.0), near_fraction=0.05)
    for m in (B.Method.BF16X3_GRAM, B.Method.FP32_GRAM):
        first = B.crmsd_all_vs_all(x, m, ctx=gpu_ctx).upper
        B.crmsd_all_vs_all_streamed(x, m, tile=64, ctx=gpu_ctx)
        B.crmsd_batch([B.generate_ensemble(4201, 90, 700), B.generate_ensemble(4202, 33, 2100)], m,
                      ctx=gpu_ctx)
        B.crmsd_all_vs_all(B.generate_ensemble(4203, 500, 40), m, ctx=gpu_ctx)
        again = B.crmsd_all_vs_all(x, m, ctx=gpu_ctx).upper
        assert np.array_equal(first, again), m.name


@pytest.mark.parametrize("n", [17, 40, 100, 250, 700])
def test_plain_tile_blocks(gpu_ctx, n):
    """The per-pair kernel picks a tile side of 16, 32 or 64 conformations by
    ensemble size; every choice in both precisions against the oracle."""
    x = B.generate_ensemble(4100 + n, n, 57, sigma=(0.5, 3.0))
    want = O.all_vs_all(x, nthreads=8)
    for method in (B.Method.FP64_PLAIN, B.Method.FP32_PLAIN):
        err = np.abs(run(gpu_ctx, x, method) - want)
        assert err.max() <= TOL[method], (method.name, n, err.max())


def test_single_conformation_and_errors(gpu_ctx):
    x = B.generate_ensemble(1, 1, 10)
    assert B.crmsd_all_vs_all(x, ctx=gpu_ctx).upper.size == 0
    with pytest.raises(B.BioshellGpuError) as e:
        B.crmsd_all_vs_all(np.zeros((3, 0, 3)), ctx=gpu_ctx)
    assert e.value.status == -1


# ---------------------------------------------------------------- precision
def test_config1_shape_gram_vs_plain(gpu_ctx):
    """2,000 x 150 CA, sigma U(0.5, 4): FP32 plain vs FP32-MFMA Gram, both vs FP64."""
    x = B.generate_ensemble(1, 2000, 150, sigma=(0.5, 4.0))
    want = O.all_vs_all(x, nthreads=8)
    for method in (B.Method.FP32_PLAIN, B.Method.FP32_GRAM, B.Method.BF16X3_GRAM):
        err = np.abs(run(gpu_ctx, x, method) - want)
        report(f"config1 {method.name}", err)
        assert err.max() <= TOL[method]


def test_near_identical_pairs(gpu_ctx):
    """5 % near-identical conformations (sigma 0.01 A) stress degenerate QCP roots."""
    x = B.generate_ensemble(2, 600, 300, sigma=(0.5, 4.0), near_fraction=0.05, near_sigma=0.01)
    want = O.all_vs_all(x, nthreads=8)
    for method in (B.Method.FP64_PLAIN, B.Method.FP32_GRAM, B.Method.BF16X3_GRAM):
        err = np.abs(run(gpu_ctx, x, method) - want)
        report(f"near-identical {method.name}", err)
        assert err.max() <= TOL[method]
        small = want < 0.1
        assert small.sum() > 100
        assert err[small].max() <= TOL[method]


def test_near_identical_large_ensemble_not_flagged(gpu_ctx):
    """4,000 conformations: the shifted frame's reference is the mean of 500
    evenly spaced samples.  The near-identical conformations' residuals are
    that mean's noise (about 0.1 A); the error model keeps their pairs out of
    FP64 re-evaluation, and every value still meets the gate."""
    x = B.generate_ensemble(12, 4000, 100, sigma=(0.5, 4.0), near_fraction=0.05, near_sigma=0.01)
    got = run(gpu_ctx, x, B.Method.BF16X3_GRAM)
    flagged = gpu_ctx.last_stats()["flagged_pairs"]
    want = O.all_vs_all(x, nthreads=8)
    err = np.abs(got - want)
    report("near-identical n=4000 BF16X3_GRAM", err)
    print("flagged", flagged, flush=True)
    assert err.max() <= TOL[B.Method.BF16X3_GRAM]
    small = want < 0.1
    assert small.sum() > 10000
    assert err[small].max() <= TOL[B.Method.BF16X3_GRAM]
    assert flagged == 0


def test_duplicates_far_from_reference(gpu_ctx):
    """Exact and near duplicates of strongly perturbed conformations: the split
    products' error is largest here, so these pairs must be re-evaluated."""
    base = B.generate_ensemble(4, 40, 120, sigma=(3.0, 4.0))
    rng = np.random.default_rng(4)
    dup = base + rng.normal(size=base.shape) * 1e-3
    x = np.concatenate([base, base, dup])
    want = O.all_vs_all(x, nthreads=8)
    for method in (B.Method.FP32_GRAM, B.Method.BF16X3_GRAM):
        err = np.abs(run(gpu_ctx, x, method) - want)
        report(f"duplicates {method.name}", err)
        assert err.max() <= TOL[method]
        assert gpu_ctx.last_stats()["flagged_pairs"] >= 40


def test_order_and_rigid_invariance(gpu_ctx):
    x = B.generate_ensemble(5, 90, 64, sigma=(0.5, 3.0))
    n = len(x)
    d0 = B.crmsd_all_vs_all(x, B.Method.BF16X3_GRAM, ctx=gpu_ctx).to_dense()
    perm = np.random.default_rng(5).permutation(n)
    d1 = B.crmsd_all_vs_all(x[perm], B.Method.BF16X3_GRAM, ctx=gpu_ctx).to_dense()
    assert np.abs(d1 - d0[np.ix_(perm, perm)]).max() <= 2e-4
    y = x.copy()
    Q = np.linalg.qr(np.random.default_rng(6).normal(size=(3, 3)))[0]
    if np.linalg.det(Q) < 0:
        Q[:, 0] *= -1
    y[7] = y[7] @ Q.T + 30.0
    d2 = B.crmsd_all_vs_all(y, B.Method.BF16X3_GRAM, ctx=gpu_ctx).to_dense()
    assert np.abs(d2 - d0).max() <= 2e-4


def test_budget_is_enforced():
    with B.Context(0, hbm_budget_bytes=1 << 16) as ctx:
        x = B.generate_ensemble(6, 300, 100)
        with pytest.raises(B.BioshellGpuError) as e:
            B.crmsd_all_vs_all(x, B.Method.BF16X3_GRAM, ctx=ctx)
        assert e.value.status == -4


def test_device_pointer_entry(gpu_ctx):
    import torch

    x = B.generate_ensemble(7, 257, 90, sigma=(0.5, 3.0), near_fraction=0.05)
    want = O.all_vs_all(x, nthreads=8)
    dx = torch.from_numpy(x).to("cuda:0")
    dout = torch.empty(len(want), dtype=torch.float32, device="cuda:0")
    torch.cuda.synchronize()
    B.crmsd_all_vs_all_device(dx.data_ptr(), 257, 90, dout.data_ptr(), B.Method.BF16X3_GRAM, ctx=gpu_ctx)
    gpu_ctx.synchronize()  # the device entry point is asynchronous on the context's stream
    got = dout.cpu().numpy().astype(np.float64)
    st = gpu_ctx.last_stats()
    assert st["pairs"] == 257 * 256 // 2 and st["main_ms"] > 0.0
    assert len(gpu_ctx.history(3)) == 3
    assert np.abs(got - want).max() <= 1e-4


def test_default_context_through_the_c_abi():
    """A null context runs on the process-wide default context -- the binding
    of the context-free Rust crmsd(a, b) and crmsd_all_vs_all(models, method);
    the Python default_context() is that same context."""
    import ctypes

    L = B._native.lib()
    x = B.generate_ensemble(21, 30, 40, sigma=(0.5, 3.0))
    out = ctypes.c_double()
    assert L.bs_crmsd_pair_f64(None, x[0].ctypes.data_as(B._native.c_dp), x[1].ctypes.data_as(B._native.c_dp),
                               40, ctypes.byref(out)) == 0
    assert abs(out.value - O.crmsd(x[0], x[1])) <= 1e-9
    up = np.empty(30 * 29 // 2, np.float32)
    assert L.bs_crmsd_all_vs_all(None, x.ctypes.data_as(B._native.c_dp), 30, 40, int(B.Method.BF16X3_GRAM),
                                 up.ctypes.data_as(ctypes.c_void_p), 0) == 0
    assert np.abs(up.astype(np.float64) - O.all_vs_all(x)).max() <= TOL[B.Method.BF16X3_GRAM]
    h1, h2 = ctypes.c_void_p(), ctypes.c_void_p()
    assert L.bs_default_context(ctypes.byref(h1)) == 0 and L.bs_default_context(ctypes.byref(h2)) == 0
    assert h1.value and h1.value == h2.value == B.default_context().handle.value
    # the stats of the null-context call are on the default context
    assert B.default_context().last_stats()["pairs"] == 30 * 29 // 2
    # closing the Python wrapper does not destroy the library's context
    B.default_context().close()
    assert B.crmsd(x[0], x[1]) == out.value
    # threads share it safely (each call takes the context's lock)
    import threading
    res = [None] * 4

    def work(t):
        res[t] = B.crmsd_all_vs_all(x, B.Method.BF16X3_GRAM).upper

    th = [threading.Thread(target=work, args=(t,)) for t in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for r in res:
        assert np.array_equal(r, up)


def test_gram_kernel_outputs_f32_and_f64():
    """The Gram kernel with f32 and f64 output, both split modes."""
    x = B.generate_ensemble(8, 700, 120, sigma=(0.5, 4.0), near_fraction=0.05)
    want = O.all_vs_all(x, nthreads=8)
    with B.Context(0) as ctx:
        for method in (B.Method.BF16X3_GRAM, B.Method.FP32_GRAM):
            for f64 in (False, True):
                got = B.crmsd_all_vs_all(x, method, out_f64=f64, ctx=ctx).upper.astype(np.float64)
                err = np.abs(got - want)
                report(f"{method.name} f64={f64}", err)
                assert err.max() <= TOL[method]


def test_flag_queue_overflow_goes_through_bitmap_scan(gpu_ctx):
    """Rigid copies of two different chains (1,200 each, 12 atoms): residuals
    from the common reference are large, so duplicate pairs and, at this small
    atom count, all others are flagged.  More pairs are flagged than the
    2^20-entry re-evaluation queue holds, so the rest go through the bitmap
    scan; every value must still match the oracle."""
    x = np.concatenate([B.generate_ensemble(1, 1200, 12, sigma=0.0),
                        B.generate_ensemble(2, 1200, 12, sigma=0.0)])
    got = B.crmsd_all_vs_all(x, B.Method.BF16X3_GRAM, ctx=gpu_ctx).upper
    st = gpu_ctx.last_stats()
    print("flagged", st["flagged_pairs"], "of", got.size)
    assert st["flagged_pairs"] > (1 << 20)
    want = O.all_vs_all(x)
    assert np.abs(got.astype(np.float64) - want).max() <= TOL[B.Method.BF16X3_GRAM]


@pytest.mark.parametrize("n,L,tile", [(300, 40, 64), (1000, 77, 256), (130, 20, 4096), (257, 300, 128)])
def test_streamed_equals_one_shot(gpu_ctx, n, L, tile):
    """Blocks of tile x tile pairs (diagonal blocks as triangles, edge blocks
    partial) reassemble into exactly the one-shot triangle."""
    x = B.generate_ensemble(31 + n, n, L, sigma=(0.5, 4.0), near_fraction=0.05)
    for m in (B.Method.BF16X3_GRAM, B.Method.FP32_GRAM):
        one = B.crmsd_all_vs_all(x, m, ctx=gpu_ctx).upper
        st = B.crmsd_all_vs_all_streamed(x, m, tile=tile, ctx=gpu_ctx).upper
        assert np.array_equal(one, st), (m, n, tile)
    nb = (n + min(tile, (n + 63) // 64 * 64) - 1) // min(tile, (n + 63) // 64 * 64)
    assert gpu_ctx.last_stats()["main_launches"] == nb * (nb + 1) // 2
    want = O.all_vs_all(x)
    assert np.abs(st.astype(np.float64) - want).max() <= TOL[B.Method.FP32_GRAM]


def test_streamed_plain_and_errors(gpu_ctx):
    x = B.generate_ensemble(8, 90, 30, sigma=1.0)
    st = B.crmsd_all_vs_all_streamed(x, B.Method.FP64_PLAIN, tile=64, ctx=gpu_ctx).upper
    assert np.abs(st.astype(np.float64) - O.all_vs_all(x)).max() <= 1e-5 + 1e-6 * 20
    with pytest.raises(B.BioshellGpuError):
        B.crmsd_all_vs_all_streamed(x, tile=100, ctx=gpu_ctx)  # not a multiple of 64
    with pytest.raises(B.BioshellGpuError):
        B.crmsd_all_vs_all_streamed(x, tile=0, ctx=gpu_ctx)
