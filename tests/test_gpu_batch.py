"""Batches of independent ensembles (uneven n and L) through bs_crmsd_batch*,
against the FP64 oracle and against the single-ensemble path."""
import json

import numpy as np
import pytest

import bioshell_amd as B
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL_GRAM = 1e-4   # |dRMSD| A, BF16x3 / FP32 Gram vs FP64 oracle
TOL_FP64 = 1e-5   # FP64 plain


def make_batch(seed, shapes, sigma=(0.5, 3.0)):
    return [B.generate_ensemble(seed + k, n, L, sigma=sigma) for k, (n, L) in enumerate(shapes)]


# ragged on purpose: 1-conformation ensembles (no pairs), 1-3 atoms, sizes
# around the tile sides (31/32/33, 63/64/65) and the atom-block side (16)
SHAPES = [(2, 5), (1, 40), (3, 1), (5, 2), (7, 3), (31, 16), (32, 17), (33, 15), (63, 64),
          (64, 65), (65, 100), (100, 31), (150, 200), (20, 300), (97, 47), (130, 129)]


@pytest.fixture(scope="module")
def batch():
    xs = make_batch(500, SHAPES)
    want = [O.all_vs_all(x) for x in xs]
    return xs, want


def tile_table(tmp_path, tile):
    """Autotune cache that sends every ensemble to `tile` x `tile` Gram tiles."""
    p = tmp_path / f"tile{tile}.json"
    p.write_text(json.dumps({"arch": "gfx950", "method": 3, "cells": [
        {"l_max": 1 << 20, "n_max": 1 << 21, "tile": tile}]}))
    return str(p)


@pytest.mark.parametrize("method", [B.Method.BF16X3_GRAM, B.Method.FP32_GRAM])
def test_batch_matches_oracle(gpu_ctx, batch, method):
    xs, want = batch
    got = B.crmsd_batch(xs, method, ctx=gpu_ctx)
    worst = 0.0
    for x, g, w in zip(xs, got, want):
        assert g.shape == w.shape
        if g.size:
            worst = max(worst, float(np.abs(g.astype(np.float64) - w).max()))
    print(f"{method.name} batch: max |dRMSD| {worst:.3e}")
    assert worst <= TOL_GRAM


def test_batch_fp64_plain(gpu_ctx, batch):
    xs, want = batch
    got = B.crmsd_batch(xs, B.Method.FP64_PLAIN, ctx=gpu_ctx)
    for g, w in zip(got, want):
        if g.size:
            # f32 output: compare at f32 resolution of the value plus the FP64 gate
            assert np.abs(g.astype(np.float64) - w).max() <= TOL_FP64 + 1e-6 * max(1.0, w.max())


@pytest.mark.parametrize("tile", [32, 64])
def test_batch_bit_identical_to_single(tmp_path, batch, tile):
    """Each ensemble's triangle from the batch equals the single-ensemble call
    bit for bit, with either tile shape (same preparation, same accumulation
    order)."""
    xs, _ = batch
    ctx = B.Context(0)
    ctx.autotune_load(tile_table(tmp_path, tile))
    got = B.crmsd_batch(xs, B.Method.BF16X3_GRAM, ctx=ctx)
    for x, g in zip(xs, got):
        if x.shape[0] < 2:
            continue
        single = B.crmsd_all_vs_all(x, B.Method.BF16X3_GRAM, ctx=ctx).upper
        assert np.array_equal(g, single), f"shape {x.shape}"
    st = ctx.last_stats()
    assert st["pairs"] > 0
    ctx.close()


def test_batch_large_ensemble_sampled_reference(gpu_ctx):
    """An ensemble above 512 conformations takes its reference mean over n/8
    samples, in the batch as in the single call: both give the same triangle,
    within the gate of the oracle."""
    # sample counts around their switch points: S = n up to 64, then 64,
    # then n / 8 above 512 (65 at 520 and 521)
    xs = make_batch(700, [(1100, 24), (300, 40), (2, 9), (64, 20), (65, 20), (519, 10),
                          (520, 10), (521, 10)])
    got = B.crmsd_batch(xs, B.Method.BF16X3_GRAM, ctx=gpu_ctx)
    for x, g in zip(xs, got):
        single = B.crmsd_all_vs_all(x, B.Method.BF16X3_GRAM, ctx=gpu_ctx).upper
        assert np.array_equal(g, single), f"shape {x.shape}"
        assert np.abs(g.astype(np.float64) - O.all_vs_all(x, nthreads=8)).max() <= TOL_GRAM


def test_batch_mixed_fit_shapes(gpu_ctx):
    """Ensembles whose chains take every preparation shape (16 .. 256 threads
    per conformation, and the streamed form beyond 2048 atoms) in one batch:
    each triangle equals the single-ensemble call bit for bit and meets the
    oracle gate; the per-pair kernel's batched launch (tile sides by ensemble
    size) meets the FP64 gate."""
    xs = make_batch(800, [(70, 30), (40, 128), (33, 200), (50, 400), (20, 900), (12, 1500),
                          (9, 2100), (17, 129), (5, 2049)])
    got = B.crmsd_batch(xs, B.Method.BF16X3_GRAM, ctx=gpu_ctx)
    plain = B.crmsd_batch(xs, B.Method.FP64_PLAIN, ctx=gpu_ctx)
    for x, g, p in zip(xs, got, plain):
        single = B.crmsd_all_vs_all(x, B.Method.BF16X3_GRAM, ctx=gpu_ctx).upper
        assert np.array_equal(g, single), f"shape {x.shape}"
        want = O.all_vs_all(x, nthreads=8)
        assert np.abs(g.astype(np.float64) - want).max() <= TOL_GRAM, f"shape {x.shape}"
        assert np.abs(p.astype(np.float64) - want).max() <= TOL_FP64 + 1e-6 * max(1.0, want.max())


def test_batch_graph_replay_and_rebuild(gpu_ctx, batch):
    """The captured graph is replayed for a repeated shape and rebuilt for a new
    one; every replay gives the eager-equivalent (first) result."""
    xs, _ = batch
    first = B.crmsd_batch(xs, B.Method.BF16X3_GRAM, ctx=gpu_ctx)
    other = make_batch(900, [(40, 50), (70, 20)])
    other_first = B.crmsd_batch(other, B.Method.BF16X3_GRAM, ctx=gpu_ctx)
    for _ in range(3):
        again = B.crmsd_batch(xs, B.Method.BF16X3_GRAM, ctx=gpu_ctx)
        for a, b in zip(first, again):
            assert np.array_equal(a, b)
        o2 = B.crmsd_batch(other, B.Method.BF16X3_GRAM, ctx=gpu_ctx)
        for a, b in zip(other_first, o2):
            assert np.array_equal(a, b)


def test_batch_device_entry(gpu_ctx, batch):
    import torch

    xs, _ = batch
    dev = torch.device("cuda:0")
    keep = [(x, n) for x, n in zip(xs, [x.shape[0] for x in xs]) if n >= 2]
    d_in = [torch.from_numpy(x).to(dev) for x, _ in keep]
    d_out = [torch.full((n * (n - 1) // 2,), -1.0, dtype=torch.float32, device=dev) for _, n in keep]
    torch.cuda.synchronize()
    for rep in range(3):  # replays with eager work between them
        for t in d_out:
            t.fill_(-1.0)
        B.crmsd_batch_device([t.data_ptr() for t in d_in], [x.shape[0] for x, _ in keep],
                             [x.shape[1] for x, _ in keep], [t.data_ptr() for t in d_out],
                             B.Method.BF16X3_GRAM, ctx=gpu_ctx)
        gpu_ctx.synchronize()
        host = B.crmsd_batch([x for x, _ in keep], B.Method.BF16X3_GRAM, ctx=gpu_ctx)
        for t, h in zip(d_out, host):
            assert np.array_equal(t.cpu().numpy(), h), f"replay {rep}"


def test_batch_flagged_pairs_reevaluated(gpu_ctx):
    """Exact duplicates inside ensembles are flagged and re-evaluated in FP64."""
    xs = []
    for k, (n, L) in enumerate([(40, 60), (90, 33)]):
        x = B.generate_ensemble(70 + k, n, L, sigma=(1.0, 2.0))
        x[1] = x[0]
        x[n - 1] = x[3]
        xs.append(x)
    got = B.crmsd_batch(xs, B.Method.BF16X3_GRAM, ctx=gpu_ctx)
    assert gpu_ctx.last_stats()["flagged_pairs"] >= 2
    for x, g in zip(xs, got):
        w = O.all_vs_all(x)
        assert np.abs(g.astype(np.float64) - w).max() <= TOL_GRAM
        assert g[O.tri_index(x.shape[0], 0, 1)] < 1e-6


def test_batch_argument_errors(gpu_ctx):
    with pytest.raises(ValueError):
        B.crmsd_batch([np.zeros((4, 5))], ctx=gpu_ctx)  # not n x L x 3
    # zero atoms is rejected by the library with a status, not a crash
    lib = B._native.lib()
    import ctypes
    nc = (ctypes.c_size_t * 1)(3)
    na = (ctypes.c_size_t * 1)(0)
    px = (B._native.c_dp * 1)()
    po = (B._native.c_fp * 1)()
    assert lib.bs_crmsd_batch(gpu_ctx.handle, px, nc, na, 1, 3, po) == -1
    assert b"n_atoms" in lib.bs_last_error()
    assert lib.bs_crmsd_batch(gpu_ctx.handle, px, nc, na, 0, 3, po) == 0  # empty batch


def test_autotune_writes_and_loads(tmp_path):
    ctx = B.Context(0)
    path = str(tmp_path / "tune.json")
    ctx.autotune(B.Method.BF16X3_GRAM, path)
    cfg = json.loads(open(path).read())
    assert cfg["arch"] == "gfx950" and cfg["method"] == 3
    assert len(cfg["cells"]) == 25
    for c in cfg["cells"]:
        assert c["tile"] in (32, 64)
        assert c["ms_tile32"] > 0 and c["ms_tile64"] > 0
        assert c["tile"] == (32 if c["ms_tile32"] <= c["ms_tile64"] else 64)
    print("autotune:", [(c["l_max"], c["n_max"], c["tile"]) for c in cfg["cells"]])
    ctx2 = B.Context(0)
    ctx2.autotune_load(path)
    xs = make_batch(77, [(50, 80), (300, 120)])
    a = B.crmsd_batch(xs, ctx=ctx)
    b = B.crmsd_batch(xs, ctx=ctx2)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
    with pytest.raises(B.BioshellGpuError):
        ctx2.autotune_load(str(tmp_path / "missing.json"))
    bad = tmp_path / "bad.json"
    bad.write_text('{"arch": "gfx942", "method": 3, "cells": []}')
    with pytest.raises(B.BioshellGpuError):
        ctx2.autotune_load(str(bad))
    ctx.close()
    ctx2.close()
