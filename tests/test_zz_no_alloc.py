"""The no-allocation variant of the device all-vs-all entry point
(bs_crmsd_all_vs_all_device_noalloc), through the C ABI: its results equal
bs_crmsd_all_vs_all_device's bit for bit on two calls in a row, the second
with a smaller input (the shared scratch keeps the first call's contents);
an input beyond its workspace is rejected with a clear error."""
import ctypes

import numpy as np
import pytest

import bioshell_amd as B
from bioshell_amd import _native
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def device_call(fn, ctx, d_x, n, L, method, d_out, out_f64):
    return fn(ctx.handle, ctypes.c_void_p(d_x.data_ptr()), n, L, int(method), ctypes.c_void_p(d_out.data_ptr()),
              int(out_f64))


@pytest.mark.parametrize("method,out_f64", [(B.Method.BF16X3_GRAM, False), (B.Method.FP32_GRAM, True),
                                            (B.Method.FP64_PLAIN, True)])
def test_noalloc_equals_device_entry_twice(method, out_f64):
    import torch

    lib = _native.lib()
    ctx = B.Context(0)
    dt = torch.float64 if out_f64 else torch.float32
    for seed, n, L in ((31, 1500, 120), (32, 300, 77)):  # the second call is smaller
        print(f"noalloc {method.name} f64={out_f64}: n={n} L={L}", flush=True)
        x = B.generate_ensemble(seed, n, L, sigma=(0.5, 3.0), near_fraction=0.05)
        d_x = torch.from_numpy(x).to("cuda:0")
        m = n * (n - 1) // 2
        want = torch.full((m,), float("nan"), dtype=dt, device="cuda:0")
        got = torch.full((m,), float("nan"), dtype=dt, device="cuda:0")
        torch.cuda.synchronize()
        assert device_call(lib.bs_crmsd_all_vs_all_device, ctx, d_x, n, L, method, want, out_f64) == 0
        ctx.synchronize()
        used = ctx.bytes_in_use()
        assert device_call(lib.bs_crmsd_all_vs_all_device_noalloc, ctx, d_x, n, L, method, got, out_f64) == 0
        ctx.synchronize()
        assert ctx.bytes_in_use() == used  # nothing allocated through the context
        g, w = got.cpu().numpy(), want.cpu().numpy()
        assert np.array_equal(g.view(np.uint64 if out_f64 else np.uint32), w.view(np.uint64 if out_f64 else np.uint32))
        tol = 1e-5 if method == B.Method.FP64_PLAIN else 1e-4
        assert np.abs(g.astype(np.float64) - O.all_vs_all(x, nthreads=16)).max() <= tol
    ctx.close()


def test_noalloc_headline_size_and_limit():
    """10,000 x 300 fits the workspace and matches the allocating entry point;
    an input that needs more than the workspace fails with a clear error."""
    import torch

    lib = _native.lib()
    ctx = B.Context(0)
    n, L = 10_000, 300
    x = B.generate_ensemble(20260, n, L, sigma=(0.5, 4.0), near_fraction=0.05, near_sigma=0.01)
    d_x = torch.from_numpy(x).to("cuda:0")
    m = n * (n - 1) // 2
    want = torch.empty(m, dtype=torch.float32, device="cuda:0")
    got = torch.empty(m, dtype=torch.float32, device="cuda:0")
    torch.cuda.synchronize()
    assert device_call(lib.bs_crmsd_all_vs_all_device, ctx, d_x, n, L, B.Method.BF16X3_GRAM, want, False) == 0
    assert device_call(lib.bs_crmsd_all_vs_all_device_noalloc, ctx, d_x, n, L, B.Method.BF16X3_GRAM, got,
                       False) == 0
    ctx.synchronize()
    assert torch.equal(got.view(torch.int32), want.view(torch.int32))
    del got, want
    big = torch.empty((12_000, 400, 3), dtype=torch.float64, device="cuda:0")
    out = torch.empty(1, dtype=torch.float32, device="cuda:0")
    st = device_call(lib.bs_crmsd_all_vs_all_device_noalloc, ctx, big, 12_000, 400, B.Method.BF16X3_GRAM, out,
                     False)
    assert st == -4
    assert b"no-allocation" in lib.bs_last_error()
    ctx.close()
