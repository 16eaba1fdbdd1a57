"""Device-resident libraries, query-vs-library cRMSD, top-k and the serving loop."""
import threading

import numpy as np
import pytest

import bioshell_amd as B
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-4  # |dRMSD| A, BF16x3 Gram vs FP64 oracle


def make_bucket(seed, n_lib, n_q, L):
    """Library and queries drawn from one base chain (queries are extra conformations)."""
    x = B.generate_ensemble(seed, n_lib + n_q, L, sigma=(0.5, 3.0), near_fraction=0.05)
    return x[:n_lib], x[n_lib:]


@pytest.fixture(scope="module")
def bucket():
    lib_x, q = make_bucket(31, 1500, 24, 64)
    want = O.query_vs_library(q, lib_x)
    return lib_x, q, want


def test_query_all_matches_oracle(gpu_ctx, bucket):
    lib_x, q, want = bucket
    lib = B.Library(lib_x, ctx=gpu_ctx)
    got = lib.query_all(q).astype(np.float64)
    assert got.shape == want.shape
    err = np.abs(got - want)
    print(f"query_all: max |dRMSD| {err.max():.3e}")
    assert err.max() <= TOL
    lib.close()


def check_topk(idx, val, want, k):
    n_q = want.shape[0]
    for a in range(n_q):
        order = np.argsort(want[a], kind="stable")
        kth = want[a, order[k - 1]]
        assert np.all(np.diff(val[a]) >= 0), "top-k not ascending"
        assert len(set(idx[a].tolist())) == k, "duplicate indices"
        # every returned member's true RMSD matches its reported value
        assert np.abs(want[a, idx[a]] - val[a]).max() <= TOL
        # and no member outside the set is clearly better than the k-th
        assert val[a, -1] <= kth + 2 * TOL


@pytest.mark.parametrize("k", [1, 7, 100])
def test_query_topk(gpu_ctx, bucket, k):
    lib_x, q, want = bucket
    lib = B.Library(lib_x, ctx=gpu_ctx)
    idx, val = lib.query_topk(q, k)
    check_topk(idx, val.astype(np.float64), want, k)
    lib.close()


def test_topk_ties_resolve_to_lower_index(gpu_ctx):
    base = B.generate_ensemble(5, 40, 32, sigma=(1.0, 2.0))
    lib_x = np.concatenate([base, base, base])  # every member appears three times
    lib = B.Library(lib_x, ctx=gpu_ctx)
    idx, val = lib.query_topk(base[:3], 9)
    for a in range(3):
        # the query itself is at a, a+40, a+80, each at RMSD ~0, lowest index first
        assert list(idx[a, :3]) == [a, a + 40, a + 80]
        assert np.all(val[a, :3] < 1e-3)
    lib.close()


@pytest.mark.parametrize("n_base,copies,k", [(7000, 3, 1), (7000, 3, 100), (7000, 3, 1024), (8193, 1, 100),
                                             (2731, 3, 1000)])
def test_topk_selection_is_exact_over_chunks(gpu_ctx, n_base, copies, k):
    """The selection against the device's own score rows: the k smallest by
    (value, index), i.e. a stable argsort, across chunk boundaries and with
    every value repeated in several chunks."""
    base = B.generate_ensemble(17, n_base, 24, sigma=(0.5, 3.0))
    lib_x = np.concatenate([base] * copies)
    lib = B.Library(lib_x, ctx=gpu_ctx)
    q = np.concatenate([base[:3], B.generate_ensemble(18, 3, 24, sigma=(0.5, 3.0))])
    scores = lib.query_all(q)
    idx, val = lib.query_topk(q, k)
    for a in range(q.shape[0]):
        want = np.argsort(scores[a], kind="stable")[:k]
        assert np.array_equal(idx[a], want.astype(np.uint32)), a
        assert np.array_equal(val[a].view(np.uint32), scores[a, want].view(np.uint32))
    lib.close()


def test_topk_argument_errors(gpu_ctx, bucket):
    lib_x, q, _ = bucket
    lib = B.Library(lib_x[:50], ctx=gpu_ctx)
    with pytest.raises(B.BioshellGpuError):
        lib.query_topk(q, 51)  # k > library size
    with pytest.raises(ValueError):
        lib.query_topk(q[:, :10], 5)  # wrong atom count
    lib.close()


def test_server_results_equal_alone_and_in_mix(gpu_ctx):
    """Continuous batching: each request's result is the same whether it ran
    alone or merged with others (bit-identical)."""
    libs_x, queries = [], []
    for b, L in enumerate((32, 64)):
        lx, qq = make_bucket(100 + b, 800, 40, L)
        libs_x.append(lx)
        queries.append(qq)
    libs = [B.Library(lx, ctx=gpu_ctx) for lx in libs_x]
    rng = np.random.default_rng(3)
    reqs = []
    for _ in range(24):
        b = int(rng.integers(0, 2))
        n = int(rng.integers(1, 9))
        s = int(rng.integers(0, 40 - n))
        reqs.append((b, queries[b][s:s + n]))
    # alone: synchronous query calls
    alone = [libs[b].query_topk(q, 20) for b, q in reqs]
    # in a mix: all submitted at once from two threads
    srv = B.Server(libs, max_inflight=3)
    tickets = [None] * len(reqs)

    def submit(part):
        for r in part:
            b, q = reqs[r]
            tickets[r] = srv.submit(b, q, 20)

    th = [threading.Thread(target=submit, args=(list(range(p, len(reqs), 2)),)) for p in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for r, t in enumerate(tickets):
        idx, val = t.wait()
        assert np.array_equal(idx, alone[r][0])
        assert np.array_equal(val, alone[r][1])
        assert t.latency_ms() > 0.0
    # and against the oracle
    for r, (b, q) in enumerate(reqs[:4]):
        check_topk(tickets[r].idx, tickets[r].rmsd.astype(np.float64), O.query_vs_library(q, libs_x[b]), 20)
    srv.close()
    for l in libs:
        l.close()


def test_server_large_requests_beside_small_ones(gpu_ctx):
    """Requests larger than the batch size slots reserve (256 queries) run as
    batches of their own and grow the slot; merges of small requests stop at
    the reserved size.  Every result equals the synchronous call."""
    lx, qq = make_bucket(140, 300, 420, 32)
    lib = B.Library(lx, ctx=gpu_ctx)
    sizes = [3, 400, 7, 200, 200, 1, 60, 300, 5]
    reqs, s0 = [], 0
    for n in sizes:
        reqs.append(qq[s0:s0 + n] if s0 + n <= len(qq) else qq[:n])
        s0 = (s0 + 37) % 20
    alone = [lib.query_topk(q, 15) for q in reqs]
    srv = B.Server([lib], max_inflight=2)
    tickets = [srv.submit(0, q, 15) for q in reqs]
    for r, t in enumerate(tickets):
        idx, val = t.wait()
        assert np.array_equal(idx, alone[r][0]), f"request {r} ({len(reqs[r])} queries)"
        assert np.array_equal(val, alone[r][1])
    srv.close()
    lib.close()


def test_server_slots_start_clean_on_recycled_memory(gpu_ctx):
    """Slots size their workspaces up front, in memory that may hold anything
    (here: pages just freed after being filled with set bits).  What that
    memory held must not queue any pair for FP64 re-evaluation: the
    re-evaluated value rounds its square root correctly while the epilogue
    takes v_sqrt_f32, so a spurious re-evaluation shows as a 1-ulp change.
    Every result equals the synchronous call."""
    import torch

    lx, qq = make_bucket(141, 200, 64, 32)
    lib = B.Library(lx, ctx=gpu_ctx)
    alone = [lib.query_topk(qq[a:a + 16], 10) for a in range(0, 64, 16)]
    junk = torch.full((64 << 20,), -1, dtype=torch.int32, device="cuda")  # 256 MiB of ones
    torch.cuda.synchronize()
    del junk
    torch.cuda.empty_cache()
    srv = B.Server([lib], max_inflight=2)
    tickets = [srv.submit(0, qq[a:a + 16], 10) for a in range(0, 64, 16)]
    for r, t in enumerate(tickets):
        idx, val = t.wait()
        assert np.array_equal(idx, alone[r][0]), r
        assert np.array_equal(val.view(np.uint32), alone[r][1].view(np.uint32)), r
    srv.close()
    lib.close()


def test_handles_close_in_any_order():
    """Libraries and servers hold their context (and servers their libraries):
    closing the context's and the library's handles first leaves the server
    working, and the memory goes with the last handle."""
    ctx = B.Context()
    lx, qq = make_bucket(151, 100, 8, 24)
    lib = B.Library(lx, ctx=ctx)
    want = lib.query_topk(qq, 5)
    srv = B.Server([lib], max_inflight=1)
    ctx.close()
    lib.close()
    idx, val = srv.submit(0, qq, 5).wait()
    assert np.array_equal(idx, want[0]) and np.array_equal(val, want[1])
    srv.close()


def test_server_budget_and_cancel(gpu_ctx):
    lx, qq = make_bucket(7, 600, 16, 48)
    lib = B.Library(lx, ctx=gpu_ctx)
    tiny = B.Server([lib], hbm_budget_bytes=1 << 12, max_inflight=1)
    t = tiny.submit(0, qq[:4], 5)
    with pytest.raises(B.BioshellGpuError) as e:
        t.wait()
    assert e.value.status == -4
    tiny.close()
    srv = B.Server([lib], max_inflight=1)
    ts = [srv.submit(0, qq, 5) for _ in range(30)]
    cancelled = 0
    try:
        ts[-1].cancel()
        cancelled = 1
    except B.BioshellGpuError as err:
        assert err.status == -6  # already admitted
    for k, t in enumerate(ts):
        if k == len(ts) - 1 and cancelled:
            with pytest.raises(B.BioshellGpuError) as e2:
                t.wait()
            assert e2.value.status == -7
        else:
            t.wait()
    srv.close()
    lib.close()


def test_server_tickets_are_released_and_bounded(gpu_ctx):
    """Ticket records go when the caller releases them (dropping a Python
    Ticket does): many requests leave the server's table bounded.  A ticket
    dropped while queued or running never writes its (freed) buffers."""
    import gc

    lx, qq = make_bucket(55, 500, 16, 32)
    lib = B.Library(lx, ctx=gpu_ctx)
    srv = B.Server([lib], max_inflight=2)
    want = lib.query_topk(qq[:2], 5)
    for r in range(300):
        t = srv.submit(0, qq[:2], 5)
        if r % 3 == 0:
            idx, val = t.wait()
            assert np.array_equal(idx, want[0])
            assert t.latency_ms() > 0.0
        del t  # finished or not: released
        if r % 50 == 0:
            gc.collect()
    gc.collect()
    assert srv.ticket_count() <= 4, srv.ticket_count()
    # explicit release of a waited ticket, then its id is unknown
    t = srv.submit(0, qq, 5)
    t.wait()
    tid = t.id
    t.release()
    assert B._native.lib().bs_server_poll(srv.handle, tid) == -1
    # a queue of tickets dropped at once, then the server still answers
    burst = [srv.submit(0, qq, 5) for _ in range(40)]
    del burst
    gc.collect()
    idx, val = srv.submit(0, qq[:2], 5).wait()
    assert np.array_equal(idx, want[0]) and np.array_equal(val, want[1])
    srv.close()
    lib.close()


def test_server_rejects_k_beyond_selection_at_submit(gpu_ctx):
    """k the top-k selection cannot take for a library fails at submit, so it
    can never take requests merged with it down; small k beside it still run."""
    lx = B.generate_ensemble(66, 300_000, 8, sigma=(0.5, 3.0))
    lib = B.Library(lx, ctx=gpu_ctx)
    srv = B.Server([lib], max_inflight=1)
    q = lx[:2]
    small = [srv.submit(0, q, 10) for _ in range(3)]
    with pytest.raises(B.BioshellGpuError) as e:
        srv.submit(0, q, 1024)  # 300k members at k = 1024: beyond the merge's 8192 candidates
    assert e.value.status == -5
    alone = lib.query_topk(q, 10)
    for t in small:
        idx, val = t.wait()
        assert np.array_equal(idx, alone[0]) and np.array_equal(val, alone[1])
    srv.close()
    lib.close()
