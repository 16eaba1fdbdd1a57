"""Host-side mirror of the intended ``bioshell_pdb::calc`` crmsd API.

The reference announces the function (bioshell-pdb/Cargo.toml:27, "nalgebra
... for helical axis and crmsd") but does not implement it; its shape here
follows the crate's idioms:

* ``crmsd(a, b)`` -- one pair, generic over "anything with a position" like
  ``SubstructureAxis::from_3d_points<T: HasCartesians>(coords: &[T])``
  (bioshell-pdb/src/calc/substructures.rs:17); inputs are (L, 3) f64 arrays,
  the Vec3 layout of bioshell-core/src/vec3.rs:43-51.
* ``crmsd_all_vs_all(models, method)`` -- an ensemble in the layout of
  ``Structure::model_coordinates`` (bioshell-pdb/src/structure.rs:101); the
  result is an :class:`RmsdMatrix` whose ``distance(i, j)`` plugs into
  ``hierarchical_clustering`` exactly like ``DataMatrixDistance::distance``
  (bioshell-clustering/src/hierarchical/distance_matrix.rs:71-77).
* ``superpose(a, b)`` -- a :class:`Rototranslation` with ``x' = R x + t``
  (bioshell-pdb/src/calc/transformation.rs:22-37, 171-174).

Every call goes through the C ABI of ``libbioshell_gpu.so``; there is no CPU
implementation behind it.
"""
from __future__ import annotations

import ctypes
import enum
from dataclasses import dataclass

import numpy as np

from . import _native


class Method(enum.IntEnum):
    """How pairs are evaluated (bs_method)."""

    FP64_PLAIN = 0
    FP32_PLAIN = 1
    FP32_GRAM = 2
    BF16X3_GRAM = 3


class BioshellGpuError(RuntimeError):
    """A non-zero bs_status, with the library's message."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[{status}] {message}")
        self.status = status


STATUS = {
    0: "OK",
    -1: "INVALID_ARGUMENT",
    -2: "NO_DEVICE",
    -3: "HIP",
    -4: "OUT_OF_BUDGET",
    -5: "UNSUPPORTED",
    -6: "NOT_READY",
    -7: "CANCELLED",
    -8: "INTERNAL",
    -9: "PARSE",
    -10: "IO",
}


def check(status: int) -> None:
    if status != 0:
        msg = _native.lib().bs_last_error()
        raise BioshellGpuError(status, (msg or b"").decode(errors="replace"))


def _f64(x) -> np.ndarray:
    return np.ascontiguousarray(x, dtype=np.float64)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_native.c_dp)


class Context:
    """One device, one stream, workspaces (bs_context)."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0, _handle=None):
        self._owned = _handle is None
        if _handle is not None:
            self._h = _handle
            return
        h = ctypes.c_void_p()
        check(_native.lib().bs_context_create(device, hbm_budget_bytes, ctypes.byref(h)))
        self._h = h

    @property
    def handle(self):
        if self._h is None:
            raise ValueError("context is closed")
        return self._h

    def close(self) -> None:
        """Destroys the context (the library's default context is never destroyed)."""
        if getattr(self, "_h", None) is not None:
            if self._owned:
                _native.lib().bs_context_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def synchronize(self) -> None:
        check(_native.lib().bs_context_synchronize(self.handle))

    @property
    def stream(self) -> int:
        return _native.lib().bs_context_stream(self.handle) or 0

    def last_stats(self) -> dict:
        st = _native.RunStats()
        check(_native.lib().bs_context_last_stats(self.handle, ctypes.byref(st)))
        return st.as_dict()

    def history(self, max_calls: int = 64) -> list[dict]:
        """Stats of the most recent calls, oldest first (synchronises on them)."""
        buf = (_native.RunStats * max_calls)()
        n = ctypes.c_size_t()
        check(_native.lib().bs_context_history(self.handle, buf, max_calls, ctypes.byref(n)))
        return [buf[k].as_dict() for k in range(n.value)]

    def bytes_in_use(self) -> int:
        return _native.lib().bs_context_bytes_in_use(self.handle)

    def autotune(self, method: "Method" = None, cache_path: str | None = None) -> None:
        """Times 32x32 and 64x64 Gram tiles per (atom count, ensemble size) cell,
        keeps the faster per cell for batches on this context, and writes the
        table to ``cache_path`` (JSON) when given (bs_autotune)."""
        m = Method.BF16X3_GRAM if method is None else method
        path = cache_path.encode() if cache_path else None
        check(_native.lib().bs_autotune(self.handle, int(m), path))

    def autotune_load(self, cache_path: str) -> None:
        """Loads a table written by :meth:`autotune` (bs_autotune_load)."""
        check(_native.lib().bs_autotune_load(self.handle, cache_path.encode()))


_default: Context | None = None


def default_context() -> Context:
    """The library's process-wide default context (bs_default_context: device 0,
    no budget, kept until the process ends) -- the one the C entry points use
    when given a null context, as the context-free Rust functions do."""
    global _default
    if _default is None or _default._h is None:
        h = ctypes.c_void_p()
        check(_native.lib().bs_default_context(ctypes.byref(h)))
        _default = Context(_handle=h)
    return _default


def device_count() -> int:
    return _native.lib().bs_device_count()


def version() -> str:
    return _native.lib().bs_version().decode()


# --------------------------------------------------------------- single pair
def crmsd(a, b, ctx: Context | None = None) -> float:
    """cRMSD of two conformations after optimal superposition, FP64 on the GPU."""
    a, b = _f64(a), _f64(b)
    if a.shape != b.shape or a.ndim != 2 or a.shape[1] != 3:
        raise ValueError(f"expected two (L, 3) arrays of equal shape, got {a.shape} and {b.shape}")
    ctx = ctx or default_context()
    out = ctypes.c_double()
    check(_native.lib().bs_crmsd_pair_f64(ctx.handle, _dp(a), _dp(b), a.shape[0], ctypes.byref(out)))
    return out.value


@dataclass
class Rototranslation:
    """x' = R x + t (bioshell-pdb/src/calc/transformation.rs:22-37)."""

    rotation: np.ndarray  # (3, 3), row-major
    translation: np.ndarray  # (3,)

    def apply(self, xyz) -> np.ndarray:
        return _f64(xyz) @ self.rotation.T + self.translation

    def apply_inverse(self, xyz) -> np.ndarray:
        return (_f64(xyz) - self.translation) @ self.rotation


def superpose(a, b, ctx: Context | None = None) -> tuple[Rototranslation, float]:
    """Optimal superposition of b onto a; returns (transform, rmsd) with a ~= transform.apply(b)."""
    a, b = _f64(a), _f64(b)
    if a.shape != b.shape or a.ndim != 2 or a.shape[1] != 3:
        raise ValueError(f"expected two (L, 3) arrays of equal shape, got {a.shape} and {b.shape}")
    ctx = ctx or default_context()
    R = np.empty(9)
    t = np.empty(3)
    r = ctypes.c_double()
    check(_native.lib().bs_superpose_f64(ctx.handle, _dp(a), _dp(b), a.shape[0], _dp(R), _dp(t), ctypes.byref(r)))
    return Rototranslation(R.reshape(3, 3), t), r.value


# --------------------------------------------------------------- all-vs-all
def tri_index(n: int, i, j):
    """Packed upper-triangle index of (i, j), i < j."""
    return i * n - i * (i + 1) // 2 + (j - i - 1)


class RmsdMatrix:
    """Packed upper triangle of an all-vs-all cRMSD matrix.

    ``distance(i, j)`` is 0 on the diagonal and symmetric, the contract of the
    ``Fn(usize, usize) -> f32`` closure that ``hierarchical_clustering``
    densifies (bioshell-clustering/src/hierarchical/clustering_matrix.rs:11-21).
    """

    def __init__(self, n: int, upper: np.ndarray):
        if upper.shape != (n * (n - 1) // 2,):
            raise ValueError("packed triangle has the wrong length")
        self._n = n
        self.upper = upper

    def n_elements(self) -> int:
        return self._n

    def distance(self, i: int, j: int) -> float:
        if i == j:
            return 0.0
        if i > j:
            i, j = j, i
        return float(self.upper[tri_index(self._n, i, j)])

    def to_dense(self) -> np.ndarray:
        n = self._n
        d = np.zeros((n, n), dtype=self.upper.dtype)
        iu = np.triu_indices(n, 1)
        d[iu] = self.upper
        d[(iu[1], iu[0])] = self.upper
        return d


def _ensemble(models) -> np.ndarray:
    x = _f64(models if isinstance(models, np.ndarray) else np.stack([_f64(m) for m in models]))
    if x.ndim != 3 or x.shape[2] != 3:
        raise ValueError(f"expected an (n_conf, n_atoms, 3) ensemble, got shape {x.shape}")
    return x


def crmsd_all_vs_all(models, method: Method = Method.BF16X3_GRAM, out_f64: bool = False,
                     ctx: Context | None = None) -> RmsdMatrix:
    """All-vs-all cRMSD of an ensemble (host in, host out)."""
    x = _ensemble(models)
    n, L, _ = x.shape
    ctx = ctx or default_context()
    out = np.empty(max(n * (n - 1) // 2, 0), dtype=np.float64 if out_f64 else np.float32)
    check(_native.lib().bs_crmsd_all_vs_all(ctx.handle, _dp(x), n, L, int(method),
                                            out.ctypes.data_as(ctypes.c_void_p), int(out_f64)))
    return RmsdMatrix(n, out)


def crmsd_all_vs_all_device(d_xyz: int, n_conf: int, n_atoms: int, d_out: int,
                            method: Method = Method.BF16X3_GRAM, out_f64: bool = False,
                            ctx: Context | None = None) -> None:
    """All-vs-all on device pointers (e.g. ``tensor.data_ptr()``).

    Asynchronous on the context's stream: call ``ctx.synchronize()`` (or a
    device-wide synchronisation) before reading the output.
    """
    ctx = ctx or default_context()
    check(_native.lib().bs_crmsd_all_vs_all_device(ctx.handle, ctypes.c_void_p(d_xyz), n_conf, n_atoms,
                                                   int(method), ctypes.c_void_p(d_out), int(out_f64)))


def crmsd_all_vs_all_row_stats_device(d_xyz: int, n_conf: int, n_atoms: int, d_out: int, d_row_min: int,
                                      d_row_max: int, method: Method = Method.BF16X3_GRAM,
                                      ctx: Context | None = None) -> None:
    """All-vs-all on device pointers with f32 output, and in the same call the
    row minima and maxima of that matrix (what ``bs_rmsd_row_stats_device``
    gives for it, bit for bit; ``retrieve_outliers`` and ``medoid_by_min_max``,
    hierarchical.rs:198-217 and 105-131, read them).  The Gram methods take
    them in the kernel's epilogue.  Asynchronous on the context's stream.
    """
    ctx = ctx or default_context()
    check(_native.lib().bs_crmsd_all_vs_all_row_stats_device(
        ctx.handle, ctypes.c_void_p(d_xyz), n_conf, n_atoms, int(method), ctypes.c_void_p(d_out),
        ctypes.c_void_p(d_row_min), ctypes.c_void_p(d_row_max)))


def crmsd_all_vs_all_streamed(models, method: Method = Method.BF16X3_GRAM, tile: int = 4096,
                              ctx: Context | None = None) -> RmsdMatrix:
    """All-vs-all streamed in tile x tile blocks to a host f32 triangle."""
    x = _ensemble(models)
    n, L, _ = x.shape
    ctx = ctx or default_context()
    out = np.empty(max(n * (n - 1) // 2, 0), dtype=np.float32)
    check(_native.lib().bs_crmsd_all_vs_all_streamed(ctx.handle, _dp(x), n, L, int(method), tile,
                                                     out.ctypes.data_as(_native.c_fp)))
    return RmsdMatrix(n, out)


def crmsd_batch(ensembles, method: Method = Method.BF16X3_GRAM,
                ctx: Context | None = None) -> list[np.ndarray]:
    """All-vs-all cRMSD of many independent ensembles (uneven n and L) in one
    call (bs_crmsd_batch).  Returns each ensemble's packed f32 triangle."""
    xs = [_ensemble(e) for e in ensembles]
    ctx = ctx or default_context()
    E = len(xs)
    outs = [np.empty(max(x.shape[0] * (x.shape[0] - 1) // 2, 0), dtype=np.float32) for x in xs]
    px = (_native.c_dp * max(E, 1))(*[_dp(x) for x in xs])
    po = (_native.c_fp * max(E, 1))(*[o.ctypes.data_as(_native.c_fp) for o in outs])
    nc = (ctypes.c_size_t * max(E, 1))(*[x.shape[0] for x in xs])
    na = (ctypes.c_size_t * max(E, 1))(*[x.shape[1] for x in xs])
    check(_native.lib().bs_crmsd_batch(ctx.handle, px, nc, na, E, int(method), po))
    return outs


def crmsd_batch_device(d_xyz, n_conf, n_atoms, d_out, method: Method = Method.BF16X3_GRAM,
                       ctx: Context | None = None) -> None:
    """Batch on device pointers: ``d_xyz[e]`` / ``d_out[e]`` are device addresses
    (e.g. ``tensor.data_ptr()``).  Asynchronous on the context's stream
    (bs_crmsd_batch_device)."""
    ctx = ctx or default_context()
    E = len(n_conf)
    if not (len(d_xyz) == len(n_atoms) == len(d_out) == E):
        raise ValueError("d_xyz, n_conf, n_atoms and d_out must have one entry per ensemble")
    px = (ctypes.c_void_p * max(E, 1))(*[int(p) for p in d_xyz])
    po = (ctypes.c_void_p * max(E, 1))(*[int(p) for p in d_out])
    nc = (ctypes.c_size_t * max(E, 1))(*[int(v) for v in n_conf])
    na = (ctypes.c_size_t * max(E, 1))(*[int(v) for v in n_atoms])
    check(_native.lib().bs_crmsd_batch_device(ctx.handle, px, nc, na, E, int(method), po))


# --------------------------------------------------------------- synthetic
def generate_ensemble(seed: int, n_conf: int, n_atoms: int, sigma=(1.5, 1.5), near_fraction: float = 0.0,
                      near_sigma: float = 0.01) -> np.ndarray:
    """Deterministic synthetic ensemble (see bs_generate_ensemble)."""
    lo, hi = (sigma, sigma) if np.isscalar(sigma) else sigma
    out = np.empty((n_conf, n_atoms, 3), dtype=np.float64)
    check(_native.lib().bs_generate_ensemble(seed, n_conf, n_atoms, lo, hi, near_fraction, near_sigma, _dp(out)))
    return out


# --------------------------------------------------------------- libraries
class Library:
    """Device-resident conformations of one atom count (bs_library)."""

    def __init__(self, xyz, ctx: Context | None = None):
        x = _ensemble(xyz)
        self.ctx = ctx or default_context()
        h = ctypes.c_void_p()
        check(_native.lib().bs_library_create(self.ctx.handle, _dp(x), x.shape[0], x.shape[1], ctypes.byref(h)))
        self._h = h
        self.n, self.n_atoms = x.shape[0], x.shape[1]

    @property
    def handle(self):
        if self._h is None:
            raise ValueError("library is closed")
        return self._h

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            _native.lib().bs_library_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _queries(self, q) -> np.ndarray:
        q = _f64(q)
        if q.ndim == 2:
            q = q[None]
        if q.ndim != 3 or q.shape[1:] != (self.n_atoms, 3):
            raise ValueError(f"queries must be (n_q, {self.n_atoms}, 3), got {q.shape}")
        return np.ascontiguousarray(q)

    def query_all(self, q) -> np.ndarray:
        """cRMSD of every query against every library member: (n_q, n) f32."""
        q = self._queries(q)
        out = np.empty((q.shape[0], self.n), dtype=np.float32)
        check(_native.lib().bs_query_all(self.ctx.handle, self.handle, _dp(q), q.shape[0],
                                         out.ctypes.data_as(_native.c_fp)))
        return out

    def query_topk(self, q, k: int = 100) -> tuple[np.ndarray, np.ndarray]:
        """Per query, the k members of lowest cRMSD: (idx u32, rmsd f32), each (n_q, k)."""
        q = self._queries(q)
        idx = np.empty((q.shape[0], k), dtype=np.uint32)
        val = np.empty((q.shape[0], k), dtype=np.float32)
        check(_native.lib().bs_query_topk(self.ctx.handle, self.handle, _dp(q), q.shape[0], k,
                                          idx.ctypes.data_as(_native.c_u32p), val.ctypes.data_as(_native.c_fp)))
        return idx, val


# --------------------------------------------------------------- serving
class Ticket:
    """A submitted request.  It owns its result buffers; the server copied the
    queries at submit.  Dropping the ticket releases it (bs_server_release): a
    request still running then finishes without writing the buffers."""

    def __init__(self, server: "Server", tid: int, idx: np.ndarray, val: np.ndarray):
        self.server, self.id, self.idx, self.rmsd = server, tid, idx, val

    def release(self) -> None:
        """Gives the ticket up; its id is invalid afterwards."""
        if self.id is not None and self.server._h is not None:
            _native.lib().bs_server_release(self.server._h, self.id)
        self.id = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    def poll(self) -> bool:
        st = _native.lib().bs_server_poll(self.server.handle, self.id)
        if st == -6:
            return False
        check(st)
        return True

    def wait(self) -> tuple[np.ndarray, np.ndarray]:
        check(_native.lib().bs_server_wait(self.server.handle, self.id))
        return self.idx, self.rmsd

    def cancel(self) -> None:
        check(_native.lib().bs_server_cancel(self.server.handle, self.id))

    def latency_ms(self) -> float:
        ms = ctypes.c_double()
        check(_native.lib().bs_server_latency_ms(self.server.handle, self.id, ctypes.byref(ms)))
        return ms.value


class Server:
    """Continuously batching query server over device-resident libraries (bs_server)."""

    def __init__(self, libraries: list[Library], hbm_budget_bytes: int = 0, max_inflight: int = 4,
                 ctx: Context | None = None):
        self.ctx = ctx or libraries[0].ctx
        self.libraries = list(libraries)
        arr = (ctypes.c_void_p * len(libraries))(*[l.handle.value for l in libraries])
        h = ctypes.c_void_p()
        check(_native.lib().bs_server_create(self.ctx.handle, arr, len(libraries), hbm_budget_bytes,
                                             max_inflight, ctypes.byref(h)))
        self._h = h

    @property
    def handle(self):
        if self._h is None:
            raise ValueError("server is closed")
        return self._h

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            _native.lib().bs_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, lib_index: int, q, k: int = 100) -> Ticket:
        q = self.libraries[lib_index]._queries(q)
        idx = np.empty((q.shape[0], k), dtype=np.uint32)
        val = np.empty((q.shape[0], k), dtype=np.float32)
        t = ctypes.c_uint64()
        check(_native.lib().bs_server_submit(self.handle, lib_index, _dp(q), q.shape[0], k,
                                             idx.ctypes.data_as(_native.c_u32p),
                                             val.ctypes.data_as(_native.c_fp), ctypes.byref(t)))
        return Ticket(self, t.value, idx, val)

    def ticket_count(self) -> int:
        """Ticket records the server holds (bs_server_ticket_count)."""
        return _native.lib().bs_server_ticket_count(self.handle)
