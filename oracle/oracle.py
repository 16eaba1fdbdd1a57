"""ctypes loader for the FP64 CPU oracle (``liboracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of ``bench.py`` -- always as the checker or the
timed CPU baseline, never as part of the product path.

The C source (``crmsd_oracle.c``) restates the optimal-superposition cRMSD
that the reference intends for ``bioshell_pdb::calc`` (bioshell-pdb/Cargo.toml:27)
in the reference's conventions (bioshell-core/src/vec3.rs:43-51,
bioshell-core/src/matrix.rs:7-12, bioshell-pdb/src/calc/substructures.rs:27-45).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.or_crmsd.restype = ctypes.c_double
        L.or_crmsd.argtypes = [_dp, _dp, ctypes.c_long]
        L.or_crmsd_kabsch.restype = ctypes.c_double
        L.or_crmsd_kabsch.argtypes = [_dp, _dp, ctypes.c_long]
        L.or_centre.restype = None
        L.or_centre.argtypes = [_dp, ctypes.c_long, _dp, _dp]
        L.or_covariance.restype = None
        L.or_covariance.argtypes = [_dp, _dp, ctypes.c_long, _dp, _dp, _dp]
        L.or_qcp_lambda.restype = ctypes.c_double
        L.or_qcp_lambda.argtypes = [_dp, ctypes.c_double, ctypes.POINTER(ctypes.c_int)]
        L.or_lambda_jacobi4.restype = ctypes.c_double
        L.or_lambda_jacobi4.argtypes = [_dp]
        L.or_lambda_kabsch.restype = ctypes.c_double
        L.or_lambda_kabsch.argtypes = [_dp]
        L.or_key_matrix.restype = None
        L.or_key_matrix.argtypes = [_dp, _dp]
        L.or_superpose.restype = ctypes.c_double
        L.or_superpose.argtypes = [_dp, _dp, ctypes.c_long, _dp, _dp]
        L.or_all_vs_all.restype = ctypes.c_int
        L.or_all_vs_all.argtypes = [_dp, ctypes.c_long, ctypes.c_long, _dp, ctypes.c_int]
        L.or_all_vs_all_rows.restype = ctypes.c_int
        L.or_all_vs_all_rows.argtypes = [_dp, ctypes.c_long, ctypes.c_long, ctypes.c_long,
                                         ctypes.c_long, _dp, ctypes.c_int]
        L.or_query_vs_library.restype = ctypes.c_int
        L.or_query_vs_library.argtypes = [_dp, ctypes.c_long, _dp, ctypes.c_long, ctypes.c_long, _dp]
        _lib = L
    return _lib


def _f64(x) -> np.ndarray:
    return np.ascontiguousarray(x, dtype=np.float64)


def _p(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def crmsd(a, b) -> float:
    a, b = _f64(a), _f64(b)
    assert a.shape == b.shape and a.ndim == 2 and a.shape[1] == 3
    return lib().or_crmsd(_p(a), _p(b), a.shape[0])


def crmsd_kabsch(a, b) -> float:
    a, b = _f64(a), _f64(b)
    assert a.shape == b.shape
    return lib().or_crmsd_kabsch(_p(a), _p(b), a.shape[0])


def centre(xyz):
    xyz = _f64(xyz)
    out = np.empty_like(xyz)
    c = np.empty(3)
    lib().or_centre(_p(xyz), xyz.shape[0], _p(out), _p(c))
    return out, c


def covariance(a, b):
    a, b = _f64(a), _f64(b)
    M = np.empty(9)
    ga = ctypes.c_double()
    gb = ctypes.c_double()
    lib().or_covariance(_p(a), _p(b), a.shape[0], _p(M), ctypes.byref(ga), ctypes.byref(gb))
    return M.reshape(3, 3), ga.value, gb.value


def qcp_lambda(M, E0):
    M = _f64(M).reshape(9)
    it = ctypes.c_int()
    lam = lib().or_qcp_lambda(_p(M), float(E0), ctypes.byref(it))
    return lam, it.value


def lambda_jacobi4(M) -> float:
    M = _f64(M).reshape(9)
    return lib().or_lambda_jacobi4(_p(M))


def lambda_kabsch(M) -> float:
    M = _f64(M).reshape(9)
    return lib().or_lambda_kabsch(_p(M))


def key_matrix(M) -> np.ndarray:
    M = _f64(M).reshape(9)
    K = np.empty(16)
    lib().or_key_matrix(_p(M), _p(K))
    return K.reshape(4, 4)


def superpose(a, b):
    """Returns (R, t, rmsd) with a ~= b @ R.T + t."""
    a, b = _f64(a), _f64(b)
    R = np.empty(9)
    t = np.empty(3)
    r = lib().or_superpose(_p(a), _p(b), a.shape[0], _p(R), _p(t))
    return R.reshape(3, 3), t, r


def all_vs_all(xyz, nthreads: int | None = None) -> np.ndarray:
    """Packed upper triangle (row-major, i<j) of all-vs-all cRMSD, float64."""
    xyz = _f64(xyz)
    n, L, _ = xyz.shape
    out = np.empty(n * (n - 1) // 2)
    if nthreads is None:
        nthreads = os.cpu_count() or 1
    rc = lib().or_all_vs_all(_p(xyz), n, L, _p(out), int(nthreads))
    if rc != 0:
        raise ValueError(f"or_all_vs_all failed ({rc})")
    return out


def all_vs_all_rows(xyz, row0: int, row1: int, nthreads: int | None = None) -> np.ndarray:
    """Dense rows [row0,row1) x n; entries j<=i are left as NaN."""
    xyz = _f64(xyz)
    n, L, _ = xyz.shape
    out = np.full((row1 - row0, n), np.nan)
    if nthreads is None:
        nthreads = os.cpu_count() or 1
    rc = lib().or_all_vs_all_rows(_p(xyz), n, L, row0, row1, _p(out), int(nthreads))
    if rc != 0:
        raise ValueError(f"or_all_vs_all_rows failed ({rc})")
    return out


def query_vs_library(q_xyz, lib_xyz) -> np.ndarray:
    q, l = _f64(q_xyz), _f64(lib_xyz)
    out = np.empty((q.shape[0], l.shape[0]))
    lib().or_query_vs_library(_p(q), q.shape[0], _p(l), l.shape[0], q.shape[1], _p(out))
    return out


def tri_index(n: int, i, j):
    """Packed upper-triangle index of (i, j), i < j."""
    return i * n - i * (i + 1) // 2 + (j - i - 1)


# ----------------------------------------------------------- clustering
LINKAGES = ("single", "complete", "average", "median", "centroid", "wards")


def hclust(upper, n: int, linkage: str, planted_mistake: int = 0):
    """Merge sequence of BioShell's agglomerative clustering (cluster_oracle.c):
    arrays i, j (current matrix indices, i < j) and the merging distance of
    each of the n - 1 steps.  ``planted_mistake`` 1 moves sizes with the row
    on a move, 2 passes sizes[k] to the rule (for tests that show a check
    rejects them); 0 is the reference's behaviour."""
    up = np.ascontiguousarray(upper, dtype=np.float32)
    if up.shape != (n * (n - 1) // 2,):
        raise ValueError("packed triangle has the wrong length")
    L = lib()
    fn = L.or_hclust_variant
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_long, ctypes.c_int, ctypes.c_int,
                   ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                   ctypes.POINTER(ctypes.c_float)]
    m = max(n - 1, 0)
    mi = np.zeros(m, dtype=np.int32)
    mj = np.zeros(m, dtype=np.int32)
    md = np.zeros(m, dtype=np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    ip = ctypes.POINTER(ctypes.c_int)
    if fn(up.ctypes.data_as(fp), n, LINKAGES.index(linkage), int(planted_mistake), mi.ctypes.data_as(ip),
          mj.ctypes.data_as(ip), md.ctypes.data_as(fp)) != 0:
        raise MemoryError("or_hclust")
    return mi, mj, md


def row_stats(upper, n: int):
    """Row minimum and maximum over j != i of a packed f32 triangle."""
    up = np.ascontiguousarray(upper, dtype=np.float32)
    L = lib()
    fn = L.or_row_stats
    fp = ctypes.POINTER(ctypes.c_float)
    fn.restype = None
    fn.argtypes = [fp, ctypes.c_long, fp, fp]
    lo = np.zeros(n, dtype=np.float32)
    hi = np.zeros(n, dtype=np.float32)
    fn(up.ctypes.data_as(fp), n, lo.ctypes.data_as(fp), hi.ctypes.data_as(fp))
    return lo, hi

