"""The clust-format TSV writer (bs_write_distance_tsv) and reader.

Pins: cities.tsv (bioshell-clustering/tests/test_files, read by the
DataMatrixDistance doc-test at bioshell-clustering/src/hierarchical/
distance_matrix.rs:11-33: 15 elements, d(0, 1) = 1149.36, labels Tokyo and
Seoul) is reproduced byte for byte by read + write; d5.tsv gives the
clustering of test_hierarchical.rs:45-69.  The shortest form is checked
against numpy's shortest round-trip positional formatting of float32 (the
form Rust's `{}` prints), fixed decimals against Python's formatting of the
exact f32 value.  The reading rules of the reference's DataMatrixBuilder
(crate data-matrix, not vendored) beyond these fixtures are unpinned.
"""
import os

import numpy as np
import pytest

import bioshell_amd as B
from bioshell_amd import clustering as C
from oracle import oracle as O
from tests.conftest import GOLDEN


def test_cities_reads_like_the_doc_test_and_writes_back(tmp_path):
    src = os.path.join(GOLDEN, "cities.tsv")
    m, labels = C.read_tsv(src)
    assert m.n_elements() == 15 and labels[:2] == ["Tokyo", "Seoul"]
    assert m.distance(0, 1) == np.float32(1149.36) and m.distance(1, 0) == np.float32(1149.36)
    assert m.distance(0, 0) == 0.0
    out = tmp_path / "c.tsv"
    C.write_tsv(m, str(out), labels=labels)
    # the same pairs in the same order (the fixture lists i < j in row order)
    orig = [l.split("\t") for l in open(src).read().splitlines()]
    got = [l.split("\t") for l in out.read_text().splitlines()]
    assert [r[:2] for r in got] == [r[:2] for r in orig]
    assert [np.float32(float(a[2])) for a in got] == [np.float32(float(b[2])) for b in orig]
    assert sum(r[2] == o[2] for r, o in zip(got, orig)) >= 95  # texts differ only where the fixture wrote "x.0"


def test_d5_clusters(tmp_path):
    m, labels = C.read_tsv(os.path.join(GOLDEN, "d5.tsv"))
    mi, mj, md = O.hclust(m.upper, m.n_elements(), "single")
    root = C.tree_from_merges(m.n_elements(), mi, mj, md)
    clusters = sorted(C.retrieve_clusters(root, 0.11), key=lambda c: c.cluster_size)
    assert [c.cluster_size for c in clusters] == [1, 1, 1, 2]
    # written and read back: the same matrix, hence the same clustering
    out = tmp_path / "d5.tsv"
    C.write_tsv(m, str(out), labels=labels)
    m2, l2 = C.read_tsv(str(out))
    assert l2 == labels and np.array_equal(m2.upper, m.upper)


def values():
    rng = np.random.default_rng(5)
    v = np.concatenate([rng.uniform(0, 20, 3000), 10.0 ** rng.uniform(-12, 12, 2000),
                        [0.0, -0.0, 1.0, 0.1, 1e-7, 3.4028235e38, 1.4e-45, 2.5, 100.0, 1149.36]])
    v[::7] *= -1
    return v.astype(np.float32)


def matrix_of(v):
    # the smallest n whose triangle holds the values
    n = int(np.ceil((1 + np.sqrt(1 + 8 * len(v))) / 2))
    up = np.zeros(n * (n - 1) // 2, np.float32)
    up[:len(v)] = v
    return B.RmsdMatrix(n, up)


def parse(path):
    rows = [l.split("\t") for l in open(path).read().splitlines()]
    return rows


def test_shortest_round_trip_form(tmp_path):
    v = values()
    m = matrix_of(v)
    out = tmp_path / "v.tsv"
    C.write_tsv(m, str(out))
    rows = parse(out)
    assert len(rows) == len(m.upper)
    got = [r[2] for r in rows]
    want = [np.format_float_positional(x, unique=True, trim="-") for x in m.upper]
    assert got == want
    back = np.array([np.float32(float(s)) for s in got], dtype=np.float32)
    assert np.array_equal(back.view(np.uint32), m.upper.view(np.uint32))
    n = m.n_elements()
    assert [(int(r[0]), int(r[1])) for r in rows[:3]] == [(0, 1), (0, 2), (0, 3)]
    assert (int(rows[-1][0]), int(rows[-1][1])) == (n - 2, n - 1)


@pytest.mark.parametrize("decimals", [0, 3, 6, 12])
def test_fixed_decimals(tmp_path, decimals):
    v = values()[:2000]
    m = matrix_of(v)
    out = tmp_path / "f.tsv"
    C.write_tsv(m, str(out), decimals=decimals)
    got = [r[2] for r in parse(out)]
    assert got == [f"{float(x):.{decimals}f}" for x in m.upper]


def test_both_triangles_and_special_values(tmp_path):
    up = np.array([1.5, np.nan, np.inf, -np.inf, 0.25, 7.0], np.float32)
    m = B.RmsdMatrix(4, up)
    out = tmp_path / "b.tsv"
    C.write_tsv(m, str(out), labels=["a", "b", "c", "d"], both_triangles=True)
    rows = parse(out)
    assert len(rows) == 12
    d = {(r[0], r[1]): r[2] for r in rows}
    assert d[("a", "b")] == d[("b", "a")] == "1.5"
    assert d[("a", "c")] == "NaN" and d[("a", "d")] == "inf" and d[("b", "c")] == "-inf"
    m2, _ = C.read_tsv(str(out), symmetric=False)
    assert np.array_equal(m2.upper, m.upper, equal_nan=True)


def test_threads_and_size(tmp_path):
    rng = np.random.default_rng(1)
    n = 2500
    m = B.RmsdMatrix(n, rng.uniform(0.1, 12.0, n * (n - 1) // 2).astype(np.float32))
    a, b = tmp_path / "a.tsv", tmp_path / "b.tsv"
    C.write_tsv(m, str(a), n_threads=1)
    C.write_tsv(m, str(b), n_threads=8)
    assert a.read_bytes() == b.read_bytes()
    vals = np.loadtxt(str(a), usecols=2, dtype=np.float64).astype(np.float32)
    assert np.array_equal(vals, m.upper)


def test_small_and_invalid(tmp_path):
    out = tmp_path / "e.tsv"
    C.write_tsv(B.RmsdMatrix(1, np.zeros(0, np.float32)), str(out))
    assert out.read_bytes() == b""
    m = B.RmsdMatrix(3, np.ones(3, np.float32))
    with pytest.raises(ValueError):
        C.write_tsv(m, str(out), labels=["a", "b"])
    with pytest.raises(ValueError):
        C.write_tsv(m, str(out), labels=["a", "b\tc", "d"])
    with pytest.raises(B.BioshellGpuError):
        C.write_tsv(m, str(tmp_path / "missing_dir" / "x.tsv"))
    with pytest.raises(B.BioshellGpuError):
        C.write_tsv(m, str(out), decimals=-2)
