"""Writes the committed golden fixtures under tests/golden/.

Run in the development container (it reads the reference's test data from
/root/reference, which does not exist on GPU hosts):

    python tests/golden/make_golden.py

Fixtures (all data: inputs plus expected outputs):

* ``structures/*.gz`` -- the reference's structure test files
  (bioshell-pdb/tests/test_files/), gzip-compressed data: multi-model 2jqb,
  2gb1 as both mmCIF and PDB, and files with chains, alternate locations,
  insertion codes, quoted atom names and ions.
* ``2jqb_ensembles.npz`` -- the 20 NMR models of 2jqb (the reference's
  multi-model test at bioshell-pdb/tests/test_structure.rs:156-166 loads the
  same file, 255 atoms per model), read as the reference reads them (first
  model's atoms, every model's coordinates, ATOM and HETATM records) with the
  IsCA predicate (20 x 15 x 3) and with all atoms (20 x 255 x 3), with their
  all-vs-all cRMSD (packed upper triangle, f64).
* ``config0.npz`` -- all-vs-all cRMSD of the synthetic 200 x 76 ensemble
  (seed 0, sigma 1.5 A) and a SHA-256 of the generated coordinates.
* ``d5.tsv``, ``cities.tsv`` -- copied verbatim from
  bioshell-clustering/tests/test_files/ (the distance-matrix fixtures the
  reference's clustering tests and DataMatrixDistance doc-test read).

Expected cRMSD values come from numpy's SVD (Kabsch: lambda = s1 + s2 +
sign(det M) s3), computed independently of the project's QCP oracle.
"""
from __future__ import annotations

import gzip
import hashlib
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"


def kabsch_all_vs_all(x: np.ndarray) -> np.ndarray:
    """Packed upper triangle of cRMSD by batched numpy SVD (float64)."""
    n, L, _ = x.shape
    c = x - x.mean(axis=1, keepdims=True)
    g = np.einsum("nkc,nkc->n", c, c)
    iu, ju = np.triu_indices(n, 1)
    out = np.empty(iu.size)
    step = 4096
    for s in range(0, iu.size, step):
        i, j = iu[s:s + step], ju[s:s + step]
        M = np.einsum("pkc,pkd->pcd", c[i], c[j])
        sv = np.linalg.svd(M, compute_uv=False)
        d = np.sign(np.linalg.det(M))
        d[d == 0] = 1.0
        lam = sv[:, 0] + sv[:, 1] + d * sv[:, 2]
        out[s:s + step] = np.sqrt(np.maximum(0.0, (g[i] + g[j] - 2.0 * lam) / L))
    return out


STRUCTURES = ("2jqb.cif", "2gb1.cif", "2gb1.pdb", "2fdo.cif", "1c5n.cif", "5edw.cif", "6ins.cif",
              "1akd_fragm.pdb", "alphaHelix.pdb")


def main() -> None:
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    from oracle import ingest_oracle as I

    sdir = os.path.join(HERE, "structures")
    os.makedirs(sdir, exist_ok=True)
    for name in STRUCTURES:
        src = os.path.join(REF, "bioshell-pdb/tests/test_files", name)
        with open(src, "rb") as f, gzip.GzipFile(os.path.join(sdir, name + ".gz"), "wb", mtime=0) as g:
            g.write(f.read())
    text = I.read_text(os.path.join(sdir, "2jqb.cif.gz"))
    ca = I.ensemble(text, selection="ca").xyz
    allat = I.ensemble(text, selection="all").xyz
    np.savez_compressed(
        os.path.join(HERE, "2jqb_ensembles.npz"),
        ca=ca, ca_rmsd=kabsch_all_vs_all(ca),
        all_atom=allat, all_atom_rmsd=kabsch_all_vs_all(allat),
    )
    print("2jqb:", ca.shape, allat.shape)

    import bioshell_amd as B  # the generator is host code; no GPU needed

    x = B.generate_ensemble(0, 200, 76, sigma=1.5)
    np.savez_compressed(
        os.path.join(HERE, "config0.npz"),
        rmsd=kabsch_all_vs_all(x),
        sha256=np.frombuffer(hashlib.sha256(x.tobytes()).digest(), dtype=np.uint8),
        first_atoms=x[:2, :4].copy(),
    )
    print("config0:", x.shape)

    for name in ("d5.tsv", "cities.tsv"):
        shutil.copyfile(os.path.join(REF, "bioshell-clustering/tests/test_files", name), os.path.join(HERE, name))


if __name__ == "__main__":
    main()
