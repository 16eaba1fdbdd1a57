"""Writes the committed golden fixtures under tests/golden/.

Run in the development container (it reads the reference's test data from
/root/reference, which does not exist on GPU hosts):

    python tests/golden/make_golden.py

Fixtures (all data: inputs plus expected outputs):

* ``2jqb_ensembles.npz`` -- the 20 NMR models of 2jqb
  (bioshell-pdb/tests/test_files/2jqb.cif; the reference's multi-model test
  at bioshell-pdb/tests/test_structure.rs:156-166 loads the same file), as
  CA-only (20 x n_ca x 3) and all-ATOM (20 x n_atom x 3) coordinate arrays,
  with their all-vs-all cRMSD (packed upper triangle, f64).
* ``config0.npz`` -- all-vs-all cRMSD of the synthetic 200 x 76 ensemble
  (seed 0, sigma 1.5 A) and a SHA-256 of the generated coordinates.
* ``d5.tsv`` -- copied verbatim from bioshell-clustering/tests/test_files/d5.tsv
  (a distance-matrix fixture the reference's clustering tests read).

Expected cRMSD values come from numpy's SVD (Kabsch: lambda = s1 + s2 +
sign(det M) s3), computed independently of the project's QCP oracle.
"""
from __future__ import annotations

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


def read_2jqb(path: str):
    """CA and all-ATOM coordinates per model from the mmCIF _atom_site loop."""
    cols, rows = [], []
    in_loop = False
    with open(path) as f:
        for line in f:
            if line.startswith("_atom_site."):
                cols.append(line.strip().split(".", 1)[1])
                in_loop = True
                continue
            if in_loop and (line.startswith("ATOM") or line.startswith("HETATM")):
                rows.append(line.split())
            elif in_loop and rows and line.startswith("#"):
                break
    ix = {c: k for k, c in enumerate(cols)}
    models: dict[int, dict[str, list]] = {}
    for r in rows:
        if r[ix["group_PDB"]] != "ATOM":
            continue
        m = int(r[ix["pdbx_PDB_model_num"]])
        xyz = [float(r[ix["Cartn_x"]]), float(r[ix["Cartn_y"]]), float(r[ix["Cartn_z"]])]
        d = models.setdefault(m, {"ca": [], "all": []})
        d["all"].append(xyz)
        if r[ix["label_atom_id"]] == "CA":
            d["ca"].append(xyz)
    keys = sorted(models)
    ca = np.array([models[k]["ca"] for k in keys])
    allat = np.array([models[k]["all"] for k in keys])
    return ca, allat


def main() -> None:
    src = os.path.join(REF, "bioshell-pdb/tests/test_files/2jqb.cif")
    ca, allat = read_2jqb(src)
    np.savez_compressed(
        os.path.join(HERE, "2jqb_ensembles.npz"),
        ca=ca, ca_rmsd=kabsch_all_vs_all(ca),
        all_atom=allat, all_atom_rmsd=kabsch_all_vs_all(allat),
    )
    print("2jqb:", ca.shape, allat.shape)

    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    import bioshell_amd as B  # the generator is host code; no GPU needed

    x = B.generate_ensemble(0, 200, 76, sigma=1.5)
    np.savez_compressed(
        os.path.join(HERE, "config0.npz"),
        rmsd=kabsch_all_vs_all(x),
        sha256=np.frombuffer(hashlib.sha256(x.tobytes()).digest(), dtype=np.uint8),
        first_atoms=x[:2, :4].copy(),
    )
    print("config0:", x.shape)

    shutil.copyfile(os.path.join(REF, "bioshell-clustering/tests/test_files/d5.tsv"),
                    os.path.join(HERE, "d5.tsv"))


if __name__ == "__main__":
    main()
