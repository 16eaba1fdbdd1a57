"""Conformational ensembles read from multi-model structure files, and their
superposition onto a reference.

Host mirror of the ``bs_ensemble_*`` and ``bs_superpose_ensemble*`` entry
points.  Reading follows what BioShell does for a multi-model deposit:
``Deposit::from_file`` (bioshell-pdb/src/deposit.rs:81-85) builds a structure
whose atoms come from the first model, and the caller sets every model in turn
(``Structure::set_model``, structure.rs:277-285) and keeps the atoms an atom
predicate accepts (``IsCA`` and friends, bioshell-pdb/src/pdb_atom_filters.rs).
The native reader does all of it in one multi-threaded pass and hands back the
``(n_models, n_atoms, 3)`` float64 array every cRMSD entry point takes.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

from . import _native
from .crmsd import Context, Rototranslation, _dp, _f64, check, default_context

FORMATS = {"auto": 0, "cif": 1, "pdb": 2}
SELECTIONS = {"all": 0, "ca": 1, "cb": 2, "not_hydrogen": 3}


@dataclass
class Atom:
    """One atom record of the first model (``bioshell_pdb::PdbAtom``)."""
    serial: int
    name: str
    alt_loc: str
    res_name: str
    chain_id: str
    res_seq: int
    i_code: str
    occupancy: float
    temp_factor: float
    element: str | None
    is_hetero_atom: bool


class Ensemble:
    """A conformational ensemble: atom records and ``xyz`` of shape
    ``(n_models, n_atoms, 3)``.  Build it with :meth:`from_file` or
    :meth:`from_text`."""

    def __init__(self, handle: int):
        self._h = ctypes.c_void_p(handle)
        info = _native.EnsembleInfo()
        check(_native.lib().bs_ensemble_get_info(self._h, ctypes.byref(info)))
        self.n_models = int(info.n_models)
        self.n_atoms = int(info.n_atoms)
        self.n_structure_atoms = int(info.n_structure_atoms)
        self.n_chains = int(info.n_chains)
        self.n_residues = int(info.n_residues)
        self.format = {1: "cif", 2: "pdb"}[int(info.format)]
        self._xyz = None
        self._atoms = None

    @staticmethod
    def _args(select, chain, fmt):
        if select not in SELECTIONS:
            raise ValueError(f"unknown selection {select!r}; one of {sorted(SELECTIONS)}")
        if fmt not in FORMATS:
            raise ValueError(f"unknown format {fmt!r}; one of {sorted(FORMATS)}")
        return FORMATS[fmt], SELECTIONS[select], (chain.encode() if chain is not None else None)

    @classmethod
    def from_file(cls, path: str, select: str = "ca", chain: str | None = None, format: str = "auto",
                  n_threads: int = 0) -> "Ensemble":
        """Reads an mmCIF or PDB file (``.gz`` is inflated)."""
        f, s, c = cls._args(select, chain, format)
        h = ctypes.c_void_p()
        check(_native.lib().bs_ensemble_load(os.fsencode(path), f, s, c, n_threads, ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def from_text(cls, text, select: str = "ca", chain: str | None = None, format: str = "auto",
                  n_threads: int = 0) -> "Ensemble":
        """Reads mmCIF or PDB text (``str`` or ``bytes``)."""
        data = text.encode() if isinstance(text, str) else bytes(text)
        f, s, c = cls._args(select, chain, format)
        h = ctypes.c_void_p()
        check(_native.lib().bs_ensemble_parse(data, len(data), f, s, c, n_threads, ctypes.byref(h)))
        return cls(h.value)

    # the reference's Structure counts
    def count_models(self) -> int:
        return self.n_models

    def count_chains(self) -> int:
        return self.n_chains

    def count_residues(self) -> int:
        return self.n_residues

    @property
    def xyz(self) -> np.ndarray:
        if self._xyz is None:
            out = np.empty((self.n_models, self.n_atoms, 3), dtype=np.float64)
            check(_native.lib().bs_ensemble_copy_coords(self._h, _dp(out)))
            self._xyz = out
        return self._xyz

    @property
    def atoms(self) -> list[Atom]:
        if self._atoms is None:
            rec = _native.AtomRecord()
            out = []
            for k in range(self.n_atoms):
                check(_native.lib().bs_ensemble_get_atom(self._h, k, ctypes.byref(rec)))
                d = lambda b: b.decode("utf-8")  # noqa: E731
                out.append(Atom(rec.serial, d(rec.name), d(rec.alt_loc), d(rec.res_name), d(rec.chain_id),
                                rec.res_seq, d(rec.i_code), rec.occupancy, rec.temp_factor,
                                d(rec.element) if rec.has_element else None, bool(rec.hetero)))
            self._atoms = out
        return self._atoms

    def upload(self, d_xyz: int, ctx: Context | None = None) -> None:
        """Copies the coordinates to device memory at ``d_xyz``."""
        ctx = ctx or default_context()
        check(_native.lib().bs_ensemble_upload(ctx.handle, self._h, ctypes.c_void_p(d_xyz)))

    def write_pdb(self, path: str, xyz=None) -> None:
        """Writes the models (or ``xyz``, e.g. superposed coordinates) in PDB
        format with the reference's ``write_pdb`` records; several models are
        framed by MODEL / ENDMDL."""
        if xyz is None:
            check(_native.lib().bs_ensemble_write_pdb(self._h, None, 0, os.fsencode(path)))
            return
        x = _f64(xyz)
        if x.ndim == 2:
            x = x[None]
        if x.shape[1:] != (self.n_atoms, 3):
            raise ValueError(f"expected (n_models, {self.n_atoms}, 3), got {x.shape}")
        check(_native.lib().bs_ensemble_write_pdb(self._h, _dp(x), x.shape[0], os.fsencode(path)))

    def close(self) -> None:
        if self._h:
            _native.lib().bs_ensemble_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def load_ensemble(path: str, select: str = "ca", chain: str | None = None, format: str = "auto",
                  n_threads: int = 0) -> Ensemble:
    """``Ensemble.from_file``."""
    return Ensemble.from_file(path, select=select, chain=chain, format=format, n_threads=n_threads)


@dataclass
class EnsembleFit:
    """Per-model superposition onto a reference: ``rotation`` (n, 3, 3),
    ``translation`` (n, 3) with ref ~= R x + t, ``rmsd`` (n,) and the
    superposed coordinates (n, L, 3) when requested."""
    rotation: np.ndarray
    translation: np.ndarray
    rmsd: np.ndarray
    superposed: np.ndarray | None

    def rototranslation(self, m: int) -> Rototranslation:
        return Rototranslation(self.rotation[m], self.translation[m])


def superpose_ensemble(xyz, reference=None, ref_index: int = 0, coordinates: bool = True,
                       ctx: Context | None = None) -> EnsembleFit:
    """Superposes every conformation of ``xyz`` (n, L, 3) onto ``reference``
    (L, 3), or onto conformation ``ref_index`` (bs_superpose_ensemble)."""
    x = _f64(xyz)
    if x.ndim != 3 or x.shape[2] != 3:
        raise ValueError(f"expected (n, L, 3), got {x.shape}")
    n, L = x.shape[:2]
    ref = None
    if reference is not None:
        ref = _f64(reference)
        if ref.shape != (L, 3):
            raise ValueError(f"expected a ({L}, 3) reference, got {ref.shape}")
    ctx = ctx or default_context()
    rot = np.empty((n, 3, 3))
    tr = np.empty((n, 3))
    rm = np.empty(n)
    sp = np.empty_like(x) if coordinates else None
    check(_native.lib().bs_superpose_ensemble(ctx.handle, _dp(x), n, L, _dp(ref) if ref is not None else None,
                                               ref_index, _dp(rot), _dp(tr), _dp(rm),
                                               _dp(sp) if sp is not None else None))
    return EnsembleFit(rot, tr, rm, sp)


def superpose_ensemble_device(d_xyz: int, n_conf: int, n_atoms: int, d_ref: int | None = None,
                              ref_index: int = 0, d_rot: int | None = None, d_trans: int | None = None,
                              d_rmsd: int | None = None, d_superposed: int | None = None,
                              ctx: Context | None = None) -> None:
    """Device-pointer form, asynchronous on the context's stream."""
    ctx = ctx or default_context()
    vp = lambda p: ctypes.c_void_p(p) if p else None  # noqa: E731
    check(_native.lib().bs_superpose_ensemble_device(ctx.handle, vp(d_xyz), n_conf, n_atoms, vp(d_ref), ref_index,
                                                      vp(d_rot), vp(d_trans), vp(d_rmsd), vp(d_superposed)))
