"""Loads ``libbioshell_gpu.so`` and declares the C ABI of include/bioshell_gpu.h.

The library must have been built (``__graft_entry__.build()`` or
``python -m bioshell_amd.build``); there is no fallback implementation, so a
missing library is an error.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbioshell_gpu.so")
DIAG_LIB_PATH = os.path.join(HERE, "libbioshell_diag.so")

c_status = ctypes.c_int32
c_dp = ctypes.POINTER(ctypes.c_double)
c_fp = ctypes.POINTER(ctypes.c_float)
c_u32p = ctypes.POINTER(ctypes.c_uint32)
c_u64p = ctypes.POINTER(ctypes.c_uint64)
c_vp = ctypes.c_void_p
c_size = ctypes.c_size_t


class RunStats(ctypes.Structure):
    _fields_ = [
        ("total_ms", ctypes.c_double),
        ("prep_ms", ctypes.c_double),
        ("main_ms", ctypes.c_double),
        ("fallback_ms", ctypes.c_double),
        ("pairs", ctypes.c_uint64),
        ("flagged_pairs", ctypes.c_uint64),
        ("main_launches", ctypes.c_uint64),
        ("tile_m", ctypes.c_int32),
        ("tile_n", ctypes.c_int32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class EnsembleInfo(ctypes.Structure):
    _fields_ = [
        ("n_models", ctypes.c_size_t),
        ("n_atoms", ctypes.c_size_t),
        ("n_structure_atoms", ctypes.c_size_t),
        ("n_chains", ctypes.c_size_t),
        ("n_residues", ctypes.c_size_t),
        ("format", ctypes.c_int),
    ]


class AtomRecord(ctypes.Structure):
    _fields_ = [
        ("serial", ctypes.c_int32),
        ("res_seq", ctypes.c_int32),
        ("occupancy", ctypes.c_double),
        ("temp_factor", ctypes.c_double),
        ("name", ctypes.c_char * 16),
        ("alt_loc", ctypes.c_char * 8),
        ("res_name", ctypes.c_char * 16),
        ("chain_id", ctypes.c_char * 16),
        ("i_code", ctypes.c_char * 8),
        ("element", ctypes.c_char * 16),
        ("has_element", ctypes.c_int32),
        ("hetero", ctypes.c_int32),
    ]


# name -> (restype, argtypes); every symbol declared in include/bioshell_gpu.h
SIGNATURES = {
    "bs_last_error": (ctypes.c_char_p, []),
    "bs_version": (ctypes.c_char_p, []),
    "bs_device_count": (ctypes.c_int, []),
    "bs_context_create": (c_status, [ctypes.c_int, c_size, ctypes.POINTER(c_vp)]),
    "bs_context_destroy": (None, [c_vp]),
    "bs_context_synchronize": (c_status, [c_vp]),
    "bs_context_stream": (c_vp, [c_vp]),
    "bs_context_last_stats": (c_status, [c_vp, ctypes.POINTER(RunStats)]),
    "bs_context_bytes_in_use": (c_size, [c_vp]),
    "bs_context_history": (c_status, [c_vp, ctypes.POINTER(RunStats), c_size, ctypes.POINTER(c_size)]),
    "bs_default_context": (c_status, [ctypes.POINTER(c_vp)]),
    "bs_crmsd_pair_f64": (c_status, [c_vp, c_dp, c_dp, c_size, c_dp]),
    "bs_superpose_f64": (c_status, [c_vp, c_dp, c_dp, c_size, c_dp, c_dp, c_dp]),
    "bs_crmsd_all_vs_all": (c_status, [c_vp, c_dp, c_size, c_size, ctypes.c_int, c_vp, ctypes.c_int]),
    "bs_crmsd_all_vs_all_device": (c_status, [c_vp, c_vp, c_size, c_size, ctypes.c_int, c_vp, ctypes.c_int]),
    "bs_crmsd_all_vs_all_device_noalloc": (c_status, [c_vp, c_vp, c_size, c_size, ctypes.c_int, c_vp,
                                                      ctypes.c_int]),
    "bs_crmsd_all_vs_all_row_stats_device": (c_status, [c_vp, c_vp, c_size, c_size, ctypes.c_int, c_vp, c_vp,
                                                        c_vp]),
    "bs_crmsd_all_vs_all_streamed": (c_status, [c_vp, c_dp, c_size, c_size, ctypes.c_int, c_size, c_fp]),
    "bs_crmsd_batch": (c_status, [c_vp, ctypes.POINTER(c_dp), ctypes.POINTER(c_size), ctypes.POINTER(c_size),
                                  c_size, ctypes.c_int, ctypes.POINTER(c_fp)]),
    "bs_crmsd_batch_device": (c_status, [c_vp, ctypes.POINTER(c_vp), ctypes.POINTER(c_size),
                                         ctypes.POINTER(c_size), c_size, ctypes.c_int, ctypes.POINTER(c_vp)]),
    "bs_autotune": (c_status, [c_vp, ctypes.c_int, ctypes.c_char_p]),
    "bs_autotune_load": (c_status, [c_vp, ctypes.c_char_p]),
    "bs_rmsd_row_stats": (c_status, [c_vp, c_fp, c_size, c_fp, c_fp]),
    "bs_rmsd_row_stats_device": (c_status, [c_vp, c_vp, c_size, c_vp, c_vp]),
    "bs_rmsd_members_row_max": (c_status, [c_vp, c_fp, c_size, c_u32p, c_size, c_fp]),
    "bs_hierarchical_clustering": (c_status, [c_vp, c_fp, c_size, ctypes.c_int, c_u32p, c_u32p, c_fp]),
    "bs_hierarchical_clustering_device": (c_status, [c_vp, c_vp, c_size, ctypes.c_int, c_u32p, c_u32p, c_fp]),
    "bs_library_create": (c_status, [c_vp, c_dp, c_size, c_size, ctypes.POINTER(c_vp)]),
    "bs_library_destroy": (None, [c_vp]),
    "bs_library_size": (c_size, [c_vp]),
    "bs_library_atoms": (c_size, [c_vp]),
    "bs_query_topk": (c_status, [c_vp, c_vp, c_dp, c_size, c_size, c_u32p, c_fp]),
    "bs_query_all": (c_status, [c_vp, c_vp, c_dp, c_size, c_fp]),
    "bs_server_create": (c_status, [c_vp, ctypes.POINTER(c_vp), c_size, c_size, ctypes.c_int, ctypes.POINTER(c_vp)]),
    "bs_server_destroy": (None, [c_vp]),
    "bs_server_submit": (c_status, [c_vp, c_size, c_dp, c_size, c_size, c_u32p, c_fp, c_u64p]),
    "bs_server_poll": (c_status, [c_vp, ctypes.c_uint64]),
    "bs_server_wait": (c_status, [c_vp, ctypes.c_uint64]),
    "bs_server_cancel": (c_status, [c_vp, ctypes.c_uint64]),
    "bs_server_latency_ms": (c_status, [c_vp, ctypes.c_uint64, c_dp]),
    "bs_server_release": (c_status, [c_vp, ctypes.c_uint64]),
    "bs_server_ticket_count": (c_size, [c_vp]),
    "bs_superpose_ensemble": (c_status, [c_vp, c_dp, c_size, c_size, c_dp, c_size, c_dp, c_dp, c_dp, c_dp]),
    "bs_superpose_ensemble_device": (c_status, [c_vp, c_vp, c_size, c_size, c_vp, c_size, c_vp, c_vp, c_vp,
                                                c_vp]),
    "bs_ensemble_parse": (c_status, [ctypes.c_char_p, c_size, ctypes.c_int, ctypes.c_int, ctypes.c_char_p,
                                     ctypes.c_int, ctypes.POINTER(c_vp)]),
    "bs_ensemble_load": (c_status, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                    ctypes.POINTER(c_vp)]),
    "bs_ensemble_destroy": (None, [c_vp]),
    "bs_ensemble_get_info": (c_status, [c_vp, ctypes.POINTER(EnsembleInfo)]),
    "bs_ensemble_copy_coords": (c_status, [c_vp, c_dp]),
    "bs_ensemble_get_atom": (c_status, [c_vp, c_size, ctypes.POINTER(AtomRecord)]),
    "bs_ensemble_upload": (c_status, [c_vp, c_vp, c_vp]),
    "bs_ensemble_write_pdb": (c_status, [c_vp, c_dp, c_size, ctypes.c_char_p]),
    "bs_write_distance_tsv": (c_status, [c_fp, c_size, ctypes.POINTER(ctypes.c_char_p), ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_char_p]),
    "bs_generate_ensemble": (c_status, [ctypes.c_uint64, c_size, c_size, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, c_dp]),
}

_lib = None


def lib() -> ctypes.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build the HIP library first "
                "(python -c 'import __graft_entry__ as g; g.build()')"
            )
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


_diag = None


def diag_lib() -> ctypes.CDLL:
    """The diagnostics library (peak-rate probes, host builds of the device QCP
    routines): used by the benchmark and the CPU tests, never by the package."""
    global _diag
    if _diag is None:
        if not os.path.exists(DIAG_LIB_PATH):
            raise ImportError(f"{DIAG_LIB_PATH} is missing: build the libraries first")
        _diag = ctypes.CDLL(DIAG_LIB_PATH)
    return _diag
