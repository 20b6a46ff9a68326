"""ctypes binding of libmf.so (the C-ABI declared in include/mf.h).

The library is loaded from this package directory.  There is no fallback: if
``libmf.so`` is missing or does not load, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
import pathlib

import numpy as np

PKG_DIR = pathlib.Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libmf.so"

MF_SUCCESS = 0
STATUS_NAMES = {
    0: "MF_SUCCESS", 1: "MF_ERR_INVALID", 2: "MF_ERR_HIP", 3: "MF_ERR_OCCUPANCY",
    4: "MF_NOT_CONVERGED", 5: "MF_BREAKDOWN", 6: "MF_TIMEOUT", 7: "MF_ERR_ALLOC",
    8: "MF_ERR_NO_DEVICE", 9: "MF_ERR_TOO_LARGE", 10: "MF_ERR_UNSUPPORTED",
}
MF_PREC_FP64, MF_PREC_MIXED = 0, 1
MF_CLASS_LOSSLESS, MF_CLASS_ROWSCALED, MF_CLASS_ALL_FP64 = 0, 1, 2


class MfError(RuntimeError):
    def __init__(self, status, where=""):
        self.status = int(status)
        super().__init__(f"{where}: {STATUS_NAMES.get(self.status, self.status)}")


i32p = ctypes.POINTER(ctypes.c_int)
u8p = ctypes.POINTER(ctypes.c_ubyte)
f64p = ctypes.POINTER(ctypes.c_double)
f32p = ctypes.POINTER(ctypes.c_float)
vp = ctypes.c_void_p


class TileMatrix(ctypes.Structure):
    """Mirror of mf_tile_matrix (the reference Tile_matrix layout)."""
    _fields_ = [
        ("tilem", ctypes.c_int), ("tilen", ctypes.c_int), ("tilenum", ctypes.c_int),
        ("tile_ptr", i32p), ("tile_columnidx", i32p), ("tile_nnz", i32p),
        ("Format", ctypes.POINTER(ctypes.c_byte)), ("blknnz", i32p), ("blknnznnz", u8p),
        ("dnsrowptr", i32p), ("dnscolptr", i32p), ("tilewidth", ctypes.POINTER(ctypes.c_byte)),
        ("csr_offset", i32p), ("csrptr_offset", i32p), ("coo_offset", i32p), ("ell_offset", i32p),
        ("hyb_offset", i32p), ("hyb_coocount", i32p), ("dns_offset", i32p), ("dnsrow_offset", i32p),
        ("dnscol_offset", i32p), ("new_coocount", i32p),
        ("Blockcsr_Val", f64p), ("Blockcsr_Val_Low", f32p), ("Tile_csr_Col", u8p),
        ("Blockcsr_Val_mix", vp), ("Blockcsr_Ptr", u8p), ("csr_compressedIdx", u8p),
        ("csrsize", ctypes.c_int), ("csrptrlen", ctypes.c_int),
        ("Blockcoo_Val", f64p), ("Blockcoo_Val_Low", f32p), ("coo_compressed_Idx", u8p), ("coosize", ctypes.c_int),
        ("Blockell_Val", f64p), ("Blockell_Val_Low", f32p), ("ell_compressedIdx", u8p), ("ellsize", ctypes.c_int),
        ("Blockhyb_Val", f64p), ("Blockhyb_Val_Low", f32p), ("hybIdx", u8p), ("hybsize", ctypes.c_int),
        ("hybellsize", ctypes.c_int), ("hybcoosize", ctypes.c_int),
        ("Blockdense_Val", f64p), ("Blockdense_Val_Low", f32p), ("dnssize", ctypes.c_int),
        ("Blockdenserow_Val", f64p), ("Blockdenserow_Val_Low", f32p), ("denserowid", vp), ("dnsrowsize", ctypes.c_int),
        ("Blockdensecol_Val", f64p), ("Blockdensecol_Val_Low", f32p), ("densecolid", vp), ("dnscolsize", ctypes.c_int),
        ("coototal", ctypes.c_int), ("deferredcoo_val", f64p), ("deferredcoo_val_Low", f32p),
        ("deferredcoo_colidx", i32p), ("deferredcoo_ptr", i32p), ("rowblkblock", ctypes.c_int),
        ("flag_bal_tile_rowidx", ctypes.POINTER(ctypes.c_uint)), ("tile_bal_rowidx_colstart", i32p),
        ("tile_bal_rowidx_colstop", i32p), ("flag_tilerow_start", i32p), ("flag_tilerow_stop", i32p),
        ("tile_bal_rowidx_colstart_v2", i32p), ("tile_bal_rowidx_colstop_v2", i32p), ("map", i32p),
    ]


class MatrixOptions(ctypes.Structure):
    _fields_ = [("class_policy", ctypes.c_int), ("device", ctypes.c_int), ("keep_mixed", ctypes.c_int)]


class MatrixStats(ctypes.Structure):
    _fields_ = [
        ("n_rows", ctypes.c_int), ("n_cols", ctypes.c_int), ("nnz", ctypes.c_int64), ("tiles", ctypes.c_int64),
        ("chunks", ctypes.c_int64), ("tilem", ctypes.c_int), ("tilen", ctypes.c_int),
        ("tiles_by_class", ctypes.c_int64 * 4), ("nnz_by_class", ctypes.c_int64 * 4),
        ("alg_bytes_fp64", ctypes.c_double), ("alg_bytes_mixed", ctypes.c_double), ("device_bytes", ctypes.c_double),
        ("positions", ctypes.c_int64),
    ]


class SolveOptions(ctypes.Structure):
    _fields_ = [
        ("tol", ctypes.c_double), ("max_iter", ctypes.c_int), ("precision", ctypes.c_int),
        ("fixed_iterations", ctypes.c_int), ("dynamic_precision", ctypes.c_int), ("replace_delta", ctypes.c_double),
        ("max_replacements", ctypes.c_int), ("grid", ctypes.c_int), ("block", ctypes.c_int), ("variant", ctypes.c_int),
        ("use_x0", ctypes.c_int), ("watchdog_ms", ctypes.c_double), ("transpose", ctypes.c_int),
        ("layout", ctypes.c_int), ("value_cache", ctypes.c_int),
    ]


class SolveInfo(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int), ("iterations", ctypes.c_int), ("replacements", ctypes.c_int),
        ("grid", ctypes.c_int), ("block", ctypes.c_int), ("rows_per_thread", ctypes.c_int),
        ("relres", ctypes.c_double), ("relres_true", ctypes.c_double), ("time_ms", ctypes.c_double),
        ("bytes_per_iter", ctypes.c_double), ("gbs", ctypes.c_double), ("layout", ctypes.c_int),
        ("variant", ctypes.c_int), ("precision", ctypes.c_int),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_LIB = None

# name -> (restype, argtypes)
_SIGS = {
    "mf_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "mf_tile_create": (ctypes.c_int, [ctypes.POINTER(TileMatrix), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      i32p, i32p, f64p, f32p]),
    "mf_tile_destroy": (None, [ctypes.POINTER(TileMatrix)]),
    "mf_tile_prep_tables": (ctypes.c_int, [ctypes.POINTER(TileMatrix), ctypes.c_int, i32p, i32p, i32p,
                                           ctypes.POINTER(ctypes.c_uint), i32p, i32p]),
    "mf_tile_row_lists": (ctypes.c_int, [ctypes.POINTER(TileMatrix), ctypes.c_int, i32p, u8p, u8p, i32p]),
    "mf_encode": (None, [u8p, u8p, ctypes.c_int, ctypes.c_int]),
    "mf_decode": (None, [u8p, u8p, ctypes.c_int]),
    "mf_value_copies": (None, [f64p, ctypes.c_int64, f32p, ctypes.POINTER(ctypes.c_uint16),
                               ctypes.POINTER(ctypes.c_int8)]),
    "mf_tile_classes": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, i32p, i32p, f64p, ctypes.c_int, u8p,
                                       ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]),
    "mf_read_matrix": (ctypes.c_int, [ctypes.c_char_p, i32p, i32p, i32p, ctypes.POINTER(i32p),
                                      ctypes.POINTER(i32p), ctypes.POINTER(f64p), i32p]),
    "mf_write_cbd": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p, i32p, f64p]),
    "mf_free": (None, [vp]),
    "mf_solve_options_default": (None, [ctypes.POINTER(SolveOptions)]),
    "mf_matrix_options_default": (None, [ctypes.POINTER(MatrixOptions)]),
    "mf_matrix_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, i32p, i32p, f64p,
                                        ctypes.POINTER(MatrixOptions), vp, ctypes.POINTER(vp)]),
    "mf_matrix_destroy": (ctypes.c_int, [vp]),
    "mf_matrix_get_stats": (ctypes.c_int, [vp, ctypes.POINTER(MatrixStats)]),
    "mf_matrix_update_values": (ctypes.c_int, [vp, f64p, vp]),
    "mf_spmv": (ctypes.c_int, [vp, vp, vp, ctypes.c_int, ctypes.c_int, vp]),
    "mf_dyn_classify": (ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, vp, vp]),
    "mf_pattern_outer": (ctypes.c_int, [vp, vp, vp, ctypes.c_double, vp, vp]),
    "mf_cg_solve": (ctypes.c_int, [vp, vp, vp, ctypes.POINTER(SolveOptions), ctypes.POINTER(SolveInfo), vp]),
    "mf_cg_solve_noalloc": (ctypes.c_int, [vp, vp, vp, ctypes.POINTER(SolveOptions), ctypes.POINTER(SolveInfo),
                                           vp]),
    "mf_bicgstab_solve": (ctypes.c_int, [vp, vp, vp, ctypes.POINTER(SolveOptions), ctypes.POINTER(SolveInfo), vp]),
    "mf_cg_autotune": (ctypes.c_int, [vp, vp, vp, ctypes.c_int, i32p, i32p, f64p, vp]),
    "mf_cg_solve_host": (ctypes.c_int, [vp, f64p, f64p, ctypes.POINTER(SolveOptions), ctypes.POINTER(SolveInfo)]),
    "mf_bicgstab_solve_host": (ctypes.c_int, [vp, f64p, f64p, ctypes.POINTER(SolveOptions),
                                              ctypes.POINTER(SolveInfo)]),
    "mf_cg_solve_multikernel": (ctypes.c_int, [vp, vp, vp, ctypes.POINTER(SolveOptions),
                                               ctypes.POINTER(SolveInfo), vp]),
    "mf_bicgstab_solve_multikernel": (ctypes.c_int, [vp, vp, vp, ctypes.POINTER(SolveOptions),
                                                     ctypes.POINTER(SolveInfo), vp]),
    "mf_batch_create": (ctypes.c_int, [ctypes.c_int, i32p, ctypes.POINTER(i32p), ctypes.POINTER(i32p),
                                       ctypes.POINTER(f64p), ctypes.POINTER(MatrixOptions), vp, ctypes.POINTER(vp)]),
    "mf_batch_destroy": (ctypes.c_int, [vp]),
    "mf_batch_row_offsets": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int64)]),
    "mf_cg_solve_batched": (ctypes.c_int, [vp, vp, vp, ctypes.POINTER(SolveOptions), ctypes.c_int,
                                           ctypes.POINTER(SolveInfo), f64p, vp]),
    "mf_batch_set_groups": (ctypes.c_int, [vp, ctypes.c_int]),
    "mf_batch_autotune": (ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.c_int, i32p, i32p, f64p, vp]),
    "mf_batch_timeline": (ctypes.c_int, [vp, f64p, f64p, i32p]),
    "mf_phase_read": (ctypes.c_int, [ctypes.POINTER(ctypes.c_ulonglong), ctypes.c_int]),
    "mf_device_count": (ctypes.c_int, []),
    "mf_device_name": (ctypes.c_int, [ctypes.c_int, ctypes.c_char_p, ctypes.c_int]),
}


def use_library(name):
    """Load ``name`` (a library of this directory built from the same sources, e.g.
    ``libmf_phase.so`` with per-phase timers) instead of libmf.so; only before the first
    call into the library."""
    global LIB_PATH
    if _LIB is not None:
        raise RuntimeError(f"{LIB_PATH.name} is already loaded")
    LIB_PATH = PKG_DIR / name


def lib():
    """Load libmf.so (raises when it is missing: there is no CPU fallback)."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (make -C mille-feuille_amd)")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def exported_symbols():
    L = lib()
    return {name for name in _SIGS if getattr(L, name, None) is not None}


def check(status, where):
    if status != MF_SUCCESS:
        raise MfError(status, where)


def ptr(a, t):
    return a.ctypes.data_as(t)


def as_i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def as_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)
