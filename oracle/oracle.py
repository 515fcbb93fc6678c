"""CPU ORACLE loader -- test infrastructure only.

Wraps oracle/lib/liblpm_oracle.so (the plain-C FP64 restatement in
oracle/*.c).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB = Path(__file__).resolve().parent / "lib" / "liblpm_oracle.so"
_lib = None
_vp = C.c_void_p


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise ImportError(f"{LIB} is not built")
        L = C.CDLL(str(LIB))
        L.oracle_bs_direct.argtypes = [_vp, _vp, _vp, C.c_int64, _vp, _vp, _vp, _vp, C.c_int64, _vp,
                                       C.c_double, C.c_double, _vp, _vp, _vp]
        L.oracle_bs_direct.restype = None
        L.oracle_greens_direct.argtypes = [_vp, _vp, _vp, C.c_int64, _vp, _vp, _vp, _vp, C.c_int64, _vp,
                                           C.c_double, _vp]
        L.oracle_greens_direct.restype = None
        L.oracle_num_threads.restype = C.c_int
        _declare_tree(L)
        _lib = L
    return _lib


def _declare_tree(L: C.CDLL) -> None:
    if hasattr(L, "oracle_tree_build"):
        L.oracle_tree_build.argtypes = [_vp, _vp, _vp, C.c_int64, C.c_double, C.c_int, C.c_int]
        L.oracle_tree_build.restype = C.c_int64
        L.oracle_tree_get.argtypes = [_vp] * 9
        L.oracle_tree_get.restype = None
        L.oracle_tree_moments.argtypes = [_vp, _vp, _vp, _vp, C.c_int, C.c_int, _vp]
        L.oracle_tree_moments.restype = None
        L.oracle_tree_eval.argtypes = [_vp, _vp, _vp, _vp, _vp, C.c_int, C.c_double, C.c_double, C.c_double,
                                       _vp, _vp, _vp, C.c_int64, _vp, _vp, _vp, _vp, C.c_int64]
        L.oracle_tree_eval.restype = C.c_int64
        L.oracle_tree_free.argtypes = []
        L.oracle_tree_free.restype = None
        L.oracle_coeffs_lex.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_int, _vp]
        L.oracle_coeffs_lex.restype = C.c_int
        L.oracle_coeffs_series.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                           C.c_double, C.c_int, _vp]
        L.oracle_coeffs_series.restype = C.c_int


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def bs_direct(targets, sources, gamma, skip=None, radius=1.0, delta=0.0):
    """Reference-order FP64 velocity sum; returns (ux, uy, uz)."""
    tx, ty, tz = (_c(a) for a in targets)
    sx, sy, sz = (_c(a) for a in sources)
    g = _c(gamma)
    sk = None if skip is None else _c(skip, np.int64)
    ux, uy, uz = np.empty(tx.size), np.empty(tx.size), np.empty(tx.size)
    lib().oracle_bs_direct(tx.ctypes.data, ty.ctypes.data, tz.ctypes.data, tx.size, sx.ctypes.data,
                           sy.ctypes.data, sz.ctypes.data, g.ctypes.data, sx.size,
                           None if sk is None else sk.ctypes.data, radius, delta,
                           ux.ctypes.data, uy.ctypes.data, uz.ctypes.data)
    return ux, uy, uz


def greens_direct(targets, sources, gamma, skip=None, radius=1.0):
    """Reference-order FP64 streamfunction sum; returns psi."""
    tx, ty, tz = (_c(a) for a in targets)
    sx, sy, sz = (_c(a) for a in sources)
    g = _c(gamma)
    sk = None if skip is None else _c(skip, np.int64)
    psi = np.empty(tx.size)
    lib().oracle_greens_direct(tx.ctypes.data, ty.ctypes.data, tz.ctypes.data, tx.size, sx.ctypes.data,
                               sy.ctypes.data, sz.ctypes.data, g.ctypes.data, sx.size,
                               None if sk is None else sk.ctypes.data, radius, psi.ctypes.data)
    return psi


def num_threads() -> int:
    return lib().oracle_num_threads()


def rel_l2(a, b) -> float:
    """||a - b||_2 / ||b||_2 over stacked components."""
    a = np.concatenate([np.ravel(x) for x in (a if isinstance(a, (tuple, list)) else [a])])
    b = np.concatenate([np.ravel(x) for x in (b if isinstance(b, (tuple, list)) else [b])])
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def rel_max(a, b) -> float:
    """max_i |a_i - b_i| / max_i |b_i| (vector magnitudes for 3-tuples)."""
    if isinstance(a, (tuple, list)):
        da = np.sqrt(sum((x - y) ** 2 for x, y in zip(a, b)))
        nb = np.sqrt(sum(y ** 2 for y in b))
        return float(da.max() / nb.max())
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
