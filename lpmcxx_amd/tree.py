"""Treecode entry points (lpm_tree_* in include/lpm_gpu.h) for Python callers.

Mirrors TreeSum's mesh path (src/LpmTreeSum.cpp:221-238 constructor,
:182-197 computeMeshMoments, :55-86 meshSolve):
    build(ctx, sources, max_radius, leaf_max)     TreeSum(p, leaf_max, mesh, ...)
    moments(ctx, gamma, order, moment_sign)       computeMeshMoments()
    evaluate(ctx, targets, theta, ...)            meshSolve(..., mp_tol=theta, macCounter)
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import LPM_F64, LPM_SERIES_CARTESIAN, Context, _as_host, _dev, _numel, _ptr


@dataclass
class TreeResult:
    velocity: tuple
    mac_count: int
    n_nodes: int
    depth: int


def _sources(ctx: Context) -> int:
    """Source count of the context's tree (set by the build wrappers)."""
    ns = getattr(ctx, "tree_sources", None)
    if ns is None:
        raise ValueError("no tree: build it first")
    return ns


def build(ctx: Context, sources, max_radius: float, leaf_max: int):
    sx, sy, sz = _as_host({"sx": sources[0], "sy": sources[1], "sz": sources[2]})
    nn, depth = C.c_int64(), C.c_int()
    ctx.check(ctx.lib.lpm_tree_build(ctx.h, _ptr(sx), _ptr(sy), _ptr(sz), sx.size, max_radius, leaf_max,
                                     C.byref(nn), C.byref(depth)))
    ctx.tree_sources = sx.size
    return nn.value, depth.value


def build_box(ctx: Context, sources, box_min, box_max, control: int, param: int):
    """General root box and depth control (control 0: max per node, 1: max depth)."""
    sx, sy, sz = _as_host({"sx": sources[0], "sy": sources[1], "sz": sources[2]})
    lo, hi = _as_host({"box_min": box_min, "box_max": box_max}, 3)
    nn, depth = C.c_int64(), C.c_int()
    ctx.check(ctx.lib.lpm_tree_build_box(ctx.h, _ptr(sx), _ptr(sy), _ptr(sz), sx.size, _ptr(lo), _ptr(hi), control,
                                         param, C.byref(nn), C.byref(depth)))
    ctx.tree_sources = sx.size
    return nn.value, depth.value


def build_dev(ctx: Context, sx, sy, sz, max_radius: float, leaf_max: int):
    nn, depth = C.c_int64(), C.c_int()
    _dev(ctx, {"sx": sx, "sy": sy, "sz": sz}, _numel(sx, "sx"))
    ctx.after_torch()
    ctx.check(ctx.lib.lpm_tree_build_dev(ctx.h, _ptr(sx), _ptr(sy), _ptr(sz), sx.numel(), max_radius, leaf_max,
                                         C.byref(nn), C.byref(depth)))
    ctx.tree_sources = sx.numel()
    return nn.value, depth.value


def moments(ctx: Context, gamma, order: int, moment_sign: int = 1) -> None:
    (g,) = _as_host({"gamma": gamma}, _sources(ctx))
    ctx.check(ctx.lib.lpm_tree_moments(ctx.h, _ptr(g), order, moment_sign))


def moments_dev(ctx: Context, gamma, order: int, moment_sign: int = 1) -> None:
    _dev(ctx, {"gamma": gamma}, _sources(ctx))
    ctx.after_torch()
    ctx.check(ctx.lib.lpm_tree_moments_dev(ctx.h, _ptr(gamma), order, moment_sign))


def evaluate(ctx: Context, targets, theta: float, delta: float = 0.0, prec: int = LPM_F64,
             series: int = LPM_SERIES_CARTESIAN):
    tx, ty, tz = _as_host({"tx": targets[0], "ty": targets[1], "tz": targets[2]})
    n = tx.size
    ux, uy, uz = np.zeros(n), np.zeros(n), np.zeros(n)
    mac = C.c_int64()
    ctx.check(ctx.lib.lpm_tree_eval(ctx.h, prec, series, theta, delta, _ptr(tx), _ptr(ty), _ptr(tz), n,
                                    _ptr(ux), _ptr(uy), _ptr(uz), C.byref(mac)))
    return (ux, uy, uz), mac.value


def evaluate_dev(ctx: Context, tx, ty, tz, out, theta: float, delta: float = 0.0, prec: int = LPM_F64,
                 series: int = LPM_SERIES_CARTESIAN, want_mac: bool = False):
    mac = C.c_int64()
    _dev(ctx, {"tx": tx, "ty": ty, "tz": tz, "ux": out[0], "uy": out[1], "uz": out[2]}, _numel(tx, "tx"))
    ctx.after_torch()
    ctx.check(ctx.lib.lpm_tree_eval_dev(ctx.h, prec, series, theta, delta, _ptr(tx), _ptr(ty), _ptr(tz), tx.numel(),
                                        _ptr(out[0]), _ptr(out[1]), _ptr(out[2]),
                                        C.byref(mac) if want_mac else None))
    return mac.value


def last_stats(ctx: Context):
    """(far pairs, near pairs) of the last evaluation."""
    f, n = C.c_int64(), C.c_int64()
    ctx.check(ctx.lib.lpm_tree_last_stats(ctx.h, C.byref(f), C.byref(n)))
    return f.value, n.value


LPM_STAT_FAR, LPM_STAT_NEAR, LPM_STAT_TERMS, LPM_STAT_MODEL = 0, 1, 2, 3
LPM_STAT_FAR_HIGH, LPM_STAT_TERMS_HIGH, LPM_STAT_FAR_STEREO, LPM_STAT_TERMS_STEREO = 4, 5, 6, 7
LPM_TREE_NSTATS = 8


def stats(ctx: Context) -> dict:
    """Counters of the last evaluation: far pairs, near pairs, far-field series
    terms, order-model calls, and (learned evaluation) the far pairs and terms
    handled by the order 8..12 list kernel and the far pairs evaluated as
    stereographic series with their complex terms."""
    out = np.zeros(LPM_TREE_NSTATS, np.int64)
    ctx.check(ctx.lib.lpm_tree_stats(ctx.h, _ptr(out)))
    return {"far": int(out[0]), "near": int(out[1]), "terms": int(out[2]), "model": int(out[3]),
            "far_high": int(out[4]), "terms_high": int(out[5]), "far_stereo": int(out[6]),
            "terms_stereo": int(out[7])}


def evaluate_batch(ctx: Context, targets, theta: float, order: int, delta: float = 0.0, prec: int = LPM_F64):
    """Batch-acceptance treecode (lpm_tree_eval_batch): one decision per 64
    Morton-sorted targets, Cartesian series of the given order."""
    tx, ty, tz = _as_host({"tx": targets[0], "ty": targets[1], "tz": targets[2]})
    n = tx.size
    ux, uy, uz = np.zeros(n), np.zeros(n), np.zeros(n)
    ctx.check(ctx.lib.lpm_tree_eval_batch(ctx.h, prec, order, theta, delta, _ptr(tx), _ptr(ty), _ptr(tz), n,
                                          _ptr(ux), _ptr(uy), _ptr(uz)))
    return ux, uy, uz


def evaluate_batch_dev(ctx: Context, tx, ty, tz, out, theta: float, order: int, delta: float = 0.0,
                       prec: int = LPM_F64) -> None:
    _dev(ctx, {"tx": tx, "ty": ty, "tz": tz, "ux": out[0], "uy": out[1], "uz": out[2]}, _numel(tx, "tx"))
    ctx.after_torch()
    ctx.check(ctx.lib.lpm_tree_eval_batch_dev(ctx.h, prec, order, theta, delta, _ptr(tx), _ptr(ty), _ptr(tz),
                                              tx.numel(), _ptr(out[0]), _ptr(out[1]), _ptr(out[2])))


def moments_psi(ctx: Context, gamma, order: int) -> None:
    """Streamfunction moments sum_j gamma_j (y_j - c)^k of the current tree."""
    (g,) = _as_host({"gamma": gamma}, _sources(ctx))
    ctx.check(ctx.lib.lpm_tree_moments_psi(ctx.h, _ptr(g), order))


def moments_psi_dev(ctx: Context, gamma, order: int) -> None:
    _dev(ctx, {"gamma": gamma}, _sources(ctx))
    ctx.after_torch()
    ctx.check(ctx.lib.lpm_tree_moments_psi_dev(ctx.h, _ptr(gamma), order))


def evaluate_psi(ctx: Context, targets, theta: float, order: int, prec: int = LPM_F64) -> np.ndarray:
    """Streamfunction treecode (lpm_tree_eval_psi), batch acceptance, log
    series of the given order."""
    tx, ty, tz = _as_host({"tx": targets[0], "ty": targets[1], "tz": targets[2]})
    psi = np.zeros(tx.size)
    ctx.check(ctx.lib.lpm_tree_eval_psi(ctx.h, prec, order, theta, _ptr(tx), _ptr(ty), _ptr(tz), tx.size, _ptr(psi)))
    return psi


def evaluate_psi_dev(ctx: Context, tx, ty, tz, psi, theta: float, order: int, prec: int = LPM_F64) -> None:
    _dev(ctx, {"tx": tx, "ty": ty, "tz": tz, "psi": psi}, _numel(tx, "tx"))
    ctx.after_torch()
    ctx.check(ctx.lib.lpm_tree_eval_psi_dev(ctx.h, prec, order, theta, _ptr(tx), _ptr(ty), _ptr(tz), tx.numel(),
                                            _ptr(psi)))


def info(ctx: Context):
    nn, depth, nl = C.c_int64(), C.c_int(), C.c_int64()
    ctx.check(ctx.lib.lpm_tree_info(ctx.h, C.byref(nn), C.byref(depth), C.byref(nl)))
    return nn.value, depth.value, nl.value


def export(ctx: Context, n_sources: int) -> dict:
    nn, _, _ = info(ctx)
    out = {
        "box": np.empty((nn, 6)), "cent": np.empty((nn, 4)), "level": np.empty(nn, np.int32),
        "parent": np.empty(nn, np.int32), "first_kid": np.empty(nn, np.int32), "n_kids": np.empty(nn, np.int32),
        "range": np.empty((nn, 2), np.int32), "perm": np.empty(n_sources, np.int32),
    }
    ctx.check(ctx.lib.lpm_tree_export(ctx.h, _ptr(out["box"]), _ptr(out["cent"]), _ptr(out["level"]),
                                      _ptr(out["parent"]), _ptr(out["first_kid"]), _ptr(out["n_kids"]),
                                      _ptr(out["range"]), _ptr(out["perm"])))
    return out


def export_moments(ctx: Context, order: int) -> np.ndarray:
    nn, _, _ = info(ctx)
    out = np.empty((nn, order * (order + 1) * (order + 2) // 6, 3))
    ctx.check(ctx.lib.lpm_tree_export_moments(ctx.h, _ptr(out)))
    return out


def dfs_order(first_kid, n_kids) -> np.ndarray:
    """Node ids in depth-first pre-order with children in octant order (the
    reference's traversal order, src/LpmTreeSum.cpp:132-134)."""
    out, stack = [], [0]
    while stack:
        i = stack.pop()
        out.append(i)
        k, f = int(n_kids[i]), int(first_kid[i])
        stack.extend(range(f + k - 1, f - 1, -1))
    return np.array(out, dtype=np.int64)


def tree_velocity(ctx: Context, sources, gamma, targets, *, max_radius: float, leaf_max: int, order: int,
                  theta: float, moment_sign: int = 1, prec: int = LPM_F64, delta: float = 0.0,
                  series: int = LPM_SERIES_CARTESIAN) -> TreeResult:
    nn, depth = build(ctx, sources, max_radius, leaf_max)
    moments(ctx, gamma, order, moment_sign)
    u, mac = evaluate(ctx, targets, theta, delta, prec, series)
    return TreeResult(u, mac, nn, depth)
