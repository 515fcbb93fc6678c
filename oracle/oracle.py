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
        L.oracle_tree_build_box.argtypes = [_vp, _vp, _vp, C.c_int64, _vp, _vp, C.c_int, C.c_int]
        L.oracle_tree_build_box.restype = C.c_int64
        L.oracle_tree_get.argtypes = [_vp] * 9
        L.oracle_tree_get.restype = None
        L.oracle_tree_moments.argtypes = [_vp, _vp, _vp, _vp, C.c_int, C.c_int, _vp]
        L.oracle_tree_moments.restype = None
        L.oracle_tree_eval.argtypes = [_vp, _vp, _vp, _vp, _vp, C.c_int, C.c_double, C.c_double, C.c_double,
                                       _vp, _vp, _vp, C.c_int64, _vp, _vp, _vp, _vp, C.c_int64]
        L.oracle_tree_eval.restype = C.c_int64
        L.oracle_tree_free.argtypes = []
        L.oracle_tree_free.restype = None
        L.oracle_node_far.argtypes = [C.c_double] * 6 + [_vp] * 4 + [C.c_int64, C.c_int, C.c_int, C.c_int,
                                                                     C.c_double, _vp]
        L.oracle_node_far.restype = None
        L.oracle_coeffs_lex.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_int, _vp]
        L.oracle_coeffs_lex.restype = C.c_int
        L.oracle_coeffs_series.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                           C.c_double, C.c_int, _vp]
        L.oracle_coeffs_series.restype = C.c_int
        L.oracle_set_target_levels.argtypes = [C.c_int]
        L.oracle_set_target_levels.restype = None
        L.oracle_target_order.argtypes = [_vp, _vp, _vp, C.c_int64, _vp]
        L.oracle_target_order.restype = None
        L.oracle_target_batches.argtypes = [_vp, _vp, _vp, C.c_int64, _vp, _vp]
        L.oracle_target_batches.restype = C.c_int64
        L.oracle_tree_eval_batch.argtypes = [_vp] * 5 + [C.c_int, C.c_double, C.c_double] + [_vp] * 3 + \
            [C.c_int64] + [_vp] * 4
        L.oracle_tree_eval_batch.restype = None
        L.oracle_pair_errors.argtypes = [_vp] * 5 + [C.c_int, C.c_double, C.c_double] + [_vp] * 3 + \
            [C.c_int64, C.c_int64, C.c_int64, _vp]
        L.oracle_pair_errors.restype = None


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


# ---------------------------------------------------------------------- treecode
_last_nn = [0]


def nterms(p):
    return p * (p + 1) * (p + 2) // 6


def tree_build(sources, max_r, leaf_max, box=None, control=0):
    """Reference recursive octree (module state); returns a dict of node arrays
    with root 0 and the children of every node consecutive.  box=(lo3, hi3)
    and control (0 = max per node, 1 = max depth) select the general build."""
    L = lib()
    sx, sy, sz = (_c(a) for a in sources)
    if box is None:
        nn = L.oracle_tree_build(sx.ctypes.data, sy.ctypes.data, sz.ctypes.data, sx.size, max_r, leaf_max, 0)
    else:
        lo, hi = _c(box[0]), _c(box[1])
        nn = L.oracle_tree_build_box(sx.ctypes.data, sy.ctypes.data, sz.ctypes.data, sx.size, lo.ctypes.data,
                                     hi.ctypes.data, control, leaf_max)
    if nn < 0:
        raise ValueError("oracle tree build failed")
    _last_nn[0] = nn
    out = {"box": np.empty((nn, 8)), "cent": np.empty((nn, 4))}
    for k in ("level", "parent", "first_kid", "n_kids", "begin", "end"):
        out[k] = np.empty(nn, np.int64)
    out["perm"] = np.empty(sx.size, np.int64)
    L.oracle_tree_get(out["box"].ctypes.data, out["cent"].ctypes.data, out["level"].ctypes.data,
                      out["parent"].ctypes.data, out["first_kid"].ctypes.data, out["n_kids"].ctypes.data,
                      out["begin"].ctypes.data, out["end"].ctypes.data, out["perm"].ctypes.data)
    out["box"] = out["box"][:, :6].copy()
    return out


def tree_moments(sources, gamma, p, moment_sign):
    """Moments of the last built oracle tree: [n_nodes, nterms(p), 3]."""
    sx, sy, sz = (_c(a) for a in sources)
    g = _c(gamma)
    out = np.empty((_last_nn[0], nterms(p), 3))
    lib().oracle_tree_moments(sx.ctypes.data, sy.ctypes.data, sz.ctypes.data, g.ctypes.data, p, moment_sign,
                              out.ctypes.data)
    return out


def tree_eval(sources, gamma, moments, p, theta, targets, delta=0.0, radius=1.0, series=0):
    """Reference traversal; series 0 = reference expansion, 1 = Cartesian
    expansion of 2/(|t-y|^2 + 2 delta^2).  Returns ((ux, uy, uz), mac_total, mac_per_target)."""
    L = lib()
    sx, sy, sz = (_c(a) for a in sources)
    g = _c(gamma)
    m = _c(moments)
    tx, ty, tz = (_c(a) for a in targets)
    n = tx.size
    ux, uy, uz = np.empty(n), np.empty(n), np.empty(n)
    mac = np.empty(n, np.int64)
    tot = L.oracle_tree_eval(sx.ctypes.data, sy.ctypes.data, sz.ctypes.data, g.ctypes.data, m.ctypes.data, p,
                             theta, delta, radius, tx.ctypes.data, ty.ctypes.data, tz.ctypes.data, n,
                             ux.ctypes.data, uy.ctypes.data, uz.ctypes.data, mac.ctypes.data, series)
    return (ux, uy, uz), int(tot), mac


def coeffs_lex(t, c, delta, p):
    """SumNode::calc_coeffs restated: lex-ordered coefficients for |k| < p."""
    out = np.empty(nterms(p))
    lib().oracle_coeffs_lex(t[0], t[1], t[2], c[0], c[1], c[2], delta, p, out.ctypes.data)
    return out


def coeffs_series(t, s, max_p):
    """SphereGreensSeries::computeCoeffs restated: lex-ordered for |k| <= max_p."""
    out = np.empty(nterms(max_p + 1))
    lib().oracle_coeffs_series(t[0], t[1], t[2], s[0], s[1], s[2], max_p, out.ctypes.data)
    return out


def multi_indices(p):
    """(k1, k2, k3) with |k| < p in lex order."""
    return [(a, b, c) for a in range(p) for b in range(p) for c in range(p) if a + b + c < p]


def node_far(c, t, sources, gamma, p, moment_sign, series, delta=0.0):
    """Unscaled far field of one box about centroid c (sum before -1/(4 pi R))."""
    sx, sy, sz = (_c(a) for a in sources)
    g = _c(gamma)
    out = np.empty(3)
    lib().oracle_node_far(c[0], c[1], c[2], t[0], t[1], t[2], sx.ctypes.data, sy.ctypes.data, sz.ctypes.data,
                          g.ctypes.data, sx.size, p, moment_sign, series, delta, out.ctypes.data)
    return out


def set_target_levels(levels: int) -> None:
    """Key levels that order the targets of the batch functions (12: general
    targets; the tree build's sorted levels, 10 unless the tree is deeper,
    for targets that are the tree's own sources).  Module state."""
    lib().oracle_set_target_levels(levels)


def target_order(targets):
    """Morton order (top sorted key levels, ties by index) of the targets in
    the last built oracle tree's root box."""
    tx, ty, tz = (_c(a) for a in targets)
    out = np.empty(tx.size, np.int64)
    lib().oracle_target_order(tx.ctypes.data, ty.ctypes.data, tz.ctypes.data, tx.size, out.ctypes.data)
    return out


def target_batches(targets):
    """(order, bstart): Morton order and the first sorted position of every
    target batch of the batch treecode (last built oracle tree's root box)."""
    tx, ty, tz = (_c(a) for a in targets)
    order = np.empty(tx.size, np.int64)
    bstart = np.empty(max(tx.size, 1), np.int64)
    nb = lib().oracle_target_batches(tx.ctypes.data, ty.ctypes.data, tz.ctypes.data, tx.size, order.ctypes.data,
                                     bstart.ctypes.data)
    return order, bstart[:nb].copy()


def tree_eval_batch(sources, gamma, moments, p, theta, targets, delta=0.0):
    """Batch-acceptance treecode on the last built oracle tree (lex moments of
    order p, sign +1).  Returns ((ux, uy, uz), far_pairs, near_pairs)."""
    sx, sy, sz = (_c(a) for a in sources)
    g, m = _c(gamma), _c(moments)
    tx, ty, tz = (_c(a) for a in targets)
    n = tx.size
    ux, uy, uz = np.empty(n), np.empty(n), np.empty(n)
    st = np.zeros(2, np.int64)
    lib().oracle_tree_eval_batch(sx.ctypes.data, sy.ctypes.data, sz.ctypes.data, g.ctypes.data, m.ctypes.data, p,
                                 theta, delta, tx.ctypes.data, ty.ctypes.data, tz.ctypes.data, n, ux.ctypes.data,
                                 uy.ctypes.data, uz.ctypes.data, st.ctypes.data)
    return (ux, uy, uz), int(st[0]), int(st[1])


def pair_errors(sources, gamma, moments, p, targets, batch, node, u_ref, delta=0.0):
    """Order-1..p far-field errors of one (batch, node) pair: max over the
    batch of |du| / u_ref (the label definition)."""
    sx, sy, sz = (_c(a) for a in sources)
    g, m = _c(gamma), _c(moments)
    tx, ty, tz = (_c(a) for a in targets)
    out = np.empty(p)
    lib().oracle_pair_errors(sx.ctypes.data, sy.ctypes.data, sz.ctypes.data, g.ctypes.data, m.ctypes.data, p,
                             delta, u_ref, tx.ctypes.data, ty.ctypes.data, tz.ctypes.data, tx.size, batch, node,
                             out.ctypes.data)
    return out


def bve_steps(src, zeta_abs, area, tracers, dt, omega, nsteps):
    """The RK4 of rk4_steps for vortex particles src = (x, y, z) plus passive
    tracers moved by the velocity the vortex particles induce (oracle direct
    sum, FP64).  Returns (src, tracers) positions."""
    ns = len(src[0])
    P = np.concatenate([np.stack(src), np.stack(tracers)], axis=1).astype(np.float64)

    def vel(p):
        g = (zeta_abs - 2.0 * omega * p[2, :ns]) * area
        return np.stack(bs_direct((p[0], p[1], p[2]), (p[0, :ns], p[1, :ns], p[2, :ns]), g))

    def nrm(p):
        return p / np.sqrt((p * p).sum(0))

    for _ in range(nsteps):
        k1 = vel(P)
        k2 = vel(nrm(P + 0.5 * dt * k1))
        k3 = vel(nrm(P + 0.5 * dt * k2))
        k4 = vel(nrm(P + dt * k3))
        P = nrm(P + dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4))
    return tuple(P[:, :ns]), tuple(P[:, ns:])


def rk4_steps(x, y, z, zeta_abs, area, dt, omega, nsteps):
    """FP64 RK4 of the Lagrangian vorticity equation with the oracle direct sum:
    gamma = (zeta_abs - 2 omega z) area per stage, stage and final positions
    renormalized to the unit sphere.  Returns (x, y, z)."""
    x, y, z = (np.array(a, dtype=np.float64) for a in (x, y, z))

    def vel(px, py, pz):
        g = (zeta_abs - 2.0 * omega * pz) * area
        return np.stack(bs_direct((px, py, pz), (px, py, pz), g))

    def nrm(p):
        return p / np.sqrt((p * p).sum(0))

    for _ in range(nsteps):
        x0 = np.stack([x, y, z])
        k1 = vel(*x0)
        k2 = vel(*nrm(x0 + 0.5 * dt * k1))
        k3 = vel(*nrm(x0 + 0.5 * dt * k2))
        k4 = vel(*nrm(x0 + dt * k3))
        x, y, z = nrm(x0 + dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4))
    return x, y, z
