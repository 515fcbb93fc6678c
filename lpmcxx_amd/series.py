"""Taylor series of the spherical Green's function (lpm_series_* in
include/lpm_gpu.h): Python access mirroring TaylorSeries3d /
SphereGreensSeries (src/LpmTaylorSeries3d.h:20-63).  Multi-indices
|k| <= max_p in lex order (the reference's std::map order)."""
from __future__ import annotations

import numpy as np

from . import Context, _as_host, _ptr


def nterms(max_p: int) -> int:
    return (max_p + 1) * (max_p + 2) * (max_p + 3) // 6


def multi_indices(max_p: int):
    """(k1, k2, k3), |k| <= max_p, lex order."""
    return [(a, b, c) for a in range(max_p + 1) for b in range(max_p + 1 - a) for c in range(max_p + 1 - a - b)]


def coeffs(ctx: Context, max_p: int, targets, centres) -> np.ndarray:
    """SphereGreensSeries::computeCoeffs for n (target, centre) pairs: [n][nterms]."""
    tx, ty, tz = _as_host({"tx": targets[0], "ty": targets[1], "tz": targets[2]})
    cx, cy, cz = _as_host({"cx": centres[0], "cy": centres[1], "cz": centres[2]}, tx.size)
    out = np.empty((tx.size, nterms(max_p)))
    ctx.check(ctx.lib.lpm_series_coeffs(ctx.h, max_p, tx.size, _ptr(tx), _ptr(ty), _ptr(tz), _ptr(cx), _ptr(cy),
                                        _ptr(cz), _ptr(out)))
    return out


def moments(ctx: Context, max_p: int, points, strength, centre, weights=None, weight_order: int = 0) -> np.ndarray:
    """TaylorSeries3d::computeMoments: sum_i (y_i - c)^k s_i [w_i] in index order."""
    x, y, z = _as_host({"x": points[0], "y": points[1], "z": points[2]})
    s, w = _as_host({"strength": strength, "weights": weights}, x.size)
    out = np.empty(nterms(max_p))
    ctx.check(ctx.lib.lpm_series_moments(ctx.h, max_p, x.size, _ptr(x), _ptr(y), _ptr(z), _ptr(s), _ptr(w),
                                         weight_order, float(centre[0]), float(centre[1]), float(centre[2]),
                                         _ptr(out)))
    return out
