"""GPU Taylor series (lpm_series_coeffs / lpm_series_moments) against the
oracle's restatement of SphereGreensSeries::computeCoeffs
(src/LpmTaylorSeries3d.cpp:63-117; itself pinned to the closed form
|k|!/k! t^k / D^(|k|+1)) and against host sums of the moments
(src/LpmTaylorSeries3d.cpp:119-154).

Bars: coefficients bit for bit (same FP64 operations in the same order);
moments within 1e-14 of the sum in index order (the power function may
round differently by an ulp); the series sum converging to 1/(1 - t.y)."""
from math import factorial

import numpy as np
import pytest

from lpmcxx_amd import series as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("max_p", [0, 1, 4, 8, 12, 20])
def test_coeffs_bit_exact(ctx, max_p):
    rng = np.random.default_rng(max_p)
    n = 257
    t = rng.normal(size=(3, n)); t /= np.linalg.norm(t, axis=0)
    c = rng.normal(size=(3, n)) * 0.35
    a = S.coeffs(ctx, max_p, tuple(t), tuple(c))
    for i in range(0, n, 16):
        ref = O.coeffs_series(t[:, i], c[:, i], max_p)
        assert np.array_equal(a[i], ref), (max_p, i)
    # closed form on one pair
    D = 1 - t[:, 0] @ c[:, 0]
    cf = [factorial(p + q + r) / (factorial(p) * factorial(q) * factorial(r)) * t[0, 0] ** p * t[1, 0] ** q *
          t[2, 0] ** r / D ** (p + q + r + 1) for p, q, r in S.multi_indices(max_p)]
    np.testing.assert_allclose(a[0], cf, rtol=1e-12, atol=0)


@pytest.mark.parametrize("weight_order", [None, 0, 1])
def test_moments_index_order_sums(ctx, weight_order):
    rng = np.random.default_rng(3)
    n, P = 1000, 9
    y = rng.uniform(-1, 1, size=(3, n))
    s = rng.normal(size=n)
    w = rng.uniform(0.5, 1.5, size=n)
    c = np.array([0.1, -0.2, 0.05])
    m = S.moments(ctx, P, tuple(y), s, c, None if weight_order is None else w, weight_order or 0)
    h = y - c[:, None]
    for t, (a, b, q) in enumerate(S.multi_indices(P)):
        vp = np.power(h[0], a) * np.power(h[1], b) * np.power(h[2], q)
        if weight_order is None:
            terms = vp * s
        elif weight_order == 0:
            terms = vp * (s * w)
        else:
            terms = vp * s * w
        ref = np.cumsum(terms)[-1]          # sequential, index order
        assert abs(m[t] - ref) <= 1e-14 * np.abs(terms).sum()
    # one point: the moment is s h^k
    m1 = S.moments(ctx, 4, tuple(y[:, :1]), s[:1], c)
    for t, (a, b, q) in enumerate(S.multi_indices(4)):
        assert m1[t] == pytest.approx(s[0] * h[0, 0] ** a * h[1, 0] ** b * h[2, 0] ** q, rel=1e-15, abs=1e-300)


def test_series_sum_converges(ctx):
    """sum_k a_k m_k -> sum_j s_j / (1 - t.y_j) for sources near the centre."""
    rng = np.random.default_rng(8)
    t = np.array([0.6, 0.0, 0.8])
    c = np.array([0.0, 0.3, -0.2])
    y = c[:, None] + 0.05 * rng.uniform(-1, 1, size=(3, 40))
    s = rng.normal(size=40)
    exact = np.sum(s / (1 - t @ y))
    errs = []
    for P in (1, 3, 5, 7):
        a = S.coeffs(ctx, P, tuple(t[:, None]), tuple(c[:, None]))[0]
        m = S.moments(ctx, P, tuple(y), s, c)
        errs.append(abs(a @ m - exact) / abs(exact))
    assert all(b < 0.2 * a for a, b in zip(errs, errs[1:])) and errs[-1] < 1e-8, errs


def test_series_errors(ctx):
    import lpmcxx_amd as L
    with pytest.raises(L.LpmError):
        S.coeffs(ctx, 21, ([1.0], [0.0], [0.0]), ([0.0], [0.0], [0.0]))
    with pytest.raises(L.LpmError):
        S.moments(ctx, 3, ([0.0], [0.0], [0.0]), [1.0], (0, 0, 0), [1.0], 2)
