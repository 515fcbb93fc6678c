"""Streamfunction treecode (lpm_tree_moments_psi / lpm_tree_eval_psi): the
scalar tree sum the reference declares but never implements.

Bars: theta = 0 (every leaf near) is the FP64 direct sum with the face self
pair skipped, within 1e-12; the error against the direct sum falls with the
order at theta = 0.5 down to 1e-7 at p = 12 (the rate of the velocity
series; a wrong series coefficient stalls it); FP32C within 1e-6 of FP64;
psi = zeta / 30 for the degree-5 wavenumber-4 pattern to the same
discretization error as the direct sum.
"""
import numpy as np
import pytest

import lpmcxx_amd as L
from lpmcxx_amd import tree as T
from oracle import oracle as O

pytestmark = pytest.mark.gpu
PSI = L.LPM_K_GREENS_SPHERE


def _setup(ctx, mesh, level, leaf, order, seed=L.LPM_SEED_CUBED):
    m = mesh(seed, level)
    src = (m.cx, m.cy, m.cz)
    g = L.sphere_harmonic54(*src) * m.area
    T.build(ctx, src, m.max_radius, leaf)
    T.moments_psi(ctx, g, order)
    return m, src, g


def test_psi_theta_zero_is_direct_sum(ctx, mesh):
    m, src, g = _setup(ctx, mesh, 4, 32, 4)
    tgt = m.targets()
    ref = ctx.direct_sum(tgt, src, g, kernels=PSI, skip=m.face_self_skip())
    for prec, tol in ((L.LPM_F64, 1e-12), (L.LPM_F32C, 1e-6)):
        psi = T.evaluate_psi(ctx, tgt, 0.0, 4, prec=prec)
        assert T.stats(ctx)["far"] == 0
        assert O.rel_max(psi, ref) < tol


def test_psi_converges_with_order(ctx, mesh):
    m, src, g = _setup(ctx, mesh, 6, 64, 12)
    tgt = m.targets()
    ref = ctx.direct_sum(tgt, src, g, kernels=PSI, skip=m.face_self_skip())
    errs = [O.rel_l2(T.evaluate_psi(ctx, tgt, 0.5, p), ref) for p in (2, 4, 6, 8, 10, 12)]
    print("\npsi treecode rel-L2 by order:", " ".join(f"{e:.2e}" for e in errs))
    assert all(b < a for a, b in zip(errs, errs[1:])), errs
    assert errs[-1] < 1e-7
    assert T.stats(ctx)["far"] > 0
    # FP32C: same traversal, FP32 series and near field
    e32 = O.rel_l2(T.evaluate_psi(ctx, tgt, 0.5, 8, prec=L.LPM_F32C), T.evaluate_psi(ctx, tgt, 0.5, 8))
    assert e32 < 1e-6


def test_psi_rossby_haurwitz(ctx, mesh):
    m, src, g = _setup(ctx, mesh, 6, 64, 10)
    tgt = (m.cx, m.cy, m.cz)
    exact = L.sphere_harmonic54(*tgt) / 30.0
    direct = ctx.direct_sum(tgt, src, g, kernels=PSI, skip=np.arange(len(g)))
    tree = T.evaluate_psi(ctx, tgt, 0.5, 10)
    ed, et = O.rel_max(direct, exact), O.rel_max(tree, exact)
    print(f"\npsi vs zeta/30: direct {ed:.3e}, treecode {et:.3e}")
    assert et < 1.01 * ed + 1e-8 and O.rel_l2(tree, direct) < 1e-6


def test_psi_errors(ctx, mesh):
    m, src, g = _setup(ctx, mesh, 3, 32, 4)
    with pytest.raises(L.LpmError):            # order above the streamfunction moments' order
        T.evaluate_psi(ctx, src, 0.5, 6)
    T.build(ctx, src, m.max_radius, 32)        # a new tree has no streamfunction moments
    with pytest.raises(L.LpmError):
        T.evaluate_psi(ctx, src, 0.5, 2)
