"""GPU parity at the workloads' full sizes, checked against the oracle.

  * cubed-sphere level 8 (393,216 particles), single Gaussian vortex, leaf
    128, theta 0.7, p 8: the per-target FP64 treecode equals the oracle's
    traversal (1e-12, per-target macCounter exact) on 2,048 sampled targets of
    the mesh solve (vertices and faces); FP32 moments and near field within
    1e-5 of it;
  * cubed-sphere level 9 (1,572,864 particles), 8 Gaussian vortices (seed
    1234): the FP64 direct sum equals the oracle on 4,096 sampled targets
    (1e-12), and the learned-order FP32C treecode (labels, GPU training,
    evaluation) is within 1e-5 of the oracle there;
  * cubed-sphere level 10 (6,291,456 particles), solid body + Rossby-Haurwitz
    wave, Omega = 2 pi, dt = 0.01: the learned treecode within 1e-5 of the
    oracle at t = 0 on 4,096 sampled targets, then 10 learned RK4 steps keep
    |x| = 1 and move the vorticity pattern at the Rossby-Haurwitz phase speed;
  * targets clustered in a small cap in random order (the batch level stops
    at the sorted key levels): batch, learned and streamfunction treecodes
    cover every target, the batch one equal to the oracle's batch traversal;
  * the sphere precondition of LPM_F32C and of the series forms.
"""
import time

import numpy as np
import pytest

import lpmcxx_amd as L
from lpmcxx_amd import learned as LN
from lpmcxx_amd import tree as T
from oracle import oracle as O

pytestmark = pytest.mark.gpu

VEL, PSI = L.LPM_K_BIOT_SAVART_SPHERE, L.LPM_K_GREENS_SPHERE


def _stereo(ctx, on, extra=1, ratio=2.0):
    ctx.set_option(L.LPM_OPT_STEREO, 1 if on else 0)
    ctx.set_option(L.LPM_OPT_STEREO_EXTRA_ORDER, extra)
    ctx.set_option(L.LPM_OPT_STEREO_RATIO, ratio)


def _learned_model(ctx, src, g, u_ref, max_r, *, stride, tau, theta_max=0.92, max_order=7, leaf=128):
    """Labels from the FP64 treecode and GPU training of the order model
    (the benchmark's settings), then the tree and moments of the evaluation
    with the stereographic far field."""
    ctx.set_option(L.LPM_OPT_BATCH_CELL_TARGETS, 1024)
    _stereo(ctx, False)
    T.build(ctx, src, max_r, leaf)
    T.moments(ctx, g, 12, 1)
    ls = LN.labels(ctx, src, theta_max, u_ref, batch_stride=stride, keep_rate=0.3, frontier_levels=2,
                   max_samples=3 << 20)
    fr = LN.fit(ctx, ls, tau, max_order, epochs=8, batch=4096, lr=2e-3, stereo_extra=1)
    _stereo(ctx, True)
    T.build(ctx, src, max_r, leaf)
    T.moments(ctx, g, max_order, 1)
    return fr, len(ls.ids)


def _reset(ctx):
    ctx.set_option(L.LPM_OPT_BATCH_CELL_TARGETS, 256)
    _stereo(ctx, False, 1, 1.0)


def test_config3_treecode_l8(ctx):
    t0 = time.perf_counter()
    m = L.GpuMesh(ctx, L.LPM_SEED_CUBED, 8)
    assert m.n_leaves == 393216 and m.n_vertices == 393218
    g = L.gaussian_vortex(m.cx, m.cy, m.cz, m.area) * m.area
    src = (m.cx, m.cy, m.cz)
    tgt = m.targets()
    idx = np.sort(np.random.default_rng(3).choice(len(tgt[0]), 2048, replace=False))
    sub = tuple(t[idx] for t in tgt)
    T.build(ctx, src, m.max_radius, 128)
    T.moments(ctx, g, 8, 1)
    u, mac = T.evaluate(ctx, sub, 0.7, prec=L.LPM_F64, series=L.LPM_SERIES_CARTESIAN)
    u32, mac32 = T.evaluate(ctx, sub, 0.7, prec=L.LPM_F32C, series=L.LPM_SERIES_CARTESIAN)
    uf, macf = T.evaluate(ctx, sub, 0.7, prec=L.LPM_F64F, series=L.LPM_SERIES_CARTESIAN)
    R = O.tree_build(src, m.max_radius, 128)
    mom = O.tree_moments(src, g, 8, 1)
    ur, macr, mac_per = O.tree_eval(src, g, mom, 8, 0.7, sub, series=1)
    ud = O.bs_direct(sub, src, g, skip=m.face_self_skip()[idx])
    e64, e32, et = O.rel_l2(u, ur), O.rel_l2(u32, u), O.rel_l2(ur, ud)
    print(f"\ncubed L8 (nodes {len(R['level'])}): FP64 tree vs oracle tree rel-L2 {e64:.2e} max "
          f"{O.rel_max(u, ur):.2e}, mac {mac} (oracle {macr}); FP32 moments/near vs FP64 tree {e32:.2e}; "
          f"tree (theta 0.7, p 8) vs direct sum {et:.2e}; {time.perf_counter() - t0:.1f} s")
    assert mac == macr == int(mac_per.sum()) and mac32 == mac == macf
    assert e64 <= 1e-12 and O.rel_max(u, ur) <= 1e-12
    assert O.rel_l2(uf, ur) <= 1e-12 and O.rel_max(uf, ur) <= 1e-12
    assert e32 <= 1e-5
    assert et < 1e-3


def test_config4_learned_l9(ctx):
    t0 = time.perf_counter()
    m = L.GpuMesh(ctx, L.LPM_SEED_CUBED, 9)
    assert m.n_leaves == 1572864
    g = L.random_vortices(m.cx, m.cy, m.cz, m.area, n=8, seed=1234) * m.area
    src = (m.cx, m.cy, m.cz)
    u_ref = LN.u_ref_of(g)
    idx = np.sort(np.random.default_rng(7).choice(m.n_leaves, 4096, replace=False))
    sub = tuple(s[idx] for s in src)
    ref = O.bs_direct(sub, src, g)                    # the self pair contributes 0 (|den| <= 1e-13)
    u64 = ctx.direct_sum(sub, src, g, prec=L.LPM_F64)
    assert O.rel_l2(u64, ref) <= 1e-12 and O.rel_max(u64, ref) <= 1e-12
    try:
        fr, nlab = _learned_model(ctx, src, g, u_ref, m.max_radius, stride=16, tau=7e-7)
        u = LN.evaluate(ctx, src, 0.92, 0, 7, u_ref, prec=L.LPM_F32C)
        st = T.stats(ctx)
    finally:
        _reset(ctx)
    e = O.rel_l2(tuple(a[idx] for a in u), ref)
    print(f"\ncubed L9: FP64 direct vs oracle {O.rel_l2(u64, ref):.2e}; learned FP32C (labels {nlab}, accuracy "
          f"{fr.accuracy:.3f}) vs oracle on 4096 targets {e:.2e}; far {st['far']} (stereographic "
          f"{st['far_stereo']}) near {st['near']}; {time.perf_counter() - t0:.1f} s")
    assert e <= 1e-5
    assert st["far_stereo"] > 0 and st["near"] < 0.01 * m.n_leaves ** 2


def test_config5_rk4_l10(ctx):
    t0 = time.perf_counter()
    m = L.GpuMesh(ctx, L.LPM_SEED_CUBED, 10)
    n = m.n_leaves
    assert n == 6291456
    omega, alpha, dt, nsteps = 2 * np.pi, 1.0, 0.01, 10
    zrel = 2 * alpha * m.cz + L.sphere_harmonic54(m.cx, m.cy, m.cz)
    g = zrel * m.area
    za = zrel + 2 * omega * m.cz
    src = (m.cx, m.cy, m.cz)
    u_ref = LN.u_ref_of(g)
    idx = np.sort(np.random.default_rng(7).choice(n, 4096, replace=False))
    sub = tuple(s[idx] for s in src)
    ref = O.bs_direct(sub, src, g)
    try:
        _learned_model(ctx, src, g, u_ref, m.max_radius, stride=64, tau=5e-7)
        u = LN.evaluate(ctx, src, 0.92, 0, 7, u_ref, prec=L.LPM_F32C)
        e0 = O.rel_l2(tuple(a[idx] for a in u), ref)
        cfg = L.SolverConfig.default(method=L.LPM_METHOD_TREE_LEARNED, prec=L.LPM_F32C, order=7, theta=0.92,
                                     leaf_max=128, u_ref=u_ref)
        x, y, z = ctx.rk4_steps(cfg, dt, omega, nsteps, *src, za, m.area)
    finally:
        _reset(ctx)
    t = dt * nsteps
    nu = (28 * alpha - 2 * omega) / 30               # Rossby-Haurwitz phase speed, R = 4
    zeta_now = za - 2 * omega * z
    rh = lambda a, b, c: 2 * alpha * c + L.sphere_harmonic54(a, b, c)
    rot = lambda ang: (np.cos(ang) * x + np.sin(ang) * y, -np.sin(ang) * x + np.cos(ang) * y, z)
    exact = rh(*rot(nu * t))
    e = np.linalg.norm(zeta_now - exact) / np.linalg.norm(exact)
    e_frozen = np.linalg.norm(zeta_now - rh(x, y, z)) / np.linalg.norm(exact)
    e_rev = np.linalg.norm(zeta_now - rh(*rot(-nu * t))) / np.linalg.norm(exact)
    r = np.sqrt(x * x + y * y + z * z)
    print(f"\ncubed L10: learned FP32C vs oracle at t = 0 on 4096 targets {e0:.2e}; after {nsteps} RK4 steps "
          f"pattern vs propagated {e:.2e}, vs frozen {e_frozen:.2e}, vs reversed {e_rev:.2e}, "
          f"max ||x| - 1| {np.abs(r - 1).max():.1e}; {time.perf_counter() - t0:.1f} s")
    assert e0 <= 1e-5
    assert np.abs(r - 1).max() < 1e-14
    assert e < 2e-3 and e_frozen > 20 * e and e_rev > 20 * e


def test_clustered_random_order_targets(ctx, mesh):
    """200,000 targets in a 0.005-rad cap, random order.  Below the 12 sorted
    key levels neighbouring keys are in input order, so the batch level stops
    there; every target is evaluated (no batch beyond the launch bound)."""
    m = mesh(L.LPM_SEED_CUBED, 6)
    src = (m.cx, m.cy, m.cz)
    g = L.gaussian_vortex(m.cx, m.cy, m.cz, m.area) * m.area
    rng = np.random.default_rng(12)
    n = 200000
    th = 0.005 * np.sqrt(rng.uniform(size=n))
    ph = rng.uniform(0, 2 * np.pi, n)
    c0 = np.array([np.cos(np.radians(60.0)), 0.0, np.sin(np.radians(60.0))])
    e1 = np.cross(c0, [0.0, 1.0, 0.0]); e1 /= np.linalg.norm(e1)
    e2 = np.cross(c0, e1)
    P = (np.cos(th)[:, None] * c0 + (np.sin(th) * np.cos(ph))[:, None] * e1 + (np.sin(th) * np.sin(ph))[:, None] * e2)
    P /= np.linalg.norm(P, axis=1)[:, None]
    tgt = tuple(P.T.copy())
    T.build(ctx, src, m.max_radius, 64)
    T.moments(ctx, g, 8, 1)
    ub = np.stack(T.evaluate_batch(ctx, tgt, 0.5, 8))
    sb = T.stats(ctx)
    O.tree_build(src, m.max_radius, 64)
    ur, far, near = O.tree_eval_batch(src, g, O.tree_moments(src, g, 8, 1), 8, 0.5, tgt)
    ur = np.stack(ur)
    assert (sb["far"], sb["near"]) == (far, near)
    assert O.rel_max(ub, ur) <= 1e-12
    u_ref = LN.u_ref_of(g)
    LN.constant_model(ctx, 6)
    ul = np.stack(LN.evaluate(ctx, tgt, 0.5, 0, 6, u_ref, prec=L.LPM_F64))
    u6 = np.stack(T.evaluate_batch(ctx, tgt, 0.5, 6))
    assert O.rel_max(ul, u6) <= 1e-13
    T.moments_psi(ctx, g, 8)
    psi = T.evaluate_psi(ctx, tgt, 0.5, 8)
    pd = ctx.direct_sum(tgt, src, g, kernels=PSI, prec=L.LPM_F64)
    ep = np.abs(psi - pd).max() / np.abs(pd).max()
    ud = np.stack(ctx.direct_sum(tgt, src, g, prec=L.LPM_F64))
    eb = np.linalg.norm(ub - ud, axis=0).max() / np.linalg.norm(ud, axis=0).max()
    print(f"\nclustered targets: batch vs oracle batch {O.rel_max(ub, ur):.2e}, far {far} near {near}; batch vs "
          f"direct (max, per target) {eb:.2e}; streamfunction tree vs direct {ep:.2e}")
    assert eb < 1e-4 and ep < 1e-4


def test_sphere_precondition(ctx):
    """LPM_F32C and the series forms need |x| = R; off the sphere they return
    LPM_ERR_INVALID_ARG, while LPM_F64 with the reference forms computes the
    reference's general R^2 - x.y (checked against the oracle)."""
    rng = np.random.default_rng(9)
    P = rng.normal(size=(3, 600)); P /= np.linalg.norm(P, axis=0)
    on = tuple(P)
    off = tuple(P * (1.0 + 1e-6 * rng.uniform(0.5, 1.0, 600)))
    g = rng.normal(size=600)
    ref = O.bs_direct(off, off, g)
    u = ctx.direct_sum(off, off, g, prec=L.LPM_F64)
    assert O.rel_l2(u, ref) <= 1e-12
    with pytest.raises(L.LpmError) as e:
        ctx.direct_sum(off, off, g, prec=L.LPM_F32C)
    assert e.value.status == -1 and "|x| = R" in str(e.value)
    with pytest.raises(L.LpmError):
        ctx.direct_sum(on, off, g, prec=L.LPM_F32C)        # off-sphere sources only
    two = tuple(2.0 * a for a in on)                        # on the sphere of radius 2
    u2 = ctx.direct_sum(two, two, g, prec=L.LPM_F32C, radius=2.0)
    assert O.rel_l2(u2, O.bs_direct(two, two, g, radius=2.0)) <= 1e-5
    # treecode: off-sphere sources, the reference series in FP64 only
    maxr = float(np.sqrt((np.stack(off) ** 2).sum(0).max()))
    T.build(ctx, off, maxr, 16)
    T.moments(ctx, g, 4, 1)
    ut, mac = T.evaluate(ctx, off, 0.3, prec=L.LPM_F64, series=L.LPM_SERIES_REFERENCE)
    O.tree_build(off, maxr, 16)
    ur, macr, _ = O.tree_eval(off, g, O.tree_moments(off, g, 4, 1), 4, 0.3, off, series=0)
    assert mac == macr and O.rel_l2(ut, ur) <= 1e-12
    for call in (lambda: T.evaluate(ctx, off, 0.3, series=L.LPM_SERIES_CARTESIAN),
                 lambda: T.evaluate(ctx, off, 0.3, prec=L.LPM_F32C, series=L.LPM_SERIES_REFERENCE),
                 lambda: T.evaluate_batch(ctx, off, 0.3, 4)):
        with pytest.raises(L.LpmError) as e:
            call()
        assert e.value.status == -1
    # on-sphere sources, off-sphere targets
    T.build(ctx, on, 1.0, 16)
    T.moments(ctx, g, 4, 1)
    T.evaluate_batch(ctx, on, 0.3, 4)
    with pytest.raises(L.LpmError):
        T.evaluate_batch(ctx, off, 0.3, 4)
    # RK4 with an FP32C solver from off-sphere positions
    cfg = L.SolverConfig.default(method=L.LPM_METHOD_DIRECT, prec=L.LPM_F32C)
    with pytest.raises(L.LpmError):
        ctx.rk4_steps(cfg, 0.01, 0.0, 1, *off, g, np.ones(600))
    cfg64 = L.SolverConfig.default(method=L.LPM_METHOD_DIRECT, prec=L.LPM_F64)
    x, y, z = ctx.rk4_steps(cfg64, 0.01, 0.0, 1, *off, g, np.ones(600))
    assert np.abs(np.sqrt(x * x + y * y + z * z) - 1).max() < 1e-15
