"""BVE step with passive tracers (lpm_bve_*): the leaf faces of a mesh are
the vortex particles, the vertices move with their velocity.

Bars: FP64 direct-sum steps equal the oracle's RK4 (oracle.bve_steps) within
1e-12; with no tracers the step is the RK4 step of lpm_rk4_*, bit for bit;
the vertices of a Rossby-Haurwitz wave (solid body + degree-5 wavenumber-4
pattern, zeta_abs carried by the vertices) show the pattern rotating at the
Haurwitz phase speed nu = (R(R+3) alpha - 2 Omega) / ((R+1)(R+2)), R = 4.
"""
import numpy as np
import pytest

import lpmcxx_amd as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rh(x, y, z, alpha):
    return 2 * alpha * z + L.sphere_harmonic54(x, y, z)


def test_bve_matches_oracle(ctx, mesh):
    m = mesh(L.LPM_SEED_ICOS_TRI, 3)
    omega, alpha, dt = 2 * np.pi, 1.0, 0.01
    src, verts = (m.cx, m.cy, m.cz), (m.vx, m.vy, m.vz)
    za = _rh(*src, alpha) + 2 * omega * m.cz
    cfg = L.SolverConfig.default(method=L.LPM_METHOD_DIRECT, prec=L.LPM_F64)
    (fs, vs) = ctx.bve_steps(cfg, dt, omega, 2, src, za, m.area, verts)
    (fr, vr) = O.bve_steps(src, za, m.area, verts, dt, omega, 2)
    d = max(np.abs(np.stack(fs) - np.stack(fr)).max(), np.abs(np.stack(vs) - np.stack(vr)).max())
    assert d < 1e-12
    moved = np.abs(np.stack(vs) - np.stack(verts)).max()
    assert moved > 1e-3
    assert np.abs(np.sqrt(sum(a * a for a in vs)) - 1).max() < 1e-15


def test_bve_without_tracers_is_rk4(ctx, mesh):
    m = mesh(L.LPM_SEED_CUBED, 4)
    za = _rh(m.cx, m.cy, m.cz, 1.0) + 4 * np.pi * m.cz
    for cfg in (L.SolverConfig.default(method=L.LPM_METHOD_DIRECT, prec=L.LPM_F32C),
                L.SolverConfig.default(method=L.LPM_METHOD_TREE, prec=L.LPM_F64, order=8, theta=0.5, leaf_max=32)):
        (fs, _) = ctx.bve_steps(cfg, 0.01, 2 * np.pi, 3, (m.cx, m.cy, m.cz), za, m.area)
        r = ctx.rk4_steps(cfg, 0.01, 2 * np.pi, 3, m.cx, m.cy, m.cz, za, m.area)
        assert all(np.array_equal(a, b) for a, b in zip(fs, r))


@pytest.mark.parametrize("method", [L.LPM_METHOD_DIRECT, L.LPM_METHOD_TREE_BATCH])
def test_bve_vertices_follow_rossby_haurwitz(ctx, mesh, method):
    m = mesh(L.LPM_SEED_CUBED, 6)
    omega, alpha, dt, nsteps = 2 * np.pi, 1.0, 0.01, 10
    src, verts = (m.cx, m.cy, m.cz), (m.vx, m.vy, m.vz)
    za_f = _rh(*src, alpha) + 2 * omega * m.cz
    za_v = _rh(*verts, alpha) + 2 * omega * m.vz      # carried by the vertices
    cfg = L.SolverConfig.default(method=method, prec=L.LPM_F32C, order=10, theta=0.5, leaf_max=64)
    (_, (x, y, z)) = ctx.bve_steps(cfg, dt, omega, nsteps, src, za_f, m.area, verts)
    t = dt * nsteps
    nu = (28 * alpha - 2 * omega) / 30
    zeta_now = za_v - 2 * omega * z
    rot = lambda ang: (np.cos(ang) * x + np.sin(ang) * y, -np.sin(ang) * x + np.cos(ang) * y, z)
    exact = _rh(*rot(nu * t), alpha)
    e = np.linalg.norm(zeta_now - exact) / np.linalg.norm(exact)
    e0 = np.linalg.norm(zeta_now - _rh(x, y, z, alpha)) / np.linalg.norm(exact)
    e1 = np.linalg.norm(zeta_now - _rh(*rot(-nu * t), alpha)) / np.linalg.norm(exact)
    print(f"\nmethod {method}: vertex pattern vs propagated {e:.3e}, frozen {e0:.3e}, reversed {e1:.3e}")
    assert e < 0.02 and e0 > 5 * e and e1 > 5 * e
    assert np.abs(np.sqrt(x * x + y * y + z * z) - 1).max() < 1e-14


def test_bve_tree_follows_direct(ctx, mesh):
    m = mesh(L.LPM_SEED_CUBED, 5)
    src, verts = (m.cx, m.cy, m.cz), (m.vx, m.vy, m.vz)
    za = L.gaussian_vortex(m.cx, m.cy, m.cz, m.area) + 4 * np.pi * m.cz
    d = ctx.bve_steps(L.SolverConfig.default(method=L.LPM_METHOD_DIRECT), 0.01, 2 * np.pi, 5, src, za, m.area, verts)
    for cfg in (L.SolverConfig.default(method=L.LPM_METHOD_TREE, order=10, theta=0.4, leaf_max=64),
                L.SolverConfig.default(method=L.LPM_METHOD_TREE_BATCH, prec=L.LPM_F32C, order=10, theta=0.4,
                                       leaf_max=64)):
        r = ctx.bve_steps(cfg, 0.01, 2 * np.pi, 5, src, za, m.area, verts)
        for a, b in zip(r, d):
            assert np.abs(np.stack(a) - np.stack(b)).max() < 1e-7


@pytest.mark.parametrize("stereo", [False, True])
def test_bve_learned_follows_direct(ctx, mesh, stereo):
    """The learned FP32C treecode (a constant model: every pair passing
    theta_max = 0.4 at order 7), with the Cartesian and with the
    stereographic far field, follows the direct-sum BVE over 5 steps."""
    from lpmcxx_amd import learned as LN
    m = mesh(L.LPM_SEED_CUBED, 5)
    src, verts = (m.cx, m.cy, m.cz), (m.vx, m.vy, m.vz)
    za = L.gaussian_vortex(m.cx, m.cy, m.cz, m.area) + 4 * np.pi * m.cz
    d = ctx.bve_steps(L.SolverConfig.default(method=L.LPM_METHOD_DIRECT), 0.01, 2 * np.pi, 5, src, za, m.area, verts)
    u_ref = LN.u_ref_of((za - 4 * np.pi * m.cz) * m.area)
    try:
        ctx.set_option(L.LPM_OPT_STEREO, 1 if stereo else 0)
        LN.constant_model(ctx, 7)
        cfg = L.SolverConfig.default(method=L.LPM_METHOD_TREE_LEARNED, prec=L.LPM_F32C, order=7, theta=0.4,
                                     leaf_max=64, u_ref=u_ref)
        r = ctx.bve_steps(cfg, 0.01, 2 * np.pi, 5, src, za, m.area, verts)
        from lpmcxx_amd import tree as T
        assert (T.stats(ctx)["far_stereo"] > 0) == stereo
        err = max(np.abs(np.stack(a) - np.stack(b)).max() for a, b in zip(r, d))
        print(f"\nlearned BVE (stereographic {stereo}) vs direct after 5 steps: max position difference {err:.2e}")
        assert err < 3e-7  # order 7 at theta 0.4: ~1e-7 in both forms
    finally:
        ctx.set_option(L.LPM_OPT_STEREO, 0)


def test_bve_edge_cases(ctx, mesh):
    m = mesh(L.LPM_SEED_ICOS_TRI, 2)
    cfg = L.SolverConfig.default()
    verts = (m.vx, m.vy, m.vz)
    # no vortex particles: nothing moves (up to the renormalization to |x| = 1)
    (_, v) = ctx.bve_steps(cfg, 0.01, 2 * np.pi, 2, (m.cx[:0], m.cy[:0], m.cz[:0]), m.area[:0], m.area[:0], verts)
    assert np.abs(np.stack(v) - np.stack(verts)).max() < 1e-15
    # argument errors
    x = np.zeros(4)
    with pytest.raises(L.LpmError):
        ctx.check(ctx.lib.lpm_bve_steps(ctx.h, L.C.byref(cfg), 0.01, 1.0, 1, -1, 4, x.ctypes.data, x.ctypes.data,
                                        x.ctypes.data, x.ctypes.data, x.ctypes.data))
    with pytest.raises(L.LpmError):
        ctx.check(ctx.lib.lpm_bve_steps(ctx.h, L.C.byref(cfg), 0.01, 1.0, 1, 2, 2, x.ctypes.data, x.ctypes.data,
                                        x.ctypes.data, None, x.ctypes.data))
