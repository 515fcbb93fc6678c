"""RK4 time step (lpm_rk4_*): parity with the oracle's FP64 RK4, the
Rossby-Haurwitz phase speed on the rotating sphere, and the learned treecode's
step against the FP32C direct sum's under the solver's default root box.

For zeta_rel = 2 alpha z + 30 cos(4 lambda) z (z^2 - 1)^2 (solid body plus the
degree-5 wavenumber-4 Rossby-Haurwitz wave of src/LpmAnalyticFunctions.h:87-120)
the vorticity pattern of the barotropic vorticity equation rotates rigidly
with nu = (R(R+3) alpha - 2 Omega) / ((R+1)(R+2)), R = 4 (Haurwitz 1940).
Absolute vorticity zeta_abs = zeta_rel + 2 Omega z is carried by the particles.
"""
import numpy as np
import pytest

import lpmcxx_amd as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rh(x, y, z, alpha):
    return 2 * alpha * z + L.sphere_harmonic54(x, y, z)


def test_rk4_matches_oracle(ctx, mesh):
    m = mesh(L.LPM_SEED_ICOS_TRI, 4)
    omega, alpha, dt = 2 * np.pi, 1.0, 0.01
    za = _rh(m.cx, m.cy, m.cz, alpha) + 2 * omega * m.cz
    cfg = L.SolverConfig.default(method=L.LPM_METHOD_DIRECT, prec=L.LPM_F64)
    g = ctx.rk4_steps(cfg, dt, omega, 2, m.cx, m.cy, m.cz, za, m.area)
    r = O.rk4_steps(m.cx, m.cy, m.cz, za, m.area, dt, omega, 2)
    d = np.sqrt(sum((a - b) ** 2 for a, b in zip(g, r)))
    assert d.max() < 1e-12
    # the fused FP64 direct sum (LPM_F64F) to the same bar
    cfg = L.SolverConfig.default(method=L.LPM_METHOD_DIRECT, prec=L.LPM_F64F)
    gf = ctx.rk4_steps(cfg, dt, omega, 2, m.cx, m.cy, m.cz, za, m.area)
    assert np.sqrt(sum((a - b) ** 2 for a, b in zip(gf, r))).max() < 1e-12
    # LPM_F64F serves the direct sum and the per-target treecode only
    with pytest.raises(L.LpmError):
        ctx.rk4_steps(L.SolverConfig.default(method=L.LPM_METHOD_TREE_BATCH, prec=L.LPM_F64F), dt, omega, 1,
                      m.cx, m.cy, m.cz, za, m.area)
    assert np.abs(np.sqrt(g[0] ** 2 + g[1] ** 2 + g[2] ** 2) - 1).max() < 1e-15
    moved = np.sqrt((g[0] - m.cx) ** 2 + (g[1] - m.cy) ** 2 + (g[2] - m.cz) ** 2)
    assert moved.max() > 1e-3          # particles did move


@pytest.mark.parametrize("method", [L.LPM_METHOD_DIRECT, L.LPM_METHOD_TREE])
def test_rossby_haurwitz_phase_speed(ctx, mesh, method):
    m = mesh(L.LPM_SEED_CUBED, 6)
    omega, alpha, dt, nsteps = 2 * np.pi, 1.0, 0.01, 10
    za = _rh(m.cx, m.cy, m.cz, alpha) + 2 * omega * m.cz
    cfg = L.SolverConfig.default(method=method, prec=L.LPM_F32C, order=10, theta=0.5, leaf_max=64)
    x, y, z = ctx.rk4_steps(cfg, dt, omega, nsteps, m.cx, m.cy, m.cz, za, m.area)
    t = dt * nsteps
    nu = (28 * alpha - 2 * omega) / 30
    zeta_now = za - 2 * omega * z                     # relative vorticity carried to time t
    lam = np.arctan2(y, x)
    rot = lambda ang: (np.cos(ang) * x + np.sin(ang) * y, -np.sin(ang) * x + np.cos(ang) * y, z)
    exact = _rh(*rot(nu * t), alpha)                  # pattern rotated by nu t
    frozen = _rh(x, y, z, alpha)                      # pattern not moving
    wrong = _rh(*rot(-nu * t), alpha)                 # moving the wrong way
    e = np.linalg.norm(zeta_now - exact) / np.linalg.norm(exact)
    e0 = np.linalg.norm(zeta_now - frozen) / np.linalg.norm(exact)
    e1 = np.linalg.norm(zeta_now - wrong) / np.linalg.norm(exact)
    print(f"\nmethod {method}: rel err vs propagated pattern {e:.3e}, vs frozen {e0:.3e}, vs reversed {e1:.3e}")
    assert e < 0.02 and e0 > 5 * e and e1 > 5 * e
    assert np.abs(np.sqrt(x * x + y * y + z * z) - 1).max() < 1e-14


@pytest.mark.parametrize("stereo", [False, True])
def test_learned_rk4_step_under_the_default_root_box(ctx, mesh, stereo):
    """The order model is trained on the tree of the build's own root box
    (+-(maxR + 1e-5)); the RK4 step builds its trees in the solver's default
    box (+-1.00101), so its batches and nodes differ from the training ones.
    One learned step must stay within 1e-5 (relative L2 of the displacement)
    of the FP32C direct-sum step from the same start, cubed level 8, with the
    Cartesian and with the stereographic far field."""
    from lpmcxx_amd import learned as LN
    from lpmcxx_amd import tree as T
    m = mesh(L.LPM_SEED_CUBED, 8)
    omega, alpha, dt = 2 * np.pi, 1.0, 0.01
    zr = _rh(m.cx, m.cy, m.cz, alpha)
    g = zr * m.area
    u_ref = LN.u_ref_of(g)
    ctx.set_option(L.LPM_OPT_BATCH_CELL_TARGETS, 1024)
    try:
        src = (m.cx, m.cy, m.cz)
        T.build(ctx, src, m.max_radius, 128)
        T.moments(ctx, g, 12, 1)
        ls = LN.labels(ctx, src, 0.85, u_ref, batch_stride=4, keep_rate=0.5, frontier_levels=2, max_samples=1 << 21)
        # per-pair bounds with margin below 1e-5 at this level (the stereographic
        # series is more accurate at equal order, so its bound is larger)
        LN.fit(ctx, ls, 5e-7 if stereo else 2e-7, 7, epochs=8, stereo_extra=1 if stereo else None)
        ctx.set_option(L.LPM_OPT_STEREO, 1 if stereo else 0)
        ctx.set_option(L.LPM_OPT_STEREO_EXTRA_ORDER, 1)
        ctx.set_option(L.LPM_OPT_STEREO_RATIO, 1.5)
        za = zr + 2 * omega * m.cz
        cfg_l = L.SolverConfig.default(method=L.LPM_METHOD_TREE_LEARNED, prec=L.LPM_F32C, order=7, theta=0.85,
                                       leaf_max=128, u_ref=u_ref)
        assert list(cfg_l.root_box_max) != [m.max_radius + 1e-5] * 3
        cfg_d = L.SolverConfig.default(method=L.LPM_METHOD_DIRECT, prec=L.LPM_F32C)
        xl = np.stack(ctx.rk4_steps(cfg_l, dt, omega, 1, *src, za, m.area))
        xd = np.stack(ctx.rk4_steps(cfg_d, dt, omega, 1, *src, za, m.area))
        x0 = np.stack(src)
        d = np.linalg.norm(xl - xd, axis=0)
        disp = np.linalg.norm(xd - x0, axis=0)
        e = np.linalg.norm(d) / np.linalg.norm(disp)
        print(f"\nlearned RK4 step (stereographic {stereo}), default root box: displacement rel L2 err {e:.2e}, "
              f"max {d.max():.2e} (max displacement {disp.max():.2e})")
        assert e < 1e-5
        assert d.max() < 1e-4 * disp.max()
    finally:
        ctx.set_option(L.LPM_OPT_BATCH_CELL_TARGETS, 256)   # the context defaults
        ctx.set_option(L.LPM_OPT_STEREO, 0)
