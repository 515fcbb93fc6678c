"""Pins for the oracle's direct sums (oracle/lpm_oracle.c).

Each check is chosen so that a plausible mistake -- a dropped term, a wrong
sign, a transposed cross product, a missing 1/(4 pi R), a self pair left in --
fails at least one of them:
  * closed forms on hand-picked pairs (cot(theta/2) law of the point vortex,
    log law of the Green's function);
  * an independent long-double brute force on a small random set;
  * invariants: tangency u.x = 0, linearity in gamma, rotation equivariance;
  * the analytic Rossby-Haurwitz velocity and psi = zeta/30 for zeta = Y_5^4
    (src/LpmAnalyticFunctions.h:87-120, tests/LpmBiotSavartRHWaveTests.cpp:86-108,
    tests/LpmPoissonSolverTests.cpp:290-301) with first-order convergence.
"""
import numpy as np
import pytest

import lpmcxx_amd as L
from oracle import oracle as O


def _rand_sphere(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    return v[:, 0].copy(), v[:, 1].copy(), v[:, 2].copy()


def test_point_vortex_closed_form():
    # target at the north pole, source at colatitude th in the xz-plane:
    # x*y = (0, sin th, 0), 1 - x.y = 1 - cos th  ->  u_y = -G cot(th/2) / (4 pi)
    for th in [0.1, 0.7, 2.0, 3.0]:
        u = O.bs_direct(([0.0], [0.0], [1.0]), ([np.sin(th)], [0.0], [np.cos(th)]), [2.5])
        assert u[0][0] == 0.0 and u[2][0] == 0.0
        assert u[1][0] == pytest.approx(-2.5 / np.tan(th / 2) / (4 * np.pi), rel=1e-14)
    # orthogonal unit pair: k(x, y) = x*y / 1
    u = O.bs_direct(([1.0], [0.0], [0.0]), ([0.0], [1.0], [0.0]), [1.0])
    assert (u[0][0], u[1][0]) == (0.0, 0.0) and u[2][0] == pytest.approx(-1 / (4 * np.pi), rel=1e-15)


def test_greens_closed_form_and_radius():
    th = 1.3
    psi = O.greens_direct(([0.0], [0.0], [1.0]), ([np.sin(th)], [0.0], [np.cos(th)]), [3.0])
    assert psi[0] == pytest.approx(-3.0 * np.log(1 - np.cos(th)) / (4 * np.pi), rel=1e-14)
    R = 2.0
    psi = O.greens_direct(([0.0], [0.0], [R]), ([R * np.sin(th)], [0.0], [R * np.cos(th)]), [1.0], radius=R)
    assert psi[0] == pytest.approx(-np.log(R * R * (1 - np.cos(th))) / (4 * np.pi * R), rel=1e-14)


def test_zero_tol_and_self_skip():
    x, y, z = _rand_sphere(5, 1)
    g = np.arange(1.0, 6.0)
    # coincident pair (x == y): |1 - x.x| <= 1e-13 -> exactly zero contribution
    u_all = O.bs_direct((x, y, z), (x, y, z), g)
    u_skip = O.bs_direct((x, y, z), (x, y, z), g, skip=np.arange(5))
    for a, b in zip(u_all, u_skip):
        assert np.array_equal(a, b)
    # the potential skips exactly the listed index
    sk = np.array([0, -1, 2, 3, 4])
    p = O.greens_direct((x[:1], y[:1], z[:1]), (x, y, z), g, skip=sk[:1])
    ref = sum(-np.log(1 - (x[0] * x[j] + y[0] * y[j] + z[0] * z[j])) / (4 * np.pi) * g[j] for j in range(1, 5))
    assert p[0] == pytest.approx(ref, rel=1e-13)


def test_long_double_brute_force():
    tx, ty, tz = _rand_sphere(64, 2)
    sx, sy, sz = _rand_sphere(200, 3)
    g = np.random.default_rng(4).normal(size=200)
    ux, uy, uz = O.bs_direct((tx, ty, tz), (sx, sy, sz), g, delta=0.05)
    psi = O.greens_direct((tx, ty, tz), (sx, sy, sz), g)
    ld = np.longdouble
    T = np.stack([tx, ty, tz], 1).astype(ld)
    S = np.stack([sx, sy, sz], 1).astype(ld)
    G = g.astype(ld)
    cr = np.cross(T[:, None, :], S[None, :, :])
    den = 1 - T @ S.T + ld(0.05) ** 2
    U = -(cr * (G[None, :] / den)[:, :, None]).sum(1) / (4 * ld(np.pi))
    P = -(np.log(1 - T @ S.T) * G[None, :]).sum(1) / (4 * ld(np.pi))
    for a, b in zip((ux, uy, uz), U.T):
        assert np.max(np.abs(a - b.astype(float))) < 1e-13 * np.max(np.abs(U))
    assert np.max(np.abs(psi - P.astype(float))) < 1e-13 * np.max(np.abs(P))


def test_invariants():
    tx, ty, tz = _rand_sphere(50, 5)
    sx, sy, sz = _rand_sphere(80, 6)
    g = np.random.default_rng(7).normal(size=80)
    u = O.bs_direct((tx, ty, tz), (sx, sy, sz), g)
    # tangent to the sphere
    assert np.max(np.abs(u[0] * tx + u[1] * ty + u[2] * tz)) < 1e-13
    # linear in gamma
    u2 = O.bs_direct((tx, ty, tz), (sx, sy, sz), -3.0 * g)
    for a, b in zip(u, u2):
        assert np.allclose(b, -3.0 * a, rtol=1e-13, atol=1e-15)
    # rotation equivariant
    q, _ = np.linalg.qr(np.random.default_rng(8).normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] *= -1
    T = np.stack([tx, ty, tz], 1) @ q.T
    S = np.stack([sx, sy, sz], 1) @ q.T
    ur = O.bs_direct(tuple(T.T), tuple(S.T), g)
    assert np.allclose(np.stack(ur, 1), np.stack(u, 1) @ q.T, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("seed", [L.LPM_SEED_ICOS_TRI, L.LPM_SEED_CUBED])
def test_rossby_haurwitz_convergence(mesh, seed):
    errs, perrs = [], []
    for lev in (3, 4, 5):
        m = mesh(seed, lev)
        g = L.sphere_harmonic54(m.cx, m.cy, m.cz) * m.area
        tgt = m.targets()
        u = O.bs_direct(tgt, (m.cx, m.cy, m.cz), g, skip=m.face_self_skip())
        psi = O.greens_direct(tgt, (m.cx, m.cy, m.cz), g, skip=m.face_self_skip())
        errs.append(O.rel_l2(u, L.rh54_velocity(*tgt)))
        perrs.append(O.rel_max(psi, L.sphere_harmonic54(*tgt) / 30.0))
    # at least first-order convergence of velocity and potential
    assert errs[-1] < 0.01 and errs[0] / errs[1] > 1.8 and errs[1] / errs[2] > 1.8
    assert perrs[-1] < 0.06 and perrs[0] / perrs[1] > 2.0 and perrs[1] / perrs[2] > 2.0


def test_gamma_of_y54_integrates_to_zero(mesh):
    m = mesh(L.LPM_SEED_ICOS_TRI, 5)
    g = L.sphere_harmonic54(m.cx, m.cy, m.cz) * m.area
    assert abs(g.sum()) < 1e-12
