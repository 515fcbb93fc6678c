"""Pins for the oracle's treecode restatement (oracle/lpm_oracle_tree.c).

  * coefficients: the reference recurrence (SumNode::calc_coeffs,
    src/LpmTreeSum.cpp:149-171) and SphereGreensSeries::computeCoeffs
    (src/LpmTaylorSeries3d.cpp:63-117) both equal the closed form
    a_k = |k|!/k! t^k / D^{|k|+1};
  * MultiIndex known answers (tests/LpmMultiIndexUnitTests.cpp:28-35);
  * the single-box series check of tests/LpmSakajoUnitTests.cpp:65-116: the
    reference's moment sign stalls near 4e-3 for every order, the consistent
    sign converges geometrically; the Cartesian series converges too;
  * octree: brute-force descent of the box recursion for every particle,
    leaf sizes, octant boxes, ascending leaf lists;
  * traversal: theta -> 0 is the direct sum, macCounter = 0 for theta <= 0,
    Cartesian-series error falls with p at fixed theta.
"""
from math import factorial

import numpy as np
import pytest

import lpmcxx_amd as L
from oracle import oracle as O


def _closed_form(t, D, p):
    out = []
    for (a, b, c) in O.multi_indices(p):
        n = a + b + c
        out.append(factorial(n) / (factorial(a) * factorial(b) * factorial(c)) * t[0] ** a * t[1] ** b * t[2] ** c
                   / D ** (n + 1))
    return np.array(out)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_coefficients_closed_form(seed):
    rng = np.random.default_rng(seed)
    t = rng.normal(size=3); t /= np.linalg.norm(t)
    c = rng.normal(size=3) * 0.3
    delta = 0.1 * seed
    for p in (1, 2, 5, 9):
        a = O.coeffs_lex(t, c, delta, p)
        D = 1 + delta * delta - t @ c
        np.testing.assert_allclose(a, _closed_form(t, D, p), rtol=1e-13, atol=0)
    s = rng.normal(size=3) * 0.4
    for mp in (0, 1, 4, 8):
        a = O.coeffs_series(t, s, mp)
        np.testing.assert_allclose(a, _closed_form(t, 1 - t @ s, mp + 1), rtol=1e-13, atol=0)


def test_multi_index_known_answers():
    # MultiIndex(4,3,1): magnitude 8, vectorPower((2,3,0.5)) = 216, factorial 144
    k = (4, 3, 1)
    assert sum(k) == 8
    assert 2.0 ** 4 * 3.0 ** 3 * 0.5 ** 1 == 216.0
    assert factorial(4) * factorial(3) * factorial(1) == 144
    # lex order (the std::map order of src/LpmMultiIndex.h:34-43) and |k| < p count
    ks = O.multi_indices(4)
    assert ks == sorted(ks) and len(ks) == O.nterms(4) == 20


def _sakajo():
    # box and points of tests/LpmSakajoUnitTests.cpp:65-97
    lo = np.array([0.37, 0.525, 0.74]); hi = np.array([0.38, 0.55, 0.76])
    c = 0.5 * (hi + lo)
    t = -np.ones(3) / np.sqrt(3.0)
    y = np.array([0.376867, 0.538382, 0.753735]); y /= np.linalg.norm(y)
    return c, t, y


def test_single_box_series():
    c, t, y = _sakajo()
    exact = np.cross(t, y) / (1 - t @ y)
    errs = {}
    for sign in (-1, 1):
        for series in (0, 1):
            errs[sign, series] = [np.linalg.norm(O.node_far(c, t, tuple(y[:, None]), [1.0], p, sign, series) - exact)
                                  / np.linalg.norm(exact) for p in (2, 4, 8)]
    # reference sign: stalls at ~3.8e-3 whatever the order
    e = errs[-1, 0]
    assert all(2e-3 < x < 6e-3 for x in e) and abs(e[2] / e[0] - 1) < 0.2
    # consistent sign: geometric convergence to round-off
    e = errs[1, 0]
    assert e[0] < 1e-5 and e[1] < 1e-10 and e[2] < 1e-14
    e = errs[1, 1]
    assert e[0] < 1e-3 and e[1] < 1e-5 and e[2] < 1e-9 and e[0] > e[1] > e[2]


def _rand_sphere(n, seed):
    v = np.random.default_rng(seed).normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    return v


@pytest.mark.parametrize("n,leaf", [(1, 4), (50, 64), (3000, 16), (5000, 1)])
def test_octree_brute_force(n, leaf):
    P = _rand_sphere(n, n)
    maxr = float(np.sqrt((P ** 2).sum(1).max()))
    T = O.tree_build(tuple(P.T), maxr, leaf)
    nn = len(T["level"])
    lo = -maxr - 1e-5
    assert np.all(T["box"][0, :3] == lo) and np.all(T["box"][0, 3:] == maxr + 1e-5)
    seen = np.zeros(n, int)
    for i in range(nn):
        b, e = T["begin"][i], T["end"][i]
        idx = T["perm"][b:e]
        bx = T["box"][i]
        # the box holds its particles, half-open
        assert np.all((P[idx] >= bx[:3]) & (P[idx] < bx[3:]))
        if T["n_kids"][i] == 0:
            assert e - b <= leaf or T["level"][i] > 55
            assert np.all(np.diff(idx) > 0)          # ascending source index
            seen[idx] += 1
        else:
            assert e - b > leaf
            mid = 0.5 * (bx[:3] + bx[3:])
            f, k = T["first_kid"][i], T["n_kids"][i]
            octs = []
            for kid in range(f, f + k):
                kb = T["box"][kid]
                o = sum(int(kb[d] == mid[d]) << d for d in range(3))
                for d in range(3):
                    lo_d, hi_d = (mid[d], bx[3 + d]) if (o >> d) & 1 else (bx[d], mid[d])
                    assert kb[d] == lo_d and kb[3 + d] == hi_d
                octs.append(o)
            assert octs == sorted(octs)                # bisectAll order
            assert T["begin"][f] == b and T["end"][f + k - 1] == e
    assert np.all(seen == 1)
    # brute-force descent: the leaf reached by the box recursion lists the particle
    leaf_of = np.empty(n, int)
    for i in range(nn):
        if T["n_kids"][i] == 0:
            leaf_of[T["perm"][T["begin"][i]:T["end"][i]]] = i
    for j in range(0, n, max(1, n // 200)):
        i = 0
        while T["n_kids"][i] > 0:
            f, k = T["first_kid"][i], T["n_kids"][i]
            i = next(kid for kid in range(f, f + k)
                     if np.all((P[j] >= T["box"][kid, :3]) & (P[j] < T["box"][kid, 3:])))
        assert i == leaf_of[j]
    # maxRadius is the half diagonal
    half = 0.5 * (T["box"][:, 3:] - T["box"][:, :3])
    np.testing.assert_allclose(T["cent"][:, 3], np.linalg.norm(half, axis=1), rtol=1e-12)


def test_outside_root_box_rejected():
    P = _rand_sphere(100, 3)
    with pytest.raises(ValueError):
        O.tree_build(tuple(P.T), 0.5, 8)


def test_traversal_limits(mesh):
    m = mesh(L.LPM_SEED_CUBED, 4)
    g = L.gaussian_vortex(m.cx, m.cy, m.cz, m.area) * m.area
    src, tgt = (m.cx, m.cy, m.cz), m.targets()
    ref = O.bs_direct(tgt, src, g)
    O.tree_build(src, m.max_radius, 16)
    mom = O.tree_moments(src, g, 4, 1)
    for series in (0, 1):
        u, mac, per = O.tree_eval(src, g, mom, 4, 0.0, tgt, series=series)
        assert mac == 0 and np.all(per == 0)
        assert O.rel_l2(u, ref) < 1e-14
        u, mac, _ = O.tree_eval(src, g, mom, 4, -1.0, tgt, series=series)
        assert mac == 0


def test_cartesian_series_converges_in_p(mesh):
    m = mesh(L.LPM_SEED_CUBED, 5)
    g = L.gaussian_vortex(m.cx, m.cy, m.cz, m.area) * m.area
    src, tgt = (m.cx, m.cy, m.cz), m.targets()
    ref = O.bs_direct(tgt, src, g)
    O.tree_build(src, m.max_radius, 32)
    errs = []
    for p in (2, 4, 6, 8, 10):
        mom = O.tree_moments(src, g, p, 1)
        u, mac, _ = O.tree_eval(src, g, mom, p, 0.5, tgt, series=1)
        errs.append(O.rel_l2(u, ref))
    assert all(a > 3 * b for a, b in zip(errs, errs[1:]))
    assert errs[-1] < 1e-5


# ---------------------------------------------------------------- batch acceptance and labels
def _morton_keys(P, lo, hi):
    """Independent restatement of the 21-level target keys (clamped into the box)."""
    p = np.clip(P, lo, hi)
    l, h = np.tile(lo, (len(P), 1)), np.tile(hi, (len(P), 1))
    key = np.zeros(len(P), np.uint64)
    for _ in range(21):
        m = 0.5 * (l + h)
        b = p >= m
        l = np.where(b, m, l)
        h = np.where(b, h, m)
        key = (key << np.uint64(3)) | (b[:, 0] | (b[:, 1] << 1) | (b[:, 2] << 2)).astype(np.uint64)
    return key


def test_target_order_is_stable_morton(mesh):
    m = mesh(L.LPM_SEED_CUBED, 4)
    src = (m.cx, m.cy, m.cz)
    R = O.tree_build(src, m.max_radius, 32)
    rng = np.random.default_rng(1)
    P = rng.normal(size=(3000, 3)) * 0.7
    P[:20] = P[20]                               # ties keep index order
    order = O.target_order(tuple(P.T.copy()))
    lo, hi = R["box"][0, :3], R["box"][0, 3:]
    key = _morton_keys(P, lo, hi)
    # ordered on the top 12 of the 21 key levels, ties by index
    assert np.array_equal(order, np.argsort(key >> np.uint64(27), kind="stable"))
    O.set_target_levels(21)
    try:
        assert np.array_equal(O.target_order(tuple(P.T.copy())), np.argsort(key, kind="stable"))
    finally:
        O.set_target_levels(12)


def test_target_batches_are_cell_aligned(mesh):
    m = mesh(L.LPM_SEED_CUBED, 5)
    src = (m.cx, m.cy, m.cz)
    R = O.tree_build(src, m.max_radius, 32)
    tgt = m.targets()
    n = len(tgt[0])
    order, bstart = O.target_batches(tgt)
    assert np.array_equal(order, O.target_order(tgt))
    sizes = np.diff(np.append(bstart, n))
    assert bstart[0] == 0 and np.all(sizes >= 1) and np.all(sizes <= 64)
    key = _morton_keys(np.stack(tgt, 1)[order], R["box"][0, :3], R["box"][0, 3:])
    # the level of the split: shallowest with >= n/256 occupied cells
    lcp = np.array([21 if a == b else (62 - (int(a ^ b).bit_length() - 1)) // 3 for a, b in zip(key[:-1], key[1:])])
    cells = [1 + np.sum(lcp < l) for l in range(22)]
    lev = next((l for l in range(12) if cells[l] >= max(1, n // 256)), 12)
    cell_start = np.maximum.accumulate(np.where(np.append(True, lcp < lev), np.arange(n), 0))
    expect = np.flatnonzero((np.arange(n) - cell_start) % 64 == 0)
    assert np.array_equal(bstart, expect)
    # every batch inside one level-L cell, and compact
    shift = np.uint64(3 * (21 - lev))
    for b0, sz in zip(bstart, sizes):
        assert len(set((key[b0:b0 + sz] >> shift).tolist())) == 1


def test_batch_oracle_limits(mesh):
    m = mesh(L.LPM_SEED_CUBED, 4)
    src = (m.cx, m.cy, m.cz)
    g = L.random_vortices(m.cx, m.cy, m.cz, m.area) * m.area
    tgt = m.targets()
    O.tree_build(src, m.max_radius, 32)
    mom = O.tree_moments(src, g, 8, 1)
    ud = np.stack(O.bs_direct(tgt, src, g))
    u0, far, near = O.tree_eval_batch(src, g, mom, 8, 0.0, tgt)
    assert far == 0 and near == len(tgt[0]) * len(src[0])
    assert O.rel_max(np.stack(u0), ud) < 1e-13
    errs = []
    for p in (2, 4, 6, 8):
        mp = O.tree_moments(src, g, p, 1)
        u, far, near = O.tree_eval_batch(src, g, mp, p, 0.6, tgt)
        assert far > 0
        errs.append(O.rel_l2(np.stack(u), ud))
    assert all(b < 0.5 * a for a, b in zip(errs, errs[1:])), errs
    # the batch criterion is stricter than the per-target one: more accurate at equal (p, theta)
    (up, _, _) = O.tree_eval(src, g, mom, 8, 0.6, tgt, series=1)
    assert errs[-1] <= O.rel_l2(np.stack(up), ud)


def test_pair_errors_match_single_box_series(mesh):
    """Label errors equal |t x (S_q - S_exact)| / (4 pi u_ref) from the
    single-box far field (independent moment code) and the direct sum."""
    m = mesh(L.LPM_SEED_CUBED, 4)
    src = (m.cx, m.cy, m.cz)
    g = L.random_vortices(m.cx, m.cy, m.cz, m.area) * m.area
    tgt = m.targets()
    R = O.tree_build(src, m.max_radius, 32)
    mom = O.tree_moments(src, g, 12, 1)
    order, bstart = O.target_batches(tgt)
    u_ref = float(np.abs(g).sum() / (4 * np.pi))
    node = int(np.flatnonzero(R["level"] == 2)[3])
    batch = 37
    err = O.pair_errors(src, g, mom, 12, tgt, batch, node, u_ref)
    ids = R["perm"][R["begin"][node]:R["end"][node]]
    ps = tuple(a[ids] for a in src)
    c = R["cent"][node, :3]
    ref = np.zeros(12)
    for i in order[bstart[batch]:bstart[batch + 1]]:
        t = np.array([tgt[0][i], tgt[1][i], tgt[2][i]])
        y = np.stack(ps, 1)
        ex = np.cross(t, (g[ids][:, None] * y / (1.0 - y @ t)[:, None]).sum(0))
        for q in range(1, 13):
            far = O.node_far(c, t, ps, g[ids], q, 1, 1)
            ref[q - 1] = max(ref[q - 1], np.linalg.norm(far - ex) / (4 * np.pi * u_ref))
    assert np.allclose(err, ref, rtol=1e-9, atol=1e-16)
    assert err[-1] < 1e-3 * err[0]


def _node_table(R):
    """Parent-before-child node order and each node's child list from an exported oracle tree."""
    nn = len(R["level"])
    order = np.argsort(R["level"], kind="stable")
    kids = [list(range(int(R["first_kid"][i]), int(R["first_kid"][i] + R["n_kids"][i]))) for i in range(nn)]
    return order, kids


@pytest.mark.parametrize("level", [4, 5])
@pytest.mark.parametrize("theta", [0.33, 0.5, 0.7])
def test_mac_count_recount(mesh, level, theta):
    """Per-target macCounter of the reference traversal (src/LpmTreeSum.cpp:106-137,
    MAC :22-24) recounted in numpy from the oracle's node table: a node is
    counted for a target when maxRadius / distance(centroid, target) < theta
    and no ancestor was; distance in the reference's operation order."""
    m = mesh(L.LPM_SEED_CUBED, level)
    src = (m.cx, m.cy, m.cz)
    R = O.tree_build(src, m.max_radius, 32)
    tgt = m.targets()
    g = np.ones(len(src[0]))
    mom = O.tree_moments(src, g, 1, 1)
    _, total, per_target = O.tree_eval(src, g, mom, 1, theta, tgt, series=1)
    order, kids = _node_table(R)
    tx, ty, tz = tgt
    reached = {0: np.ones(tx.size, bool)}
    count = np.zeros(tx.size, np.int64)
    for i in order:
        r = reached.pop(i)
        c = R["cent"][i]
        d = np.sqrt((tx - c[0]) * (tx - c[0]) + (ty - c[1]) * (ty - c[1]) + (tz - c[2]) * (tz - c[2]))
        acc = r & (c[3] / d < theta)
        count += acc
        for k in kids[i]:
            reached[k] = r & ~acc
    assert np.array_equal(count, per_target)
    assert total == count.sum() > 0


def _butterfly_sum(v):
    """Sum over the last axis (64 lanes) in the order of a wave's xor butterfly."""
    v = v.copy()
    lanes = np.arange(64)
    for o in (32, 16, 8, 4, 2, 1):
        v = v + v[..., lanes ^ o]
    return v[..., 0]


@pytest.mark.parametrize("level", [4, 5])
@pytest.mark.parametrize("theta", [0.33, 0.5, 0.7])
def test_batch_pair_counts_recount(mesh, level, theta):
    """Far (batch, node) and near (target, source) pair counts of the batch
    traversal recounted in numpy: batch centroid by the butterfly sum times
    1/nv, radius the largest distance; node source radius exact at leaves
    and max(|c_k - c| + r_k) above; accept when r_n + r_b < theta |c_b - c_n|."""
    m = mesh(L.LPM_SEED_CUBED, level)
    src = (m.cx, m.cy, m.cz)
    R = O.tree_build(src, m.max_radius, 32)
    tgt = m.targets()
    g = np.ones(len(src[0]))
    _, far, near = O.tree_eval_batch(src, g, O.tree_moments(src, g, 1, 1), 1, theta, tgt)
    order, bstart = O.target_batches(tgt)
    n = len(tgt[0])
    bend = np.append(bstart[1:], n)
    nb = len(bstart)
    P = np.stack(tgt, 1)[order]
    T = np.zeros((nb, 64, 3))
    nv = bend - bstart
    for b in range(nb):
        T[b, :nv[b]] = P[bstart[b]:bend[b]]
    cb = np.stack([_butterfly_sum(T[:, :, d]) * (1.0 / nv) for d in range(3)], 1)
    lane_ok = np.arange(64)[None, :] < nv[:, None]
    d2 = ((T[:, :, 0] - cb[:, None, 0]) * (T[:, :, 0] - cb[:, None, 0]) +
          (T[:, :, 1] - cb[:, None, 1]) * (T[:, :, 1] - cb[:, None, 1]) +
          (T[:, :, 2] - cb[:, None, 2]) * (T[:, :, 2] - cb[:, None, 2]))
    rb = np.sqrt(np.where(lane_ok, d2, 0.0).max(1))
    nodes, kids = _node_table(R)
    S = np.stack(src, 1)
    nrad = np.zeros(len(R["level"]))
    for i in nodes[::-1]:                       # children before parents
        c = R["cent"][i, :3]
        if not kids[i]:
            q = S[R["perm"][R["begin"][i]:R["end"][i]]]
            nrad[i] = np.sqrt(((q[:, 0] - c[0]) * (q[:, 0] - c[0]) + (q[:, 1] - c[1]) * (q[:, 1] - c[1]) +
                               (q[:, 2] - c[2]) * (q[:, 2] - c[2])).max())
        else:
            nrad[i] = max(np.sqrt((R["cent"][k, 0] - c[0]) ** 2 + (R["cent"][k, 1] - c[1]) ** 2 +
                                  (R["cent"][k, 2] - c[2]) ** 2) + nrad[k] for k in kids[i])
    reached = {0: np.ones(nb, bool)}
    far_n = near_n = 0
    for i in nodes:
        r = reached.pop(i)
        c = R["cent"][i]
        s = np.sqrt((cb[:, 0] - c[0]) * (cb[:, 0] - c[0]) + (cb[:, 1] - c[1]) * (cb[:, 1] - c[1]) +
                    (cb[:, 2] - c[2]) * (cb[:, 2] - c[2]))
        acc = r & (nrad[i] + rb < theta * s)
        far_n += int(nv[acc].sum())
        rest = r & ~acc
        if not kids[i]:
            near_n += int(nv[rest].sum()) * int(R["end"][i] - R["begin"][i])
        for k in kids[i]:
            reached[k] = rest
    assert (far_n, near_n) == (far, near)
    assert far > 0 and near > 0


def test_target_batches_clustered_random_order(mesh):
    """Targets clustered in a 0.005-rad cap, in random order: the batch level
    stops at the 12 sorted key levels, every batch stays inside one level-L
    cell, and the batch count stays within n/64 + 8 max(1, n/256) (the launch
    bound of the GPU batch kernels)."""
    m = mesh(L.LPM_SEED_CUBED, 4)
    R = O.tree_build((m.cx, m.cy, m.cz), m.max_radius, 32)
    rng = np.random.default_rng(4)
    n = 20000
    th = 0.005 * np.sqrt(rng.uniform(size=n))
    ph = rng.uniform(0, 2 * np.pi, n)
    P = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)
    tgt = tuple(P.T.copy())
    order, bstart = O.target_batches(tgt)
    want = max(1, n // 256)
    assert len(bstart) <= n // 64 + 1 + 8 * want
    key = _morton_keys(P[order], R["box"][0, :3], R["box"][0, 3:])
    top = key >> np.uint64(27)
    assert np.all(top[1:] >= top[:-1])        # sorted on the top 12 levels
    sizes = np.diff(np.append(bstart, n))
    assert np.all(sizes >= 1) and np.all(sizes <= 64)
    lcp = np.array([12 if a == b else (35 - (int(a ^ b).bit_length() - 1)) // 3 for a, b in zip(top[:-1], top[1:])])
    cells = [1 + np.sum(lcp < l) for l in range(13)]
    lev = next((l for l in range(12) if cells[l] >= want), 12)
    shift = np.uint64(3 * (12 - lev))
    for b0, sz in zip(bstart, sizes):
        assert len(set((top[b0:b0 + sz] >> shift).tolist())) == 1
