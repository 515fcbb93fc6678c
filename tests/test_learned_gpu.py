"""GPU tests of the batch treecode, the order-model kernels and the learned
treecode.

Bars: batch treecode (FP64) equal to the oracle's batch traversal within
1e-12 with identical far/near pair counts; FP32C within 1e-6 of FP64; model
gradients of the BF16 MFMA kernels within 5e-3 (relative L2) of the oracle's
BF16-emulating backward pass; labels within 1e-5 of the oracle's per-pair
errors; the learned traversal reducing exactly to the plain one for a
constant model; model inputs clamped to the range training recorded.
"""
import numpy as np
import pytest

import lpmcxx_amd as L
from lpmcxx_amd import learned as LN
from lpmcxx_amd import tree as T
from oracle import mlp_oracle as MO
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _field(mesh, level, leaf, order, seed=L.LPM_SEED_CUBED):
    m = mesh(seed, level)
    src = (m.cx, m.cy, m.cz)
    g = L.random_vortices(m.cx, m.cy, m.cz, m.area) * m.area
    return m, src, g, m.targets()


def _gpu_tree(ctx, m, src, g, leaf, order):
    T.build(ctx, src, m.max_radius, leaf)
    T.moments(ctx, g, order, 1)


# ---------------------------------------------------------------- batch treecode
@pytest.mark.parametrize("level,leaf,p,theta", [(5, 64, 6, 0.5), (5, 32, 8, 0.7), (4, 16, 4, 0.9)])
def test_batch_matches_oracle(ctx, mesh, level, leaf, p, theta):
    m, src, g, tgt = _field(mesh, level, leaf, p)
    _gpu_tree(ctx, m, src, g, leaf, p)
    u = T.evaluate_batch(ctx, tgt, theta, p, prec=L.LPM_F64)
    st = T.stats(ctx)
    O.tree_build(src, m.max_radius, leaf)
    mom = O.tree_moments(src, g, p, 1)
    ur, far, near = O.tree_eval_batch(src, g, mom, p, theta, tgt)
    assert O.rel_max(np.stack(u), np.stack(ur)) < 1e-12
    assert (st["far"], st["near"]) == (far, near)
    assert st["terms"] == far * p * (p + 1) * (p + 2) // 6
    u32 = T.evaluate_batch(ctx, tgt, theta, p, prec=L.LPM_F32C)
    assert O.rel_l2(np.stack(u32), np.stack(u)) < 1e-6


def test_batch_theta_zero_is_direct_sum(ctx, mesh):
    m, src, g, tgt = _field(mesh, 4, 32, 4)
    _gpu_tree(ctx, m, src, g, 32, 4)
    u = T.evaluate_batch(ctx, tgt, 0.0, 4)
    ud = O.bs_direct(tgt, src, g)
    assert O.rel_max(np.stack(u), np.stack(ud)) < 1e-12
    assert T.stats(ctx)["far"] == 0


def test_batch_converges_and_ragged(ctx, mesh):
    m, src, g, tgt = _field(mesh, 6, 128, 12)
    _gpu_tree(ctx, m, src, g, 128, 12)
    ud = np.stack(ctx.direct_sum(tgt, src, g))
    errs = [O.rel_l2(np.stack(T.evaluate_batch(ctx, tgt, 0.5, p)), ud) for p in (2, 4, 6, 8, 10, 12)]
    assert all(b < a for a, b in zip(errs, errs[1:])), errs
    assert errs[-1] < 1e-7
    # ragged batch counts (one target, partial waves) and an empty target set, against the oracle
    O.tree_build(src, m.max_radius, 128)
    mom = O.tree_moments(src, g, 8, 1)
    for n in (1, 63, 65, 1000):
        sub = tuple(a[:n].copy() for a in tgt)
        us = T.evaluate_batch(ctx, sub, 0.5, 8)
        ur, far, near = O.tree_eval_batch(src, g, mom, 8, 0.5, sub)
        assert O.rel_max(np.stack(us), np.stack(ur)) < 1e-12
        assert (T.stats(ctx)["far"], T.stats(ctx)["near"]) == (far, near)
    T.evaluate_batch(ctx, tuple(a[:0].copy() for a in tgt), 0.5, 8)


def test_batch_errors(ctx, mesh):
    m, src, g, tgt = _field(mesh, 3, 32, 4)
    _gpu_tree(ctx, m, src, g, 32, 4)
    with pytest.raises(L.LpmError):            # order above the moments' order
        T.evaluate_batch(ctx, tgt, 0.5, 6)
    with pytest.raises(L.LpmError):
        T.evaluate_batch(ctx, tgt, 0.5, 4, prec=7)
    T.moments(ctx, g, 4, -1)
    with pytest.raises(L.LpmError):            # reference moment sign
        T.evaluate_batch(ctx, tgt, 0.5, 4)


# ---------------------------------------------------------------- model kernels
def _rand_model(ctx, seed):
    mod = LN.OrderModel(ctx)
    mod.init(seed)
    params, _, _ = mod.get_params()
    rng = np.random.default_rng(seed)
    params = params + rng.normal(0, 0.05, params.size).astype(np.float32)  # nonzero biases too
    mod.set_params(params)
    return mod, params


def test_model_gradient_matches_oracle(ctx):
    mod, params = _rand_model(ctx, 3)
    rng = np.random.default_rng(4)
    for n in (64, 333):
        x = np.zeros((n, 16), np.float32)
        x[:, :10] = rng.normal(size=(n, 10))
        y = rng.integers(0, 13, n).astype(np.int32)
        g, loss = mod.gradient(x, y)
        lr, gr = MO.loss_and_grad(params, x, y, bf=True)
        assert abs(loss - lr) < 1e-4 * max(1.0, abs(lr))
        assert np.linalg.norm(g - gr) / np.linalg.norm(gr) < 5e-3
        # every layer separately
        for name, a in LN.unpack(g).items():
            b = MO.unpack(gr)[name]
            assert np.linalg.norm(a - b) <= 1e-2 * np.linalg.norm(b) + 1e-7, name


def test_model_forward_matches_oracle(ctx):
    mod, params = _rand_model(ctx, 5)
    rng = np.random.default_rng(6)
    x = np.zeros((517, 16), np.float32)
    x[:, :10] = rng.normal(size=(517, 10))
    logits, cls = mod.forward(x)
    z, _ = MO.forward(params, x, bf=True)
    assert np.max(np.abs(logits - z)) < 1e-3 * max(1.0, np.abs(z).max())
    assert np.mean(cls == np.argmax(z[:, :13], axis=1)) > 0.99


def test_model_params_roundtrip_and_init(ctx):
    mod = LN.OrderModel(ctx)
    mod.init(11)
    a, _, _ = mod.get_params()
    mod.init(11)
    b, _, _ = mod.get_params()
    assert np.array_equal(a, b)
    mean = np.arange(16, dtype=np.float32)
    inv = np.full(16, 0.5, np.float32)
    mod.set_params(a * 2, mean, inv)
    c, m2, i2 = mod.get_params()
    assert np.array_equal(c, a * 2) and np.array_equal(m2, mean) and np.array_equal(i2, inv)
    with pytest.raises(L.LpmError):
        mod.gradient(np.zeros((4, 16), np.float32), np.array([0, 1, 2, 13], np.int32))


def test_model_training_learns_rule(ctx):
    """Adam through the GPU kernels learns a fixed piecewise rule of the inputs."""
    rng = np.random.default_rng(7)
    n = 20000
    x = np.zeros((n, 16), np.float32)
    x[:, :10] = rng.normal(size=(n, 10))
    y = np.clip(np.floor(2.0 * x[:, 0] + 6.0), 0, 12).astype(np.int32)
    mod = LN.OrderModel(ctx)
    mod.init(1)
    p, _, _ = mod.get_params()
    mean, inv = LN.normalization(x)
    mod.set_params(p, mean, inv)
    _, l0 = mod.gradient(x[:4096], y[:4096])
    loss = mod.train(x, y, epochs=15, batch=1024, lr=3e-3, seed=2)
    _, cls = mod.forward(x)
    assert loss < 0.5 * l0
    assert np.mean(np.abs(cls - y) <= 1) > 0.97


def test_model_input_range(ctx):
    """Training records the per-input range of its rows; forward clamps inputs
    to it before normalizing; set_params and set_range(None) unbound it."""
    rng = np.random.default_rng(8)
    x = np.zeros((3000, 16), np.float32)
    x[:, :10] = rng.normal(size=(3000, 10))
    y = np.clip(np.floor(2.0 * x[:, 0] + 6.0), 0, 12).astype(np.int32)
    mod = LN.OrderModel(ctx)
    mod.init(1)
    p, _, _ = mod.get_params()
    mean, inv = LN.normalization(x)
    mod.set_params(p, mean, inv)
    lo, hi = mod.get_range()
    assert (lo == -3e38).all() and (hi == 3e38).all()
    mod.train(x, y, epochs=1, batch=1024, seed=2)
    lo, hi = mod.get_range()
    assert np.array_equal(lo, x.min(0)) and np.array_equal(hi, x.max(0))
    xo = x[:200] * 40.0                              # far outside the training range
    zo, co = mod.forward(xo)
    zc, cc = mod.forward(np.clip(xo, lo, hi))
    assert np.array_equal(zo, zc) and np.array_equal(co, cc)
    mod.set_range(None)
    zu, _ = mod.forward(xo)
    assert not np.array_equal(zu, zo)
    mod.set_range(lo, hi)
    assert np.array_equal(mod.forward(xo)[0], zo)
    with pytest.raises(L.LpmError):
        mod.set_range(hi, lo)
    with pytest.raises(L.LpmError):
        mod.train(np.full((4, 16), np.nan, np.float32), np.zeros(4, np.int32), epochs=1)
    p2, _, _ = mod.get_params()
    mod.set_params(p2, mean, inv)
    assert (mod.get_range()[0] == -3e38).all()


# ---------------------------------------------------------------- labels and learned traversal
def _radius(R, src, i):
    """Source radius of oracle node i: exact at a leaf, max over children of
    |c_child - c| + r_child above."""
    c = R["cent"][i, :3]
    if R["n_kids"][i] == 0:
        ids = R["perm"][R["begin"][i]:R["end"][i]]
        return float(np.sqrt(((np.stack([a[ids] for a in src], 1) - c) ** 2).sum(1).max()))
    f = R["first_kid"][i]
    return max(float(np.linalg.norm(R["cent"][k, :3] - c)) + _radius(R, src, k)
               for k in range(f, f + R["n_kids"][i]))


def test_labels_match_oracle(ctx, mesh):
    m, src, g, tgt = _field(mesh, 5, 64, 12)
    _gpu_tree(ctx, m, src, g, 64, 12)
    u_ref = LN.u_ref_of(g)
    ls = LN.labels(ctx, tgt, 0.8, u_ref, batch_stride=7, keep_rate=1.0, max_samples=1 << 16)
    assert len(ls.ids) > 100
    assert np.median(ls.errors[:, -1] / np.maximum(ls.errors[:, 0], 1e-300)) < 0.1  # the series converges
    O.tree_build(src, m.max_radius, 64)
    mom = O.tree_moments(src, g, 12, 1)
    G = T.export(ctx, len(src[0]))
    R = O.tree_build(src, m.max_radius, 64)
    gi = T.dfs_order(G["first_kid"], G["n_kids"])
    ri = T.dfs_order(R["first_kid"], R["n_kids"])
    g2o = np.empty(len(gi), np.int64)
    g2o[gi] = ri
    rng = np.random.default_rng(0)
    for k in rng.choice(len(ls.ids), 25, replace=False):
        b, node = ls.ids[k]
        ref = O.pair_errors(src, g, mom, 12, tgt, int(b), int(g2o[node]), u_ref)
        got = ls.errors[k].astype(np.float64)
        big = ref > 1e-12
        assert np.all(np.abs(got[big] - ref[big]) <= 1e-5 * ref[big]), (k, got, ref)
        assert np.all(np.abs(got[~big] - ref[~big]) <= 2e-12)
        # features: node geometry from the oracle tree
        c = R["cent"][g2o[node]]
        assert ls.features[k, 9] == (1.0 if R["n_kids"][g2o[node]] == 0 else 0.0)
        assert abs(np.exp2(ls.features[k, 6]) - (R["end"][g2o[node]] - R["begin"][g2o[node]])) < 1e-3
        order, bstart = O.target_batches(tgt)
        b1 = bstart[b + 1] if b + 1 < len(bstart) else len(order)
        T_b = np.stack([a[order[bstart[b]:b1]] for a in tgt], 1)
        cb = T_b.mean(0)
        s = np.linalg.norm(cb - c[:3])
        assert abs(ls.features[k, 8] - np.log2(s)) < 1e-5
        rn = _radius(R, src, int(g2o[node]))
        assert abs(ls.features[k, 0] - np.log2(rn / s)) < 1e-5


def test_learned_constant_model_reduces_to_plain(ctx, mesh):
    m, src, g, tgt = _field(mesh, 5, 64, 8)
    _gpu_tree(ctx, m, src, g, 64, 8)
    u_ref = LN.u_ref_of(g)
    for prec in (L.LPM_F64, L.LPM_F32C):
        LN.constant_model(ctx, 6)
        ul = LN.evaluate(ctx, tgt, 0.6, 0, 8, u_ref, prec=prec)
        sl = T.stats(ctx)
        ub = T.evaluate_batch(ctx, tgt, 0.6, 6, prec=prec)
        sb = T.stats(ctx)
        assert O.rel_max(np.stack(ul), np.stack(ub)) < 1e-14
        assert (sl["far"], sl["near"], sl["terms"]) == (sb["far"], sb["near"], sb["terms"])
        assert sl["model"] > 0
        # margin raises the order; class 0 (always descend) is the direct sum
        LN.constant_model(ctx, 4)
        um = LN.evaluate(ctx, tgt, 0.6, 2, 8, u_ref, prec=prec)
        assert O.rel_max(np.stack(um), np.stack(ub)) < 1e-14
    LN.constant_model(ctx, 0)
    u0 = LN.evaluate(ctx, tgt, 0.6, 0, 8, u_ref, prec=L.LPM_F64)
    assert T.stats(ctx)["far"] == 0
    assert O.rel_max(np.stack(u0), np.stack(O.bs_direct(tgt, src, g))) < 1e-12
    # a class above max_order descends as well
    LN.constant_model(ctx, 9)
    u9 = LN.evaluate(ctx, tgt, 0.6, 0, 8, u_ref, prec=L.LPM_F64)
    assert T.stats(ctx)["far"] == 0 and O.rel_max(np.stack(u9), np.stack(u0)) < 1e-14


@pytest.mark.parametrize("order", [6, 8, 10, 12])
def test_learned_constant_model_all_list_kernels(ctx, mesh, order):
    """Orders 8..12 go through the second list kernel: a constant model of
    order p must reproduce the fixed-order batch treecode in both precisions."""
    m, src, g, tgt = _field(mesh, 5, 64, 12)
    _gpu_tree(ctx, m, src, g, 64, 12)
    u_ref = LN.u_ref_of(g)
    for prec in (L.LPM_F64, L.LPM_F32C):
        LN.constant_model(ctx, order)
        ul = LN.evaluate(ctx, tgt, 0.9, 0, order, u_ref, prec=prec)
        sl = T.stats(ctx)
        ub = T.evaluate_batch(ctx, tgt, 0.9, order, prec=prec)
        sb = T.stats(ctx)
        d = O.rel_max(np.stack(ul), np.stack(ub))
        print(f"\norder {order} prec {prec}: learned(const) vs batch rel max {d:.3e}, far {sl['far']} high {sl['far_high']}")
        assert (sl["far"], sl["near"], sl["terms"]) == (sb["far"], sb["near"], sb["terms"])
        assert sl["far_high"] == (sl["far"] if order > 7 else 0)
        assert d < (1e-13 if prec == L.LPM_F64 else 1e-6)


def _clustered(cap, nu=20000, nc=4096, seed=3):
    """nu uniform points on the unit sphere and nc in a cap of angular radius cap."""
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(nu, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    ax = np.array([0.3, -0.5, 0.81])
    ax /= np.linalg.norm(ax)
    t1 = np.cross(ax, [1.0, 0.0, 0.0])
    t1 /= np.linalg.norm(t1)
    t2 = np.cross(ax, t1)
    r, ph = cap * np.sqrt(rng.random(nc)), 2.0 * np.pi * rng.random(nc)
    c = ax + (r * np.cos(ph))[:, None] * t1 + (r * np.sin(ph))[:, None] * t2
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    P = np.vstack([u, c])
    return tuple(np.ascontiguousarray(P[:, k]) for k in range(3)), rng.normal(size=len(P)) * 1e-4


@pytest.mark.parametrize("cap,leaf,min_depth", [(1e-3, 32, 10), (2e-5, 16, 20)])
def test_learned_deep_tree_fewer_rows_per_call(ctx, cap, leaf, min_depth):
    """A tight cluster deepens the tree, and the planning pass then classifies
    fewer children per model call (64 rows up to depth 9, 32 up to depth 19,
    16 beyond: the bound on its stack).  With a constant model it must still
    reproduce the fixed-order batch treecode: the same far and near pairs and
    terms, the same velocity (FP64 within 1e-13, FP32C within 1e-6)."""
    src, g = _clustered(cap)
    nn, depth = T.build(ctx, src, float(np.sqrt(src[0] ** 2 + src[1] ** 2 + src[2] ** 2).max()), leaf)
    print(f"\ncluster cap {cap}, leaf {leaf}: {nn} nodes, depth {depth}")
    assert depth >= min_depth
    T.moments(ctx, g, 8, 1)
    u_ref = LN.u_ref_of(g)
    for prec in (L.LPM_F64, L.LPM_F32C):
        LN.constant_model(ctx, 6)
        ul = LN.evaluate(ctx, src, 0.6, 0, 8, u_ref, prec=prec)
        sl = T.stats(ctx)
        ub = T.evaluate_batch(ctx, src, 0.6, 6, prec=prec)
        sb = T.stats(ctx)
        assert sl["model"] > 0
        assert (sl["far"], sl["near"], sl["terms"]) == (sb["far"], sb["near"], sb["terms"])
        assert O.rel_max(np.stack(ul), np.stack(ub)) < (1e-13 if prec == L.LPM_F64 else 1e-6)


def test_learned_zero_rule_near_duplicates(ctx, mesh):
    """Sources with near-duplicates: 24 partners 2e-7 away (inside the
    reference's zero rule |d|^2 / 2 <= 1e-13: the pair contributes nothing)
    and 24 partners 1e-5 away (a full contribution), targets = sources.
    Class 0 (every pair summed in the near field): FP64 within 1e-12 and
    FP32C within 1e-3 of the oracle's direct sum with the same rule; class 6
    (far + near): FP32C equal to the fixed-order batch treecode to 1e-6."""
    m = mesh(L.LPM_SEED_CUBED, 5)
    P = np.stack([m.cx, m.cy, m.cz], 1)
    rng = np.random.default_rng(8)
    pick = rng.choice(len(P), 48, replace=False)
    extra = []
    for i, d in zip(pick, [2e-7] * 24 + [1e-5] * 24):
        t = np.cross(P[i], rng.normal(size=3))
        q = P[i] + d * t / np.linalg.norm(t)
        extra.append(q / np.linalg.norm(q))
    Q = np.vstack([P, np.array(extra)])
    src = tuple(np.ascontiguousarray(Q[:, k]) for k in range(3))
    g = np.concatenate([m.area, np.full(len(extra), m.area.mean())]) * rng.uniform(0.5, 1.5, len(Q))
    T.build(ctx, src, float(np.sqrt((Q ** 2).sum(1)).max()), 32)
    T.moments(ctx, g, 8, 1)
    u_ref = LN.u_ref_of(g)
    ud = np.stack(O.bs_direct(src, src, g))
    LN.constant_model(ctx, 0)
    u64 = np.stack(LN.evaluate(ctx, src, 0.6, 0, 8, u_ref, prec=L.LPM_F64))
    u32 = np.stack(LN.evaluate(ctx, src, 0.6, 0, 8, u_ref, prec=L.LPM_F32C))
    assert T.stats(ctx)["far"] == 0
    e64, e32 = O.rel_max(u64, ud), O.rel_l2(u32, ud)
    print(f"\nnear duplicates, all pairs near: FP64 rel max {e64:.2e}, FP32C rel-L2 {e32:.2e}")
    assert e64 < 1e-12 and e32 < 1e-3
    LN.constant_model(ctx, 6)
    ul = np.stack(LN.evaluate(ctx, src, 0.6, 0, 8, u_ref, prec=L.LPM_F32C))
    sl = T.stats(ctx)
    ub = np.stack(T.evaluate_batch(ctx, src, 0.6, 6, prec=L.LPM_F32C))
    sb = T.stats(ctx)
    assert (sl["far"], sl["near"]) == (sb["far"], sb["near"])
    assert O.rel_max(ul, ub) < 1e-6


def test_learned_planning_clamps_inputs_to_the_model_range(ctx, mesh):
    """A one-target batch has radius 0, so log2(r_b / s) is -39.9, far below
    anything in training.  A hand-made model answers order 7 when
    log2(r_b / s) > -20 and order 1 otherwise: with the input range
    [-10, 10] the planning pass clamps the feature and the evaluation equals
    the constant order-7 model bit for bit; unbounded, it takes order 1."""
    m, src, g, _ = _field(mesh, 5, 64, 8)
    _gpu_tree(ctx, m, src, g, 64, 8)
    u_ref = LN.u_ref_of(g)
    parts = {name: np.zeros(shape, np.float32) for name, shape in LN.LAYOUT}
    parts["W1"][1, 0] = 1.0      # h1 = relu(log2(r_b / s) + 20)
    parts["b1"][0] = 20.0
    parts["W2"][0, 0] = 1.0
    parts["W3"][0, 7] = 1.0      # logit 7 = h1, logit 1 = 1
    parts["b3"][1] = 1.0
    mod = LN.OrderModel(ctx)
    for i in (0, 1000, 5000):
        tgt = tuple(a[i:i + 1] for a in m.targets())
        for prec in (L.LPM_F64, L.LPM_F32C):
            LN.constant_model(ctx, 7)
            u7 = np.stack(LN.evaluate(ctx, tgt, 0.6, 0, 8, u_ref, prec=prec))
            LN.constant_model(ctx, 1)
            u1 = np.stack(LN.evaluate(ctx, tgt, 0.6, 0, 8, u_ref, prec=prec))
            assert not np.array_equal(u1, u7)
            mod.set_params(LN.pack(parts))
            ua = np.stack(LN.evaluate(ctx, tgt, 0.6, 0, 8, u_ref, prec=prec))
            assert np.array_equal(ua, u1)
            mod.set_range(np.full(16, -10.0), np.full(16, 10.0))
            uc = np.stack(LN.evaluate(ctx, tgt, 0.6, 0, 8, u_ref, prec=prec))
            assert np.array_equal(uc, u7)


# ---------------------------------------------------------------- stereographic far field
def _stereo(ctx, on, extra=1, ratio=1.0):
    ctx.set_option(L.LPM_OPT_STEREO, 1 if on else 0)
    ctx.set_option(L.LPM_OPT_STEREO_EXTRA_ORDER, extra)
    ctx.set_option(L.LPM_OPT_STEREO_RATIO, ratio)


def test_stereo_far_field_against_the_cartesian_series(ctx, mesh):
    """A constant model accepts every pair passing theta_max = 0.4 at order p.
    The pairs whose plane ratio is at most their sphere ratio become
    stereographic entries of order p + extra; over both projections (targets
    on the whole sphere: vertices and faces of cubed level 5) the result is at
    least as accurate as the Cartesian series of order p against the direct
    sum, and two more orders are no less accurate (within 1%: once the
    remaining Cartesian pairs dominate the error, the FP32 rounding of the
    sums decides the last digits)."""
    m, src, g, tgt = _field(mesh, 5, 64, 8)
    ud = np.stack(O.bs_direct(tgt, src, g))
    u_ref = LN.u_ref_of(g)
    err = {}
    try:
        for on, extra, cls in ((False, 0, 6), (False, 0, 8), (True, 0, 6), (True, 2, 6), (True, 0, 8)):
            _stereo(ctx, on, extra, 1.0)
            _gpu_tree(ctx, m, src, g, 64, 8)
            LN.constant_model(ctx, cls)
            u = np.stack(LN.evaluate(ctx, tgt, 0.4, 0, 8, u_ref, prec=L.LPM_F32C))
            st = T.stats(ctx)
            err[on, extra, cls] = O.rel_l2(u, ud)
            if on:
                assert st["far_stereo"] > 0.3 * st["far"]
                assert st["terms_stereo"] == st["far_stereo"] * (cls + extra + 1)
                assert st["terms"] == (st["far"] - st["far_stereo"]) * cls * (cls + 1) * (cls + 2) // 6
            else:
                assert st["far_stereo"] == 0
        print("\n" + ", ".join(f"stereo {k[0]} extra {k[1]} order {k[2]}: {v:.2e}" for k, v in err.items()))
        assert err[True, 0, 6] <= 1.05 * err[False, 0, 6]
        assert err[True, 0, 8] <= 1.05 * err[False, 0, 8]
        assert err[True, 2, 6] <= 1.01 * err[True, 0, 6]
    finally:
        _stereo(ctx, False)


def test_stereo_is_off_for_fp64_and_smoothing(ctx, mesh):
    """The stereographic form holds for delta = 0 on the unit sphere: with
    delta > 0, and in FP64, the option changes nothing."""
    m, src, g, tgt = _field(mesh, 4, 32, 6)
    u_ref = LN.u_ref_of(g)
    try:
        out = {}
        for on in (False, True):
            _stereo(ctx, on, 2)
            _gpu_tree(ctx, m, src, g, 32, 6)
            LN.constant_model(ctx, 5)
            out[on] = [np.stack(LN.evaluate(ctx, tgt, 0.6, 0, 6, u_ref, delta=0.01, prec=L.LPM_F32C)),
                       np.stack(LN.evaluate(ctx, tgt, 0.6, 0, 6, u_ref, prec=L.LPM_F64))]
            assert T.stats(ctx)["far_stereo"] == 0
        for a, b in zip(out[False], out[True]):
            assert np.array_equal(a, b)
        # moments made before the option was set: an error, not a silent fallback
        _stereo(ctx, False)
        _gpu_tree(ctx, m, src, g, 32, 6)
        _stereo(ctx, True)
        with pytest.raises(L.LpmError):
            LN.evaluate(ctx, tgt, 0.6, 0, 6, u_ref, prec=L.LPM_F32C)
        with pytest.raises(L.LpmError):
            ctx.set_option(L.LPM_OPT_STEREO_EXTRA_ORDER, 15)
        with pytest.raises(L.LpmError):
            ctx.set_option(L.LPM_OPT_STEREO, 3)
    finally:
        _stereo(ctx, False)


def test_stereo_only_moments(ctx, mesh):
    """LPM_OPT_STEREO 2 makes the Cartesian moments only up to order 3 (the
    model's features) beside stereographic moments of the requested order.  A
    constant model accepting every pair past theta_max = 0.4 at order 6 then
    leaves no Cartesian entry (a pair that neither projection takes descends),
    the result is no less accurate against the direct sum than with the full
    Cartesian moments (LPM_OPT_STEREO 1), and the evaluations that need the
    Cartesian series of order 6 refuse these moments."""
    m, src, g, tgt = _field(mesh, 5, 64, 8)
    ud = np.stack(O.bs_direct(tgt, src, g))
    u_ref = LN.u_ref_of(g)
    res = {}
    try:
        for mode in (1, 2):
            _stereo(ctx, True, 1, 2.0)
            ctx.set_option(L.LPM_OPT_STEREO, mode)
            _gpu_tree(ctx, m, src, g, 64, 8)
            LN.constant_model(ctx, 6)
            u = np.stack(LN.evaluate(ctx, tgt, 0.4, 0, 8, u_ref, prec=L.LPM_F32C))
            res[mode] = (O.rel_l2(u, ud), T.stats(ctx))
        (e1, s1), (e2, s2) = res[1], res[2]
        print(f"\nstereo 1: {e1:.2e} (far {s1['far']}, stereographic {s1['far_stereo']}, near {s1['near']}); "
              f"stereo 2: {e2:.2e} (far {s2['far']}, stereographic {s2['far_stereo']}, near {s2['near']})")
        assert s1["far"] > s1["far_stereo"] > 0.9 * s1["far"]
        assert s2["far"] == s2["far_stereo"] and s2["terms"] == 0
        assert s2["near"] >= s1["near"]
        assert e2 <= 1.05 * e1
        with pytest.raises(L.LpmError):
            T.evaluate_batch(ctx, tgt, 0.4, 6)
        with pytest.raises(L.LpmError):
            LN.evaluate(ctx, tgt, 0.4, 0, 6, u_ref, prec=L.LPM_F64)
        # orders up to 3 are still served by the Cartesian moments
        LN.evaluate(ctx, tgt, 0.4, 0, 3, u_ref, prec=L.LPM_F64)
    finally:
        _stereo(ctx, False)


def test_learned_errors(ctx, mesh):
    m, src, g, tgt = _field(mesh, 3, 32, 4)
    _gpu_tree(ctx, m, src, g, 32, 4)
    LN.constant_model(ctx, 2)
    with pytest.raises(L.LpmError):          # max_order above the moments' order
        LN.evaluate(ctx, tgt, 0.6, 0, 6, 1.0)
    with pytest.raises(L.LpmError):
        LN.evaluate(ctx, tgt, 0.6, -1, 4, 1.0)
    with pytest.raises(L.LpmError):
        LN.evaluate(ctx, tgt, 0.6, 0, 4, 0.0)
    with pytest.raises(L.LpmError):          # labels need order-12 moments
        LN.labels(ctx, tgt, 0.8, 1.0)


def test_learned_gain_over_fixed_order(ctx, mesh):
    """The trained model does less far- and near-field work than every
    fixed-(p, theta) batch treecode that is at least as accurate (cubed
    level 7, 8 vortices; work = 8.5 vector instructions per series term + 11
    per near pair, the kernels' instruction counts)."""
    m, src, g, _ = _field(mesh, 7, 64, 12)
    _gpu_tree(ctx, m, src, g, 64, 12)
    u_ref = LN.u_ref_of(g)
    ud = np.stack(ctx.direct_sum(src, src, g))   # the self pair contributes 0 (|den| <= 1e-13)
    ls = LN.labels(ctx, src, 0.85, u_ref, batch_stride=2, keep_rate=1.0, max_samples=1 << 21)
    fr = LN.fit(ctx, ls, 3e-7, 8, epochs=10, seed=3)
    assert fr.accuracy > 0.6
    work = lambda st: 8.5 * st["terms"] + 11.0 * st["near"]
    ul = LN.evaluate(ctx, src, 0.85, 0, 8, u_ref)
    err_l, work_l = O.rel_l2(np.stack(ul), ud), work(T.stats(ctx))
    assert err_l < 1e-5
    best = None
    for p in (4, 5, 6, 7, 8):
        for th in (0.25, 0.3, 0.35, 0.4, 0.5, 0.6, 0.7):
            ub = T.evaluate_batch(ctx, src, th, p, prec=L.LPM_F32C)
            if O.rel_l2(np.stack(ub), ud) <= err_l:
                w = work(T.stats(ctx))
                best = w if best is None else min(best, w)
    assert best is not None and work_l < 0.8 * best, (work_l, best)
