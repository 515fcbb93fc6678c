"""GPU parity of the direct sum (liblpm_gpu, include/lpm_gpu.h) against the oracle.

Bars: FP64 (LPM_F64 and the fused LPM_F64F) relative L2 and max error
<= 1e-12 against the oracle's reference-order FP64 sum; FP32C relative L2 <= 1e-5 (velocity) with the
measured error printed.  Inputs are the mesh particle sets of the workloads
(icosahedral level 4, cubed-sphere levels 5 and 7) plus ragged, empty,
smoothed and scaled cases.
"""
import numpy as np
import pytest

import lpmcxx_amd as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu

VEL, PSI = L.LPM_K_BIOT_SAVART_SPHERE, L.LPM_K_GREENS_SPHERE


def _setup(m, field="rh"):
    if field == "rh":
        zeta = L.sphere_harmonic54(m.cx, m.cy, m.cz)
    else:
        zeta = L.gaussian_vortex(m.cx, m.cy, m.cz, m.area)
    return m.targets(), (m.cx, m.cy, m.cz), zeta * m.area, m.face_self_skip()


@pytest.mark.parametrize("seed,level", [(L.LPM_SEED_ICOS_TRI, 4), (L.LPM_SEED_CUBED, 5)])
def test_full_mesh_parity(ctx, mesh, seed, level):
    m = mesh(seed, level)
    tgt, src, g, skip = _setup(m)
    u_ref = O.bs_direct(tgt, src, g, skip=skip)
    p_ref = O.greens_direct(tgt, src, g, skip=skip)
    for prec in (L.LPM_F64, L.LPM_F64F):
        (u64, p64) = ctx.direct_sum(tgt, src, g, kernels=VEL | PSI, prec=prec, skip=skip)
        assert O.rel_l2(u64, u_ref) <= 1e-12 and O.rel_max(u64, u_ref) <= 1e-12, prec
        assert O.rel_l2(p64, p_ref) <= 1e-12 and O.rel_max(p64, p_ref) <= 1e-12, prec
    (u32, p32) = ctx.direct_sum(tgt, src, g, kernels=VEL | PSI, prec=L.LPM_F32C, skip=skip)
    e_u, e_p = O.rel_l2(u32, u_ref), O.rel_l2(p32, p_ref)
    print(f"\nseed={seed} level={level} f32c rel-L2 velocity {e_u:.3e} potential {e_p:.3e} "
          f"max {O.rel_max(u32, u_ref):.3e}")
    assert e_u <= 1e-5 and e_p <= 1e-5
    # velocity-only and potential-only launches agree with the fused one
    u_only = ctx.direct_sum(tgt, src, g, kernels=VEL, prec=L.LPM_F32C)
    p_only = ctx.direct_sum(tgt, src, g, kernels=PSI, prec=L.LPM_F32C, skip=skip)
    assert O.rel_l2(u_only, u32) < 1e-12 and O.rel_l2(p_only, p32) < 1e-12


def test_cubed_level7_sampled_targets(ctx, mesh):
    """Full GPU sum at 98,306 + 98,304 targets, velocity and streamfunction;
    oracle on 2,048 sampled targets."""
    m = mesh(L.LPM_SEED_CUBED, 7)
    tgt, src, g, skip = _setup(m)
    u64, p64 = ctx.direct_sum(tgt, src, g, kernels=VEL | PSI, prec=L.LPM_F64, skip=skip)
    uf = ctx.direct_sum(tgt, src, g, kernels=VEL, prec=L.LPM_F64F)
    u32, p32 = ctx.direct_sum(tgt, src, g, kernels=VEL | PSI, prec=L.LPM_F32C, skip=skip)
    idx = np.sort(np.random.default_rng(11).choice(len(tgt[0]), 2048, replace=False))
    sub = tuple(t[idx] for t in tgt)
    u_ref = O.bs_direct(sub, src, g, skip=skip[idx])
    p_ref = O.greens_direct(sub, src, g, skip=skip[idx])
    u64s = tuple(u[idx] for u in u64)
    u32s = tuple(u[idx] for u in u32)
    assert O.rel_l2(u64s, u_ref) <= 1e-12 and O.rel_max(u64s, u_ref) <= 1e-12
    ufs = tuple(u[idx] for u in uf)
    print(f"\ncubed L7 f64f rel-L2 {O.rel_l2(ufs, u_ref):.3e} max {O.rel_max(ufs, u_ref):.3e}; whole mesh "
          f"f64f vs f64 rel-L2 {O.rel_l2(uf, u64):.3e}")
    assert O.rel_l2(ufs, u_ref) <= 1e-12 and O.rel_max(ufs, u_ref) <= 1e-12
    assert O.rel_l2(uf, u64) <= 1e-13
    assert O.rel_l2(p64[idx], p_ref) <= 1e-12 and O.rel_max(p64[idx], p_ref) <= 1e-12
    e, ep = O.rel_l2(u32s, u_ref), O.rel_l2(p32[idx], p_ref)
    print(f"\ncubed L7 f32c rel-L2 velocity {e:.3e} (max {O.rel_max(u32s, u_ref):.3e}), "
          f"streamfunction {ep:.3e} (max {O.rel_max(p32[idx], p_ref):.3e})")
    assert e <= 1e-5 and ep <= 1e-5
    # the analytic Rossby-Haurwitz velocity is approached (first-order method),
    # and psi = zeta / 30 for zeta = Y_5^4 (tests/LpmPoissonSolverTests.cpp:290-301)
    assert O.rel_l2(u64, L.rh54_velocity(*tgt)) < 2e-3
    zeta_t = L.sphere_harmonic54(*tgt)
    assert np.linalg.norm(p64 - zeta_t / 30.0) / np.linalg.norm(zeta_t / 30.0) < 5e-3  # 3.1e-2 at L5, 9.0e-3 at L6


@pytest.mark.parametrize("nt,ns", [(1, 1), (1, 257), (300, 1), (65, 513), (1000, 255), (4097, 3001)])
def test_ragged_sizes(ctx, nt, ns):
    rng = np.random.default_rng(nt * 7 + ns)
    t = rng.normal(size=(3, nt)); t /= np.linalg.norm(t, axis=0)
    s = rng.normal(size=(3, ns)); s /= np.linalg.norm(s, axis=0)
    g = rng.normal(size=ns)
    ref = O.bs_direct(tuple(t), tuple(s), g)
    pref = O.greens_direct(tuple(t), tuple(s), g)
    for prec in (L.LPM_F64, L.LPM_F64F):
        u, p = ctx.direct_sum(tuple(t), tuple(s), g, kernels=VEL | PSI, prec=prec)
        assert O.rel_l2(u, ref) <= 1e-12 and O.rel_l2(p, pref) <= 1e-12, prec
    u, p = ctx.direct_sum(tuple(t), tuple(s), g, kernels=VEL | PSI, prec=L.LPM_F32C)
    assert O.rel_l2(u, ref) <= 1e-5 and O.rel_l2(p, pref) <= 1e-5


def test_smoothing_and_radius(ctx, mesh):
    m = mesh(L.LPM_SEED_ICOS_TRI, 3)
    tgt, src, g, skip = _setup(m, "gauss")
    for R, delta in [(1.0, 0.05), (2.0, 0.0), (6371.22, 10.0)]:
        t = tuple(R * a for a in tgt)
        s = tuple(R * a for a in src)
        ref = O.bs_direct(t, s, g, radius=R, delta=delta)
        for prec in (L.LPM_F64, L.LPM_F64F):
            u = ctx.direct_sum(t, s, g, prec=prec, radius=R, delta=delta)
            assert O.rel_l2(u, ref) <= 1e-12, (R, delta, prec)
        u = ctx.direct_sum(t, s, g, prec=L.LPM_F32C, radius=R, delta=delta)
        assert O.rel_l2(u, ref) <= 1e-5, (R, delta)


def test_f64f_off_sphere_and_smoothed(ctx):
    """LPM_F64F keeps the reference's general formula: points off the sphere
    (|x| in [0.3, 0.9], so R^2 - x.y >= 0.19) and a smoothing length, against
    the oracle; and the zero rule with duplicated sources and delta = 0."""
    rng = np.random.default_rng(21)
    t = rng.normal(size=(3, 700)); t *= rng.uniform(0.3, 0.9, 700) / np.linalg.norm(t, axis=0)
    s = rng.normal(size=(3, 1100)); s *= rng.uniform(0.3, 0.9, 1100) / np.linalg.norm(s, axis=0)
    g = rng.normal(size=1100)
    for delta in (0.0, 0.1):
        ref = O.bs_direct(tuple(t), tuple(s), g, delta=delta)
        u = ctx.direct_sum(tuple(t), tuple(s), g, prec=L.LPM_F64F, delta=delta)
        assert O.rel_l2(u, ref) <= 1e-12 and O.rel_max(u, ref) <= 1e-12, delta
    sd = np.concatenate([s, s[:, :50]], axis=1)
    gd = np.concatenate([g, g[:50]])
    ref = O.bs_direct(tuple(sd), tuple(sd), gd)
    u = ctx.direct_sum(tuple(sd), tuple(sd), gd, prec=L.LPM_F64F)
    assert all(np.isfinite(a).all() for a in u) and O.rel_l2(u, ref) <= 1e-12


def test_empty_inputs(ctx):
    one = (np.array([1.0]), np.array([0.0]), np.array([0.0]))
    u = ctx.direct_sum(one, (np.empty(0),) * 3, np.empty(0))
    assert all(np.array_equal(a, [0.0]) for a in u)
    u = ctx.direct_sum((np.empty(0),) * 3, one, np.array([1.0]))
    assert all(a.size == 0 for a in u)


def test_coincident_sources_give_zero_velocity(ctx):
    # every target coincides with one source (the face self pair) and the
    # reference's |den| <= 1e-13 rule zeroes it; duplicates do not produce NaN
    rng = np.random.default_rng(3)
    s = rng.normal(size=(3, 100)); s /= np.linalg.norm(s, axis=0)
    g = rng.normal(size=100)
    ref = O.bs_direct(tuple(s), tuple(s), g)
    for prec, tol in [(L.LPM_F64, 1e-12), (L.LPM_F64F, 1e-12), (L.LPM_F32C, 1e-5)]:
        u = ctx.direct_sum(tuple(s), tuple(s), g, prec=prec)
        assert all(np.isfinite(a).all() for a in u)
        assert O.rel_l2(u, ref) <= tol


def test_deterministic(ctx, mesh):
    m = mesh(L.LPM_SEED_CUBED, 4)
    tgt, src, g, skip = _setup(m)
    for prec in (L.LPM_F64, L.LPM_F64F, L.LPM_F32C):
        a = ctx.direct_sum(tgt, src, g, prec=prec, kernels=VEL | PSI, skip=skip)
        b = ctx.direct_sum(tgt, src, g, prec=prec, kernels=VEL | PSI, skip=skip)
        assert all(np.array_equal(x, y) for x, y in zip(a[0], b[0])) and np.array_equal(a[1], b[1])


def test_device_pointer_entry(ctx, mesh):
    import torch
    m = mesh(L.LPM_SEED_CUBED, 4)
    tgt, src, g, skip = _setup(m)
    host = ctx.direct_sum(tgt, src, g, prec=L.LPM_F32C)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    T = [dev(a) for a in tgt]
    S = [dev(a) for a in src]
    G = dev(g)
    out = [torch.empty_like(T[0]) for _ in range(3)]
    torch.cuda.synchronize()
    ctx.direct_sum_dev(*T, *S, G, out, prec=L.LPM_F32C)
    ctx.synchronize()
    for a, b in zip(out, host):
        assert np.array_equal(a.cpu().numpy(), b)


def test_argument_errors(ctx):
    one = (np.array([1.0]), np.array([0.0]), np.array([0.0]))
    with pytest.raises(L.LpmError) as e:
        ctx.direct_sum(one, one, np.array([1.0]), kernels=4)
    assert e.value.status == -7  # LPM_ERR_KERNEL, the reference's "invalid kernel"
    with pytest.raises(L.LpmError) as e:
        ctx.direct_sum(one, one, np.array([1.0]), prec=9)
    assert e.value.status == -1
    with pytest.raises(L.LpmError) as e:
        ctx.direct_sum(one, one, np.array([1.0]), radius=-1.0)
    assert e.value.status == -1
    lib = L.gpu_lib()
    st = lib.lpm_direct_sum(ctx.h, VEL, 0, 1.0, 0.0, None, None, None, 5, None, None, None, None, 5,
                            None, None, None, None, None)
    assert st == -2


def test_f32c_shuffled_order(ctx, mesh):
    """FP32C on the cubed level 7 particles in a random order: lpm_direct_sum
    sorts targets and sources by position, so the result equals the mesh-order
    sum to FP32C round-off, the self skip follows the shuffled sources, and the
    pair kernel runs at the mesh-order speed (no tile falls back to FP64)."""
    m = mesh(L.LPM_SEED_CUBED, 7)
    tgt, src, g, skip = _setup(m)
    rng = np.random.default_rng(5)
    pt, ps = rng.permutation(len(tgt[0])), rng.permutation(len(src[0]))
    inv_ps = np.empty_like(ps)
    inv_ps[ps] = np.arange(len(ps))
    tgt_s = tuple(t[pt] for t in tgt)
    src_s = tuple(s[ps] for s in src)
    sk = skip[pt]
    skip_s = np.where(sk >= 0, inv_ps[np.maximum(sk, 0)], -1)
    u0, p0 = ctx.direct_sum(tgt, src, g, kernels=VEL | PSI, prec=L.LPM_F32C, skip=skip)
    ms0 = ctx.last_kernel_ms()
    u1, p1 = ctx.direct_sum(tgt_s, src_s, g[ps], kernels=VEL | PSI, prec=L.LPM_F32C, skip=skip_s)
    ms1 = ctx.last_kernel_ms()
    back = np.empty_like(pt)
    back[pt] = np.arange(len(pt))
    u1 = tuple(u[back] for u in u1)
    p1 = p1[back]
    e_u, e_p = O.rel_l2(u1, u0), O.rel_l2(p1, p0)
    print(f"\ncubed L7 f32c shuffled vs mesh order: rel-L2 velocity {e_u:.2e} streamfunction {e_p:.2e}; "
          f"pair kernel {ms1:.3f} ms shuffled, {ms0:.3f} ms in mesh order")
    assert e_u <= 1e-6 and e_p <= 1e-6
    idx = np.sort(rng.choice(len(tgt[0]), 512, replace=False))
    u_ref = O.bs_direct(tuple(t[idx] for t in tgt), src, g, skip=skip[idx])
    assert O.rel_l2(tuple(u[idx] for u in u1), u_ref) <= 1e-5
    assert ms1 <= 1.5 * ms0 + 0.05
