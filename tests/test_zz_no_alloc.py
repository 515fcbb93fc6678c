"""lpm_direct_sum_dev_noalloc: the direct sum on the library's static device
scratch (no allocation), called through the C ABI and checked against
lpm_direct_sum_dev on the same device arrays.  Three calls in a row on one
GPU share its scratch: a large one (cubed level 7, fewer splits than
lpm_direct_sum_dev), then two smaller ones (cubed level 5, icosahedral level
3) whose results must not depend on what the earlier calls left there.
Bars: FP64 bit for bit (same splits) or within 1e-12 (fewer splits); FP32C
bit for bit, or within 3e-6 (max difference over max value) where
lpm_direct_sum_dev sums copies sorted by position (65,536 or more particles)
and the FP32 round-off of each differs (each is within 1e-5 of the oracle);
FP64 within 1e-12 of the oracle.
"""
import numpy as np
import pytest
import torch

import lpmcxx_amd as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu

VEL, PSI = L.LPM_K_BIOT_SAVART_SPHERE, L.LPM_K_GREENS_SPHERE


def _run(ctx, fn, T, S, G, skip, prec):
    nt = T[0].numel()
    out = [torch.full((nt,), np.nan, dtype=torch.float64, device="cuda") for _ in range(4)]
    torch.cuda.synchronize()
    p = lambda t: None if t is None else t.data_ptr()
    ctx.check(fn(ctx.h, VEL | PSI, prec, 1.0, 0.0, *(p(t) for t in T), nt, *(p(s) for s in S), p(G), S[0].numel(),
                 p(skip), *(p(o) for o in out)))
    ctx.synchronize()
    return np.stack([o.cpu().numpy() for o in out])


@pytest.fixture(scope="module")
def own_ctx():
    c = L.Context(0)
    yield c
    c.close()


def _case(seed, level):
    m = L.Mesh(seed, level)
    tgt = m.targets()
    g = L.sphere_harmonic54(m.cx, m.cy, m.cz) * m.area
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return m, tgt, (m.cx, m.cy, m.cz), g, [dev(a) for a in tgt], [dev(a) for a in (m.cx, m.cy, m.cz)], dev(g), \
        dev(m.face_self_skip())


def test_noalloc_matches_dev_entry(own_ctx):
    ctx = own_ctx
    lib = ctx.lib
    for seed, level, exact in [(L.LPM_SEED_CUBED, 7, False), (L.LPM_SEED_CUBED, 5, True),
                               (L.LPM_SEED_ICOS_TRI, 3, True)]:
        m, tgt, src, g, T, S, G, SK = _case(seed, level)
        print(f"\nnoalloc direct sum: seed {seed} level {level}, {T[0].numel()} targets x {S[0].numel()} sources",
              flush=True)
        for prec, tol in [(L.LPM_F64, 1e-12), (L.LPM_F32C, 3e-6)]:
            a = _run(ctx, lib.lpm_direct_sum_dev_noalloc, T, S, G, SK, prec)
            b = _run(ctx, lib.lpm_direct_sum_dev, T, S, G, SK, prec)
            assert np.isfinite(a).all()
            d = np.abs(a - b).max(axis=1) / np.abs(b).max(axis=1)
            print(f"  prec {prec}: max rel difference to lpm_direct_sum_dev {d.max():.2e}")
            if exact:
                assert np.array_equal(a, b)
            else:
                assert d.max() <= tol
        if level <= 5:
            ur = np.stack(O.bs_direct(tgt, src, g, skip=m.face_self_skip()))
            pr = O.greens_direct(tgt, src, g, skip=m.face_self_skip())
            a = _run(ctx, lib.lpm_direct_sum_dev_noalloc, T, S, G, SK, L.LPM_F64)
            assert O.rel_max(a[:3], ur) <= 1e-12 and O.rel_max(a[3], pr) <= 1e-12


def test_noalloc_limits(own_ctx):
    ctx = own_ctx
    lib = ctx.lib
    n = 2_000_000                                  # 2M targets x 98,304 sources: beyond the 32 MiB scratch
    T = [torch.zeros(n, dtype=torch.float64, device="cuda") for _ in range(3)]
    T[2].fill_(1.0)
    S = [torch.zeros(98304, dtype=torch.float64, device="cuda") for _ in range(3)]
    S[0].fill_(1.0)
    G = torch.zeros(98304, dtype=torch.float64, device="cuda")
    out = [torch.empty(n, dtype=torch.float64, device="cuda") for _ in range(3)]
    torch.cuda.synchronize()
    st = lib.lpm_direct_sum_dev_noalloc(ctx.h, VEL, L.LPM_F64, 1.0, 0.0, *(t.data_ptr() for t in T), n,
                                        *(s.data_ptr() for s in S), G.data_ptr(), 98304, None,
                                        *(o.data_ptr() for o in out), None)
    assert st == -1 and b"scratch" in lib.lpm_last_error(ctx.h)
    # no sources: zero velocity, as lpm_direct_sum_dev
    o1 = [torch.full((5,), 7.0, dtype=torch.float64, device="cuda") for _ in range(3)]
    torch.cuda.synchronize()
    ctx.check(lib.lpm_direct_sum_dev_noalloc(ctx.h, VEL, L.LPM_F64, 1.0, 0.0, *(t.data_ptr() for t in T), 5,
                                             None, None, None, None, 0, None, *(o.data_ptr() for o in o1), None))
    ctx.synchronize()
    assert all(float(o.abs().max()) == 0.0 for o in o1)
    # off the sphere with LPM_F32C: the sphere check on the scratch words
    P = torch.tensor([[1.0, 0.0, 0.0], [0.0, 1.1, 0.0]], dtype=torch.float64, device="cuda").T.contiguous()
    g2 = torch.ones(2, dtype=torch.float64, device="cuda")
    o2 = [torch.empty(2, dtype=torch.float64, device="cuda") for _ in range(3)]
    torch.cuda.synchronize()
    st = lib.lpm_direct_sum_dev_noalloc(ctx.h, VEL, L.LPM_F32C, 1.0, 0.0, *(P[i].data_ptr() for i in range(3)), 2,
                                        *(P[i].data_ptr() for i in range(3)), g2.data_ptr(), 2, None,
                                        *(o.data_ptr() for o in o2), None)
    assert st == -1
