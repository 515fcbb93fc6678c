"""Benchmark of the spherical Biot-Savart sum on one MI355X.

Workload: cubed-sphere level 9, 1,572,864 leaf-face particles (targets =
sources = leaf faces, the face self pair excluded); zeta = sum of 8 Gaussian
vortices (centers uniform on the sphere, amplitudes U(-2, 2), widths U(4, 16),
numpy PCG64 seed 1234) minus its mean; gamma = zeta * area.  Inputs are
resident in HBM before timing starts.

One step = one velocity evaluation of all particles by the learned-order
FP32C treecode: tree build + moments + model-planned evaluation, the work an
RK4 stage repeats.  Accepted pairs whose plane geometry qualifies are
evaluated as stereographic Cauchy series (LPM_OPT_STEREO; --no-stereo keeps
the Cartesian series throughout).  value = N (N - 1) / step time (effective interactions/s),
valid only if the step's relative L2 error against the FP64 direct sum of
this run is <= 1e-5; otherwise the fixed-order FP32C batch treecode is the
step (the fallback), and if that misses too, the FP32C direct sum.

Measured in the same run:
  * FP64 direct sums (the error reference: the reference's pair formula with
    IEEE division; and the fused LPM_F64F form) and the FP32C direct sum, with
    the roofline of their pair kernels;
  * the plain FP64 GPU treecode (reference acceptance per target, Cartesian
    series, fixed p and theta, fused FP64 near field; the fastest such
    setting within 1e-5) with its FP64 roofline, and the fixed-order FP32C
    batch treecode;
  * training of the order model on GPU-generated labels (time reported, and
    the number of evaluations after which it has paid for itself);
  * the model trained on this field evaluated on held-out fields (another
    vortex seed, the Rossby-Haurwitz field) against models trained on those;
  * the pairs actually evaluated per second beside the effective rate;
  * RK4 steps/s with the learned treecode and with the plain FP64 treecode;
  * the roofline of the step's evaluation kernels from HIP events recorded by
    the library on its stream (lpm_last_kernel_ms);
  * a CPU baseline: the oracle (reference-order FP64 direct sum, OpenMP) on a
    bounded sample of targets.

--workload mesh prints the line of another path instead: GPU mesh generation
of the cubed sphere at --mesh-level (default 10), one step = one whole
generation, against the host mesh builder (see workload_mesh).
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

FP32_PEAK = 157.3   # TFLOP/s, MI355X FP32 vector (packed)
FP32_NONPACKED = 78.6
FP64_PEAK = 78.6    # TFLOP/s, MI355X FP64 vector
HBM_PEAK = 8000.0   # GB/s, MI355X HBM3E
DIRECT_FLOPS = 18   # velocity pair: d 3, |d|^2 5, q 2, rcp 1, w 1, acc 6
# stereographic far pair: z - z_c 2, |dz|^2 3, rcp 1, w 2, rho w 2, w S 6, sums 3
# = 19, of which the first complex term's 8 are counted with the terms
STEREO_PAIR_FLOPS = 11
PSI_NEAR_FLOPS = 12  # streamfunction pair (FP32): d 3, |d|^2 5, log 1, gamma * log 2, strength sum 1
METRIC = "Biot-Savart interactions/s + RK4 steps/s at 1.57M particles, rel L2 err ≤1e-5"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def nterms(p: int) -> int:
    return p * (p + 1) * (p + 2) // 6


def cpu_baseline(src, g, target_seconds: float):
    """Oracle (reference-order FP64, OpenMP over targets) on a sample of targets."""
    import numpy as np
    from oracle import oracle as O
    n = len(src[0])
    rng = np.random.default_rng(5)
    k = 16
    while True:
        idx = np.sort(rng.choice(n, k, replace=False))
        sub = tuple(t[idx] for t in src)
        t0 = time.perf_counter()
        O.bs_direct(sub, src, g, skip=idx)
        dt = time.perf_counter() - t0
        if dt >= target_seconds or k >= n:
            break
        k = min(n, int(k * max(2.0, 1.2 * target_seconds / max(dt, 1e-3))))
    return {"value": k * (n - 1) / dt, "unit": "interactions/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"{k} random targets x all {n} sources, FP64 reference-order oracle, {dt:.1f} s"}


def rk4_level10(ctx, args):
    """RK4 on the rotating sphere at cubed-sphere level 10 (6,291,456
    particles): zeta_rel = 2 alpha z + 30 cos(4 lambda) z (z^2 - 1)^2 (solid
    body plus the Rossby-Haurwitz wave), alpha = 1, Omega = 2 pi, dt = 0.01.
    The order model is trained on this field's own labels at t = 0; then
    `l10_steps` steps run with the learned FP32C treecode and with the plain
    FP64 treecode from the same start.  Reported: ms per step of each, the
    velocity error of both at t = 0 on 4096 random targets against the FP64
    direct sum, and the learned trajectory's distance from the plain one."""
    import numpy as np
    import torch
    import lpmcxx_amd as L
    from lpmcxx_amd import learned as LN
    from lpmcxx_amd import tree as T

    log(f"[bench] level-{args.l10_level} RK4: mesh")
    m = L.Mesh(L.LPM_SEED_CUBED, args.l10_level)
    n = m.n_leaves
    omega, alpha, dt = 2 * np.pi, 1.0, 0.01
    zrel = 2 * alpha * m.cz + L.sphere_harmonic54(m.cx, m.cy, m.cz)
    g = zrel * m.area
    u_ref = LN.u_ref_of(g)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    X, Y, Z, G = dev(m.cx), dev(m.cy), dev(m.cz), dev(g)
    ZA, A = dev(zrel + 2 * omega * m.cz), dev(m.area)

    ctx.set_option(L.LPM_OPT_BATCH_CELL_TARGETS, args.learned_cell)
    set_stereo(ctx, args, False)
    t0 = time.perf_counter()
    T.build_dev(ctx, X, Y, Z, m.max_radius, args.leaf)
    T.moments_dev(ctx, G, 12, 1)
    ls = LN.labels(ctx, (m.cx, m.cy, m.cz), args.theta_max, u_ref, batch_stride=args.l10_label_stride,
                   keep_rate=args.label_keep, frontier_levels=2, max_samples=3 << 20)
    fr = LN.fit(ctx, ls, args.l10_tau, args.max_order, epochs=args.epochs, batch=4096, lr=2e-3,
                stereo_extra=stereo_extra(args))
    t_fit = time.perf_counter() - t0
    log(f"[bench] level-{args.l10_level} labels {len(ls.ids)}, labels + training {t_fit:.1f} s")

    # velocity error at t = 0 on a sample of targets
    idx = np.sort(np.random.default_rng(7).choice(n, 4096, replace=False))
    I = dev(idx.astype(np.int64))
    S = [dev(a[idx]) for a in (m.cx, m.cy, m.cz)]
    ref = [torch.empty_like(S[0]) for _ in range(3)]
    ctx.direct_sum_dev(*S, X, Y, Z, G, ref, prec=L.LPM_F64)
    U = [torch.empty_like(X) for _ in range(3)]

    def sample_err():
        ctx.synchronize()
        d = torch.stack([u[I] for u in U]) - torch.stack(ref)
        return float(torch.linalg.norm(d) / torch.linalg.norm(torch.stack(ref)))

    set_stereo(ctx, args, True)
    T.build_dev(ctx, X, Y, Z, m.max_radius, args.leaf)
    T.moments_dev(ctx, G, args.max_order, 1)
    LN.evaluate_dev(ctx, X, Y, Z, U, args.theta_max, args.margin, args.max_order, u_ref, prec=L.LPM_F32C)
    err_l = sample_err()
    set_stereo(ctx, args, False)
    T.build_dev(ctx, X, Y, Z, m.max_radius, args.l10_base_leaf)
    T.moments_dev(ctx, G, args.l10_base_order, 1)
    T.evaluate_dev(ctx, X, Y, Z, U, args.l10_base_theta, prec=L.LPM_F64F, series=L.LPM_SERIES_CARTESIAN)
    err_b = sample_err()

    cfg_l = L.SolverConfig.default(method=L.LPM_METHOD_TREE_LEARNED, prec=L.LPM_F32C, order=args.max_order,
                                   theta=args.theta_max, leaf_max=args.leaf, margin=args.margin, u_ref=u_ref)
    cfg_b = L.SolverConfig.default(method=L.LPM_METHOD_TREE, prec=L.LPM_F64F, order=args.l10_base_order,
                                   theta=args.l10_base_theta, leaf_max=args.l10_base_leaf)

    def run(cfg):
        set_stereo(ctx, args, cfg.method == L.LPM_METHOD_TREE_LEARNED)
        warm = [X.clone(), Y.clone(), Z.clone()]
        ctx.rk4_step_dev(cfg, dt, omega, *warm, ZA, A)
        P = [X.clone(), Y.clone(), Z.clone()]
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(args.l10_steps):
            ctx.rk4_step_dev(cfg, dt, omega, *P, ZA, A)
        ctx.synchronize()
        return (time.perf_counter() - t) / args.l10_steps, P

    log(f"[bench] level-{args.l10_level} RK4: {args.l10_steps} learned steps, {args.l10_steps} plain FP64 steps")
    s_l, P_l = run(cfg_l)
    s_b, P_b = run(cfg_b)
    P0 = torch.stack([X, Y, Z])
    Pl, Pb = torch.stack(P_l), torch.stack(P_b)
    traj = float(torch.linalg.norm(Pl - Pb) / torch.linalg.norm(Pb - P0))
    log(f"[bench] level-{args.l10_level} RK4: learned {1e3 * s_l:.1f} ms/step (t=0 err {err_l:.2e}), plain FP64 "
        f"{1e3 * s_b:.1f} ms/step (t=0 err {err_b:.2e}), trajectory difference {traj:.2e}")
    return {"particles": n, "steps": args.l10_steps, "dt": dt, "omega": omega,
            "field": "solid body (alpha 1) + Rossby-Haurwitz wave 30 cos(4 lambda) z (z^2-1)^2",
            "learned_ms_per_step": 1e3 * s_l, "plain_fp64_ms_per_step": 1e3 * s_b, "speedup": s_b / s_l,
            "learned_rel_l2_err_t0": err_l, "plain_fp64_rel_l2_err_t0": err_b,
            "err_sample": "4096 random targets vs the FP64 direct sum over all sources",
            "trajectory_rel_diff": traj, "tau": args.l10_tau, "label_samples": int(len(ls.ids)),
            "labels_and_training_s": t_fit, "train_accuracy": fr.accuracy}


def generalization(ctx, args, m, X, Y, Z, steps: int) -> dict:
    """The order model trained on the seed-1234 vortex field, evaluated without
    retraining on held-out fields of the same mesh: 8 Gaussian vortices of
    seed 4321, and the solid body + Rossby-Haurwitz field 2 z + 30 cos(4
    lambda) z (z^2 - 1)^2.  Each is then evaluated again with a model
    trained on its own labels.  Reported per field and model: ms per step,
    rel-L2 error on 4096 random targets against the FP64 direct sum, near
    and far pairs; and the labels + training time of the own-field model."""
    import numpy as np
    import torch
    import lpmcxx_amd as L
    from lpmcxx_amd import learned as LN
    from lpmcxx_amd import tree as T

    om = LN.OrderModel(ctx)
    trained = om.get_params()
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    n = m.n_leaves
    idx = np.sort(np.random.default_rng(17).choice(n, 4096, replace=False))
    I = dev(idx.astype(np.int64))
    S = [dev(a[idx]) for a in (m.cx, m.cy, m.cz)]
    U = [torch.empty_like(X) for _ in range(3)]
    fields = {"vortices_seed_4321": L.random_vortices(m.cx, m.cy, m.cz, m.area, n=8, seed=4321),
              "rossby_haurwitz": 2.0 * m.cz + L.sphere_harmonic54(m.cx, m.cy, m.cz)}
    out = {}
    for name, zeta in fields.items():
        g = zeta * m.area
        G = dev(g)
        u_ref = LN.u_ref_of(g)
        ref = [torch.empty_like(S[0]) for _ in range(3)]
        ctx.direct_sum_dev(*S, X, Y, Z, G, ref, prec=L.LPM_F64)
        ctx.synchronize()  # the library's stream, not torch's
        R = torch.stack(ref)

        def step():
            T.build_dev(ctx, X, Y, Z, m.max_radius, args.leaf)
            T.moments_dev(ctx, G, args.max_order, 1)
            LN.evaluate_dev(ctx, X, Y, Z, U, args.theta_max, args.margin, args.max_order, u_ref, prec=L.LPM_F32C)

        def measure():
            set_stereo(ctx, args, True)
            step()
            ctx.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                step()
            ctx.synchronize()
            dt = (time.perf_counter() - t0) / steps
            err = float(torch.linalg.norm(torch.stack([u[I] for u in U]) - R) / torch.linalg.norm(R))
            st = T.stats(ctx)
            return {"ms_per_step": 1e3 * dt, "rel_l2_err": err, "near_pairs": st["near"], "far_pairs": st["far"]}

        om.set_params(*trained)
        res = {"seed_1234_model": measure()}
        set_stereo(ctx, args, False)
        t0 = time.perf_counter()
        T.build_dev(ctx, X, Y, Z, m.max_radius, args.leaf)
        T.moments_dev(ctx, G, 12, 1)
        ls = LN.labels(ctx, (m.cx, m.cy, m.cz), args.theta_max, u_ref, batch_stride=args.label_stride,
                       keep_rate=args.label_keep, frontier_levels=2, max_samples=3 << 20)
        LN.fit(ctx, ls, args.tau, args.max_order, epochs=args.epochs, batch=4096, lr=2e-3,
               stereo_extra=stereo_extra(args))
        t_fit = time.perf_counter() - t0
        res["own_field_model"] = measure()
        res["own_field_model"]["labels_and_training_s"] = t_fit
        a, b = res["seed_1234_model"], res["own_field_model"]
        log(f"[bench] generalization {name}: seed-1234 model {a['ms_per_step']:.2f} ms err {a['rel_l2_err']:.2e} "
            f"(near {a['near_pairs']}, far {a['far_pairs']}); own model {b['ms_per_step']:.2f} ms err "
            f"{b['rel_l2_err']:.2e} (near {b['near_pairs']}, far {b['far_pairs']}), trained in {t_fit:.2f} s")
        out[name] = res
    om.set_params(*trained)
    out["err_sample"] = "4096 random targets vs the FP64 direct sum over all sources"
    return out


def cart_far_flops(p: int) -> int:
    """FP64 flops of one accepted (target, node) pair of the per-target
    treecode at order p: R2 (6), its reciprocal (1), b_0 and the first three
    products (4), 2 d (3), then per term k != 0 the recurrence (an FMA per
    non-zero k_i, the first a multiply; a subtraction per k_i >= 2; the scaling
    by 1/R2) and three FMAs against the moments, and t x S (9)."""
    tot = 6 + 1 + 4 + 3 + 9
    for n in range(1, p):
        for k1 in range(n, -1, -1):
            for k2 in range(n - k1, -1, -1):
                k = (k1, k2, n - k1 - k2)
                n1, n2 = sum(1 for a in k if a >= 1), sum(1 for a in k if a >= 2)
                tot += 2 * n1 - 1 + n2 + 1 + 6
    return tot


def fp64_tree_roof(st: dict, ms: float, p: int) -> dict:
    """FP64 roofline of tree_eval<p> with the fused near field: DIRECT_FLOPS
    per near pair, cart_far_flops(p) per accepted node; the MAC tests of the
    descent are overhead and not counted."""
    flops = st["near"] * DIRECT_FLOPS + st["far"] * cart_far_flops(p)
    ach = flops / (ms / 1e3) / 1e12
    return {"bound": "valu", "kernel": f"tree_eval<{p}, fused FP64 near field>", "achieved": ach,
            "peak": FP64_PEAK, "unit": "TFLOP/s", "frac": ach / FP64_PEAK, "flops_per_launch": flops,
            "launch_ms": ms, "near_pairs": st["near"], "far_pairs": st["far"],
            "far_flops_per_pair": cart_far_flops(p), "near_flops_per_pair": DIRECT_FLOPS}


def flop_roof(kernel: str, ms: float, near: int, far: int, terms: int, near_flops: int, far_flops: int = 16,
              term_flops: int = 8, stereo_pairs: int = 0, stereo_terms: int = 0) -> dict:
    """FP32 VALU roofline of one treecode evaluation kernel from its pair and
    term counts and its launch time (HIP events on the library stream).
    Stereographic far pairs: STEREO_PAIR_FLOPS each plus 8 per complex term."""
    flops = (near * near_flops + far * far_flops + terms * term_flops + stereo_pairs * STEREO_PAIR_FLOPS
             + stereo_terms * 8)
    ach = flops / (ms / 1e3) / 1e12
    out = {"bound": "valu", "kernel": kernel, "achieved": ach, "peak": FP32_PEAK, "unit": "TFLOP/s",
           "frac": ach / FP32_PEAK, "frac_nonpacked": ach / FP32_NONPACKED, "traffic": None,
           "flops_per_launch": flops, "launch_ms": ms, "far_pairs": far, "near_pairs": near, "far_terms": terms}
    if stereo_pairs:
        out.update(stereo_far_pairs=stereo_pairs, stereo_terms=stereo_terms)
    return out


def lists_roof(kernel: str, ms: float, st: dict) -> dict:
    """Roofline of the low-order list kernel of a learned evaluation: near
    pairs, the Cartesian far pairs and terms of order <= 7, the stereographic
    far pairs and terms."""
    return flop_roof(kernel, ms, st["near"], st["far"] - st["far_high"] - st["far_stereo"],
                     st["terms"] - st["terms_high"], DIRECT_FLOPS, stereo_pairs=st["far_stereo"],
                     stereo_terms=st["terms_stereo"])


def set_stereo(ctx, args, on: bool) -> None:
    """The stereographic far field of the learned evaluation (LPM_OPT_STEREO);
    off for the other solvers, whose moments would otherwise include the
    stereographic ones."""
    import lpmcxx_amd as L
    ctx.set_option(L.LPM_OPT_STEREO, args.stereo_mode if (on and not args.no_stereo) else 0)
    ctx.set_option(L.LPM_OPT_STEREO_EXTRA_ORDER, args.stereo_extra)
    ctx.set_option(L.LPM_OPT_STEREO_RATIO, args.stereo_ratio)


def stereo_extra(args):
    return None if args.no_stereo else args.stereo_extra


def lists_kernel(args) -> str:
    """The low-order list kernel of the learned evaluation."""
    return "tree_eval_lists<7, true, true>" if args.no_stereo else "tree_eval_lists<7, true, true, true>"


def sample_rate(fn, n_src: int, n_tgt: int, target_seconds: float, seed: int = 5):
    """Time the oracle fn(idx) on a growing random sample of targets until it
    takes target_seconds; returns (pairs/s, sample size, seconds)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    k = 16
    while True:
        idx = np.sort(rng.choice(n_tgt, k, replace=False))
        t0 = time.perf_counter()
        fn(idx)
        dt = time.perf_counter() - t0
        if dt >= target_seconds or k >= n_tgt:
            return k * (n_src - 1) / dt, k, dt
        k = min(n_tgt, int(k * max(2.0, 1.2 * target_seconds / max(dt, 1e-3))))


def workload_psi(args) -> dict:
    """Streamfunction treecode (lpm_tree_*_psi) on the particles of the
    default workload (cubed L9, 8 Gaussian vortices): one step = tree build +
    streamfunction moments + FP32C batch evaluation of psi at every particle.
    Errors are measured against the FP64 direct streamfunction on 8,192 random targets
    (self pair excluded); the FP64 streamfunction treecode and the FP32C
    direct sum are timed beside it."""
    import numpy as np
    import torch
    import lpmcxx_amd as L
    from lpmcxx_amd import tree as T
    from oracle import oracle as O

    PSI_K = L.LPM_K_GREENS_SPHERE
    m = L.Mesh(L.LPM_SEED_CUBED, args.level)
    n = m.n_leaves
    g = L.random_vortices(m.cx, m.cy, m.cz, m.area, n=8, seed=1234) * m.area
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    X, Y, Z, G = dev(m.cx), dev(m.cy), dev(m.cz), dev(g)
    ctx = L.Context(0)
    idx = np.sort(np.random.default_rng(7).choice(n, 8192, replace=False))
    I = dev(idx.astype(np.int64))
    ref = torch.empty(len(idx), dtype=torch.float64, device="cuda")
    ctx.direct_sum_dev(*(dev(a[idx]) for a in (m.cx, m.cy, m.cz)), X, Y, Z, G, None, kernels=PSI_K,
                       prec=L.LPM_F64, skip=I, psi=ref)
    PSI = torch.empty_like(X)

    def err():
        ctx.synchronize()
        return float(torch.linalg.norm(PSI[I] - ref) / torch.linalg.norm(ref))

    def step(leaf, p, theta, prec):
        T.build_dev(ctx, X, Y, Z, m.max_radius, leaf)
        T.moments_psi_dev(ctx, G, p)
        T.evaluate_psi_dev(ctx, X, Y, Z, PSI, theta, p, prec=prec)

    cfg = (args.psi_leaf, args.psi_order, args.psi_theta, L.LPM_F32C)
    for _ in range(args.warmup):
        step(*cfg)
    ctx.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(*cfg)
    ctx.synchronize()
    torch.cuda.synchronize()
    step_s = (time.perf_counter() - t0) / args.steps
    e32 = err()
    kms = []
    for _ in range(args.steps):
        step(*cfg)
        kms.append(ctx.last_kernel_ms())
    st = T.stats(ctx)
    kms = float(np.mean(kms))
    roof = flop_roof(f"tree_eval_psi<{args.psi_order}, true>", kms, st["near"], st["far"], st["terms"],
                     PSI_NEAR_FLOPS)
    log(f"[bench] psi treecode: {1e3 * step_s:.2f} ms per step, rel-L2 {e32:.2e}, eval kernel {kms:.2f} ms")

    # the FP64 streamfunction treecode (its fastest setting within 1e-5) and the FP32C direct sum
    f64 = (64, 6, 0.4, L.LPM_F64)
    step(*f64)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(2):
        step(*f64)
    ctx.synchronize()
    s64 = (time.perf_counter() - t0) / 2
    e64 = err()
    ctx.direct_sum_dev(X, Y, Z, X, Y, Z, G, None, kernels=PSI_K, prec=L.LPM_F32C,
                       skip=torch.arange(n, device="cuda", dtype=torch.int64), psi=PSI)
    ds_ms = ctx.last_kernel_ms()
    eds = err()
    log(f"[bench] psi FP64 treecode {1e3 * s64:.1f} ms (rel-L2 {e64:.2e}); FP32C direct {ds_ms:.1f} ms "
        f"(rel-L2 {eds:.2e})")
    cpu = None
    if not args.no_cpu:
        src = (m.cx, m.cy, m.cz)
        rate, k, dt = sample_rate(lambda ix: O.greens_direct(tuple(a[ix] for a in src), src, g, skip=ix), n, n,
                                  args.cpu_seconds)
        cpu = {"value": rate, "unit": "interactions/s", "cores": O.num_threads(), "kind": "port",
               "sample": f"{k} random targets x all {n} sources, FP64 reference-order streamfunction oracle, "
                         f"{dt:.1f} s"}
    ctx.close()
    pairs = n * (n - 1)
    return {
        "metric": f"streamfunction interactions/s at {n / 1e6:.2f}M particles, rel L2 err <=1e-5",
        "value": pairs / step_s, "unit": "interactions/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * step_s, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32c+f64", "data": "synthetic",
        "config": {"workload": f"streamfunction of all {n} particles of a cubed-sphere level-{args.level} mesh, "
                               "8 Gaussian vortices",
                   "step": f"FP32C streamfunction treecode (build + moments + evaluation), leaf {args.psi_leaf}, "
                           f"order {args.psi_order}, theta {args.psi_theta}",
                   "rel_l2_err_vs_fp64_direct": e32, "err_sample": "8192 random targets, FP64 direct sum"},
        "roofline": roof, "cpu_baseline": cpu,
        "fp64_psi_tree": {"ms_per_step": 1e3 * s64, "rel_l2_err": e64, "leaf": 64, "order": 6, "theta": 0.4,
                          "speedup_of_step": s64 / step_s},
        "direct_f32c_psi": {"kernel_ms": ds_ms, "rel_l2_err": eds, "interactions_per_s": pairs / (ds_ms / 1e3)},
        "stats": st,
    }


def workload_bve(args) -> dict:
    """Barotropic vorticity RK4 of a meshed particle set (lpm_bve_step_dev):
    cubed L9, the leaf faces are the vortex particles (zeta_rel = solid body
    alpha = 1 + the Rossby-Haurwitz wave), the vertices passive tracers, Omega
    = 2 pi, dt = 0.01.  One step = one RK4 step of faces and vertices (4
    velocity evaluations at 3.15M targets from 1.57M sources) by the learned
    FP32C treecode, its order model trained on the faces' and vertices' own
    labels at t = 0.  The plain FP64 treecode steps from the same start."""
    import numpy as np
    import torch
    import lpmcxx_amd as L
    from lpmcxx_amd import learned as LN
    from lpmcxx_amd import tree as T
    from oracle import oracle as O

    m = L.Mesh(L.LPM_SEED_CUBED, args.level)
    nf, nv = m.n_leaves, m.n_vertices
    nt = nf + nv
    omega, alpha, dt = 2 * np.pi, 1.0, 0.01
    zrel = 2 * alpha * m.cz + L.sphere_harmonic54(m.cx, m.cy, m.cz)
    g = zrel * m.area
    u_ref = LN.u_ref_of(g)
    tx, ty, tz = (np.concatenate([a, b]) for a, b in ((m.cx, m.vx), (m.cy, m.vy), (m.cz, m.vz)))
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    X, Y, Z = dev(tx), dev(ty), dev(tz)
    ZA, A, G = dev(zrel + 2 * omega * m.cz), dev(m.area), dev(g)
    SX, SY, SZ = X[:nf], Y[:nf], Z[:nf]
    ctx = L.Context(0)
    ctx.set_option(L.LPM_OPT_BATCH_CELL_TARGETS, args.learned_cell)
    set_stereo(ctx, args, False)

    t0 = time.perf_counter()
    T.build_dev(ctx, SX, SY, SZ, m.max_radius, args.leaf)
    T.moments_dev(ctx, G, 12, 1)
    ls = LN.labels(ctx, (tx, ty, tz), args.theta_max, u_ref, batch_stride=args.label_stride,
                   keep_rate=args.label_keep, frontier_levels=2, max_samples=3 << 20)
    fr = LN.fit(ctx, ls, args.bve_tau, args.max_order, epochs=args.epochs, batch=4096, lr=2e-3,
                stereo_extra=stereo_extra(args))
    t_fit = time.perf_counter() - t0

    # velocity error at t = 0 on a sample of faces and vertices
    idx = np.sort(np.random.default_rng(7).choice(nt, 4096, replace=False))
    I = dev(idx.astype(np.int64))
    ref = [torch.empty(len(idx), dtype=torch.float64, device="cuda") for _ in range(3)]
    ctx.direct_sum_dev(*(dev(a[idx]) for a in (tx, ty, tz)), SX, SY, SZ, G, ref, prec=L.LPM_F64,
                       skip=dev(np.where(idx < nf, idx, -1).astype(np.int64)))
    U = [torch.empty_like(X) for _ in range(3)]

    def err():
        ctx.synchronize()
        d = torch.stack([u[I] for u in U]) - torch.stack(ref)
        return float(torch.linalg.norm(d) / torch.linalg.norm(torch.stack(ref)))

    set_stereo(ctx, args, True)
    T.build_dev(ctx, SX, SY, SZ, m.max_radius, args.leaf)
    T.moments_dev(ctx, G, args.max_order, 1)
    LN.evaluate_dev(ctx, X, Y, Z, U, args.theta_max, args.margin, args.max_order, u_ref, prec=L.LPM_F32C)
    err_l = err()
    set_stereo(ctx, args, False)
    T.build_dev(ctx, SX, SY, SZ, m.max_radius, args.base_leaf)
    T.moments_dev(ctx, G, args.base_order, 1)
    T.evaluate_dev(ctx, X, Y, Z, U, args.base_theta, prec=L.LPM_F64F, series=L.LPM_SERIES_CARTESIAN)
    err_b = err()

    cfg_l = L.SolverConfig.default(method=L.LPM_METHOD_TREE_LEARNED, prec=L.LPM_F32C, order=args.max_order,
                                   theta=args.theta_max, leaf_max=args.leaf, margin=args.margin, u_ref=u_ref)
    cfg_b = L.SolverConfig.default(method=L.LPM_METHOD_TREE, prec=L.LPM_F64F, order=args.base_order,
                                   theta=args.base_theta, leaf_max=args.base_leaf)

    def run(cfg, nsteps, warm):
        set_stereo(ctx, args, cfg.method == L.LPM_METHOD_TREE_LEARNED)
        P = [X.clone(), Y.clone(), Z.clone()]
        W = [X.clone(), Y.clone(), Z.clone()]
        for _ in range(warm):
            ctx.bve_step_dev(cfg, dt, omega, nf, *W, ZA, A)
        ctx.synchronize()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(nsteps):
            ctx.bve_step_dev(cfg, dt, omega, nf, *P, ZA, A)
        ctx.synchronize()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / nsteps, P

    log(f"[bench] BVE: {args.steps} learned steps, {args.steps} plain FP64 steps")
    s_l, P_l = run(cfg_l, args.steps, args.warmup)
    kms = ctx.last_kernel_ms()
    st = T.stats(ctx)
    roof = lists_roof(lists_kernel(args), kms, st)
    roof["launch_note"] = "last velocity evaluation (RK4 stage 4) of the last timed step"
    s_b, P_b = run(cfg_b, args.steps, 1)
    P0 = torch.stack([X, Y, Z])
    traj = float(torch.linalg.norm(torch.stack(P_l) - torch.stack(P_b)) / torch.linalg.norm(torch.stack(P_b) - P0))
    # one step from the same start by each solver against the FP32C direct sum's step
    one = {}
    for name, cfg in (("direct", L.SolverConfig.default(method=L.LPM_METHOD_DIRECT, prec=L.LPM_F32C)),
                      ("learned", cfg_l), ("plain", cfg_b)):
        _, P = run(cfg, 1, 0)
        one[name] = torch.stack(P)
    disp_err = lambda P: float(torch.linalg.norm(P - one["direct"]) / torch.linalg.norm(one["direct"] - P0))
    step_err_l, step_err_b = disp_err(one["learned"]), disp_err(one["plain"])
    log(f"[bench] BVE: learned {1e3 * s_l:.1f} ms/step (t=0 err {err_l:.2e}, one-step displacement err "
        f"{step_err_l:.2e}), plain FP64 {1e3 * s_b:.1f} ms/step (t=0 err {err_b:.2e}, one-step {step_err_b:.2e}), "
        f"trajectory difference {traj:.2e}")
    cpu = None
    if not args.no_cpu:
        src = (m.cx, m.cy, m.cz)
        tgt = (tx, ty, tz)
        rate, k, sec = sample_rate(
            lambda ix: O.bs_direct(tuple(a[ix] for a in tgt), src, g, skip=np.where(ix < nf, ix, -1)), nf, nt,
            args.cpu_seconds)
        pairs_per_step = 4 * (nt * nf - nf)
        cpu = {"value": rate / pairs_per_step, "unit": "steps/s", "cores": O.num_threads(),
               "kind": "port", "interactions_per_s": rate,
               "sample": f"{k} random targets of the {nt} x all {nf} sources, FP64 reference-order oracle "
                         f"direct sum, {sec:.1f} s; steps/s = interactions/s / (4 stages x {pairs_per_step // 4} "
                         f"pairs)"}
    ctx.close()
    return {
        "metric": f"BVE RK4 steps/s at {nf / 1e6:.2f}M vortex faces + {nv / 1e6:.2f}M tracer vertices, "
                  "rel L2 err <=1e-5",
        "value": 1.0 / s_l, "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * s_l, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32c+f64", "data": "synthetic",
        "config": {"workload": f"barotropic vorticity RK4 of a cubed-sphere level-{args.level} mesh: {nf} faces "
                               f"(vortex particles, solid body + Rossby-Haurwitz wave) and {nv} vertices (passive "
                               "tracers), Omega 2 pi, dt 0.01",
                   "step": f"learned-order FP32C treecode per stage, leaf {args.leaf}, theta_max {args.theta_max}, "
                           f"tau {args.bve_tau}, max order {args.max_order}",
                   "rel_l2_err_t0": err_l, "err_sample": "4096 random faces and vertices, FP64 direct sum",
                   "step_displacement_rel_l2_err": step_err_l,
                   "step_err_reference": "one RK4 step of all faces and vertices by the FP32C direct sum"},
        "roofline": roof, "cpu_baseline": cpu,
        "plain_fp64_tree": {"ms_per_step": 1e3 * s_b, "rel_l2_err_t0": err_b, "speedup": s_b / s_l,
                            "step_displacement_rel_l2_err": step_err_b},
        "trajectory_rel_diff_vs_plain_fp64": traj, "label_samples": int(len(ls.ids)),
        "labels_and_training_s": t_fit, "train_accuracy": fr.accuracy, "stats_last_stage": st,
    }


def workload_mesh(args) -> dict:
    """GPU mesh generation (lpm_mesh_generate_dev), cubed sphere: one step =
    the whole refinement from the seed to the leaf centroids and areas, into
    device arrays allocated before timing.  The CPU baseline is the host
    builder (liblpm's PolyMesh2d: the reference's sequential division loop)
    on the same level; its arrays double as a full-size parity check
    (vertices and centroids bit for bit, areas to a few ulp)."""
    import numpy as np
    import torch
    import lpmcxx_amd as L

    seed, level = L.LPM_SEED_CUBED, args.mesh_level
    nv, nf, nl = L.mesh_counts(seed, level)
    ctx = L.Context(0)
    out = ctx.mesh_generate_dev(seed, level)
    for _ in range(args.warmup):
        ctx.mesh_generate_dev(seed, level, out=out)
    ctx.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.mesh_generate_dev(seed, level, out=out)
    ctx.synchronize()
    torch.cuda.synchronize()
    step_s = (time.perf_counter() - t0) / args.steps
    leaf_ms = []
    for _ in range(args.steps):
        ctx.mesh_generate_dev(seed, level, out=out)
        leaf_ms.append(ctx.last_kernel_ms())
    leaf_ms = float(np.mean(leaf_ms))
    log(f"[bench] mesh level {level}: {1e3 * step_s:.3f} ms per generation, leaf kernel {leaf_ms:.3f} ms")

    # algorithmic HBM bytes of the leaf kernel: corner edge lists (4 ints per
    # leaf), edges read once (orig, dest, left: 3 ints), vertices read once
    # (3 doubles), centroid and area written (4 doubles)
    ne = nl + nv - 2
    alg_bytes = 16 * nl + 12 * ne + 24 * nv + 32 * nl
    roof = {"bound": "valu", "kernel": "mesh_leaves<4>", "launch_ms": leaf_ms, "algorithmic_bytes": alg_bytes,
            "hbm_achieved_GBps": alg_bytes / (leaf_ms / 1e3) / 1e9, "achieved": None, "peak": FP64_PEAK,
            "unit": "TFLOP/s", "frac": None, "traffic": None}
    roof["hbm_frac"] = roof["hbm_achieved_GBps"] / HBM_PEAK
    f = ROOT / "profiles" / "mesh_leaves_l10_counters.json"
    if level == 10 and f.exists():
        d = json.loads(f.read_text())
        flops = d["fp64_flops_per_launch"]
        roof.update({"achieved": flops / (leaf_ms / 1e3) / 1e12, "fp64_flops_per_launch": flops,
                     "flops_source": "executed FP64 VALU flops (2 FMA + ADD + MUL, x64 lanes) per launch from "
                                     "profiles/mesh_leaves_l10_counters.md",
                     "traffic": d.get("traffic_bytes_per_launch")})
        roof["frac"] = roof["achieved"] / FP64_PEAK

    cpu, parity = None, None
    if not args.no_cpu:
        log("[bench] host builder (CPU baseline) and full-size parity")
        t0 = time.perf_counter()
        h = L.Mesh(seed, level)
        dt = time.perf_counter() - t0
        cpu = {"value": nl / dt, "unit": "leaf faces/s", "cores": 1, "kind": "port",
               "sample": f"whole level-{level} cubed sphere by the host PolyMesh2d builder (liblpm, the "
                         f"reference's sequential division loop in C++), {dt:.2f} s"}
        g = {k: v.cpu().numpy() for k, v in out.items() if v is not None}
        parity = {"vertices_bit_exact": all(np.array_equal(g[k], getattr(h, k)) for k in ("vx", "vy", "vz")),
                  "centroids_bit_exact": all(np.array_equal(g[k], getattr(h, k)) for k in ("cx", "cy", "cz")),
                  "area_max_rel_diff": float(np.max(np.abs(g["area"] - h.area) / h.area))}
    ctx.close()
    return {
        "metric": f"GPU mesh generation: leaf faces/s (cubed sphere level {level})", "value": nl / step_s,
        "unit": "leaf faces/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * step_s, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"uniform refinement of the cubed-sphere seed to level {level}: {nv} vertices, "
                               f"{nl} leaf faces ({nf} faces with parents), radius 1",
                   "vertices": nv, "leaf_faces": nl},
        "roofline": roof, "cpu_baseline": cpu, "parity_vs_host_builder": parity,
    }


def measured_traffic(kernel: str, args) -> tuple:
    """HBM bytes per launch of `kernel` from the committed counter summary of
    the default level-9 learned step (FETCH_SIZE / WRITE_SIZE passes scaled by
    a 256 MiB copy); None for other workloads."""
    f = ROOT / "profiles" / "tree_eval_lists_l9_traffic.json"
    if not f.exists():
        return None, None
    d = json.loads(f.read_text())
    mine = {"level": args.level, "leaf": args.leaf, "theta_max": args.theta_max, "max_order": args.max_order,
            "cell": args.learned_cell, "tau": args.tau, "stereo_mode": 0 if args.no_stereo else args.stereo_mode,
            "stereo_extra": args.stereo_extra, "stereo_ratio": args.stereo_ratio}
    if d.get("kernel") != kernel or d.get("settings") != mine:
        return None, None
    return d["traffic_bytes_per_launch"], "profiles/tree_eval_lists_l9_traffic.md"


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 20 timed steps: a learned step is ~2 ms, so fewer leave the timing to a
    # handful of launches' jitter
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--level", type=int, default=9)
    # learned treecode
    ap.add_argument("--leaf", type=int, default=None, help="default 192 (stereographic), 128 (Cartesian)")
    ap.add_argument("--theta-max", type=float, default=None, help="default 0.92 (stereographic), 0.85 (Cartesian)")
    ap.add_argument("--tau", type=float, default=None,
                    help="per-pair error bound of the labels (default 2.5e-6 with the stereographic far field, "
                         "3e-7 without)")
    ap.add_argument("--no-stereo", action="store_true",
                    help="Cartesian far field only (no LPM_OPT_STEREO in the learned evaluation)")
    ap.add_argument("--stereo-mode", type=int, default=2, choices=[1, 2],
                    help="LPM_OPT_STEREO of the learned evaluation: 2 = stereographic far field only (Cartesian "
                         "moments of order 3 for the model's features), 1 = with the Cartesian series beside it")
    ap.add_argument("--stereo-extra", type=int, default=1, help="LPM_OPT_STEREO_EXTRA_ORDER")
    ap.add_argument("--stereo-ratio", type=float, default=2.0, help="LPM_OPT_STEREO_RATIO")
    ap.add_argument("--max-order", type=int, default=7)
    ap.add_argument("--margin", type=int, default=0)
    ap.add_argument("--learned-cell", type=int, default=1024, help="targets per batch cell (learned treecode)")
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--label-stride", type=int, default=16)
    ap.add_argument("--label-keep", type=float, default=0.3)
    # plain FP64 treecode (reference acceptance) and fixed-order batch treecode
    ap.add_argument("--base-order", type=int, default=6)
    ap.add_argument("--base-theta", type=float, default=0.35)
    ap.add_argument("--base-leaf", type=int, default=256)
    ap.add_argument("--batch-order", type=int, default=7)
    ap.add_argument("--batch-theta", type=float, default=0.4)
    ap.add_argument("--batch-leaf", type=int, default=256)
    ap.add_argument("--batch-cell", type=int, default=256, help="targets per batch cell (batch treecode)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-rk4", action="store_true")
    ap.add_argument("--no-generalization", action="store_true", help="skip the held-out field evaluation")
    ap.add_argument("--no-direct64", action="store_true", help="measure errors against the FP32C direct sum")
    ap.add_argument("--copy-probe", action="store_true", help="run a 256 MiB device copy (counter calibration)")
    ap.add_argument("--no-rk4-l10", action="store_true", help="skip the level-10 RK4 comparison")
    ap.add_argument("--l10-level", type=int, default=10)
    ap.add_argument("--l10-steps", type=int, default=10)
    ap.add_argument("--l10-tau", type=float, default=None, help="default 1.5e-6 (stereographic), 1.5e-7 (Cartesian)")
    ap.add_argument("--l10-label-stride", type=int, default=64)
    # the fastest plain FP64 setting within 1e-5 at level 10 on the Rossby-Haurwitz field (p 6, theta 0.35
    # gives 1.00e-5 there)
    ap.add_argument("--l10-base-order", type=int, default=7, help="plain FP64 treecode order at level 10")
    ap.add_argument("--l10-base-theta", type=float, default=0.4)
    ap.add_argument("--l10-base-leaf", type=int, default=256)
    ap.add_argument("--workload", choices=["learned", "mesh", "psi", "bve"], default="learned",
                    help="learned: the velocity step (default); mesh: GPU mesh generation; psi: the "
                         "streamfunction treecode; bve: the barotropic vorticity RK4 of faces and vertices")
    ap.add_argument("--mesh-level", type=int, default=10)
    ap.add_argument("--psi-leaf", type=int, default=256)
    ap.add_argument("--psi-order", type=int, default=6)
    ap.add_argument("--psi-theta", type=float, default=0.4)
    ap.add_argument("--bve-tau", type=float, default=None, help="default 2e-6 (stereographic), 2e-7 (Cartesian)")
    args = ap.parse_args()
    if args.theta_max is None:
        args.theta_max = 0.85 if args.no_stereo else 0.92
    if args.leaf is None:
        args.leaf = 128 if args.no_stereo else 192
    if args.tau is None:
        args.tau = 3e-7 if args.no_stereo else 2.5e-6
    if args.l10_tau is None:
        args.l10_tau = 1.5e-7 if args.no_stereo else 1.5e-6
    if args.bve_tau is None:
        args.bve_tau = 2e-7 if args.no_stereo else 2e-6
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1", "n_gpus": args.gpus}))
        return 2
    if args.workload != "learned":
        out = {"mesh": workload_mesh, "psi": workload_psi, "bve": workload_bve}[args.workload](args)
        print(json.dumps(out), flush=True)
        return 0

    import numpy as np
    import torch
    import lpmcxx_amd as L
    from lpmcxx_amd import learned as LN
    from lpmcxx_amd import tree as T

    log(f"[bench] cubed-sphere level {args.level} mesh")
    m = L.Mesh(L.LPM_SEED_CUBED, args.level)
    n = m.n_leaves
    zeta = L.random_vortices(m.cx, m.cy, m.cz, m.area, n=8, seed=1234)
    g = zeta * m.area
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    X, Y, Z, G = dev(m.cx), dev(m.cy), dev(m.cz), dev(g)
    ctx = L.Context(0)
    torch.cuda.synchronize()
    pairs = n * (n - 1)
    u_ref = LN.u_ref_of(g)

    if args.copy_probe:
        a = torch.empty(256 << 20, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log("[bench] copy probe: 256 MiB device-to-device")

    def wall(fn, reps=1):
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        ctx.synchronize()
        return (time.perf_counter() - t0) / reps

    def rel_l2(u, ref):
        return float(torch.linalg.norm(torch.stack(u) - torch.stack(ref)) / torch.linalg.norm(torch.stack(ref)))

    # --- direct sums
    U32 = [torch.empty_like(X) for _ in range(3)]
    f32c = lambda: ctx.direct_sum_dev(X, Y, Z, X, Y, Z, G, U32, prec=L.LPM_F32C)
    f32c()
    k32 = []
    for _ in range(max(1, min(args.steps, 3))):
        f32c()
        k32.append(ctx.last_kernel_ms())
    ms32 = float(np.mean(k32))
    if args.no_direct64:
        U64, ms64, ms64f, err64f = U32, None, None, None
    else:
        log("[bench] FP64 direct sums (IEEE-division form, fused form)")
        U64 = [torch.empty_like(X) for _ in range(3)]
        ctx.direct_sum_dev(X, Y, Z, X, Y, Z, G, U64, prec=L.LPM_F64)
        ms64 = ctx.last_kernel_ms()
        U64F = [torch.empty_like(X) for _ in range(3)]
        ctx.direct_sum_dev(X, Y, Z, X, Y, Z, G, U64F, prec=L.LPM_F64F)
        ms64f = ctx.last_kernel_ms()
        err64f = rel_l2(U64F, U64)
        del U64F
    err32 = rel_l2(U32, U64)
    log(f"[bench] direct: f64 {ms64} ms, f64f {ms64f} ms (rel-L2 vs f64 {err64f}), f32c {ms32:.1f} ms (pair "
        f"kernels), f32c rel-L2 {err32:.2e}")

    UT = [torch.empty_like(X) for _ in range(3)]

    # --- plain FP64 GPU treecode: reference acceptance per target, Cartesian
    # series, fused FP64 near field (the faster FP64 form); the IEEE-division
    # near field timed beside it
    def base_step(prec=L.LPM_F64F):
        T.build_dev(ctx, X, Y, Z, m.max_radius, args.base_leaf)
        T.moments_dev(ctx, G, args.base_order, 1)
        T.evaluate_dev(ctx, X, Y, Z, UT, args.base_theta, prec=prec, series=L.LPM_SERIES_CARTESIAN)

    base_step()
    base_s = wall(base_step, 3)
    base_eval_ms = ctx.last_kernel_ms()
    base_err = rel_l2(UT, U64)
    base_st = T.stats(ctx)
    base_roof = fp64_tree_roof(base_st, base_eval_ms, args.base_order)
    base_step(L.LPM_F64)
    base_ieee_s = wall(lambda: base_step(L.LPM_F64), 2)
    base_ieee_eval_ms = ctx.last_kernel_ms()
    log(f"[bench] plain FP64 treecode: {1e3 * base_s:.1f} ms/step (eval {base_eval_ms:.1f}, "
        f"{base_roof['frac']:.3f} of the FP64 peak), rel-L2 {base_err:.2e}; IEEE-division near field "
        f"{1e3 * base_ieee_s:.1f} ms/step (eval {base_ieee_eval_ms:.1f})")

    # --- fixed-order FP32C batch treecode (the fallback)
    def batch_step():
        ctx.set_option(L.LPM_OPT_BATCH_CELL_TARGETS, args.batch_cell)
        T.build_dev(ctx, X, Y, Z, m.max_radius, args.batch_leaf)
        T.moments_dev(ctx, G, args.batch_order, 1)
        T.evaluate_batch_dev(ctx, X, Y, Z, UT, args.batch_theta, args.batch_order, prec=L.LPM_F32C)

    batch_step()
    batch_s = wall(batch_step, 3)
    batch_eval_ms = ctx.last_kernel_ms()
    batch_err = rel_l2(UT, U64)
    batch_stats = T.stats(ctx)
    log(f"[bench] FP32C batch treecode: {1e3 * batch_s:.1f} ms/step (eval {batch_eval_ms:.1f}), rel-L2 {batch_err:.2e}")

    # --- learned order model: labels from the FP64 treecode, training on the GPU
    ctx.set_option(L.LPM_OPT_BATCH_CELL_TARGETS, args.learned_cell)
    set_stereo(ctx, args, False)
    t0 = time.perf_counter()
    T.build_dev(ctx, X, Y, Z, m.max_radius, args.leaf)
    T.moments_dev(ctx, G, 12, 1)
    ls = LN.labels(ctx, (m.cx, m.cy, m.cz), args.theta_max, u_ref, batch_stride=args.label_stride,
                   keep_rate=args.label_keep, frontier_levels=2, max_samples=3 << 20)
    t_label = time.perf_counter() - t0
    t0 = time.perf_counter()
    fr = LN.fit(ctx, ls, args.tau, args.max_order, epochs=args.epochs, batch=4096, lr=2e-3,
                stereo_extra=stereo_extra(args))
    t_train = time.perf_counter() - t0
    set_stereo(ctx, args, True)
    log(f"[bench] labels {len(ls.ids)} in {t_label:.1f} s, training {t_train:.1f} s, loss {fr.final_loss:.3f}, "
        f"accuracy {fr.accuracy:.3f}")

    def learned_step():
        T.build_dev(ctx, X, Y, Z, m.max_radius, args.leaf)
        T.moments_dev(ctx, G, args.max_order, 1)
        LN.evaluate_dev(ctx, X, Y, Z, UT, args.theta_max, args.margin, args.max_order, u_ref, prec=L.LPM_F32C)

    log(f"[bench] warmup {args.warmup}, {args.steps} timed learned-treecode step(s)")
    for _ in range(args.warmup):
        learned_step()
    ctx.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        learned_step()
    ctx.synchronize()
    torch.cuda.synchronize()
    learned_s = (time.perf_counter() - t0) / args.steps
    learned_err = rel_l2(UT, U64)
    st = T.stats(ctx)
    # duration of the list kernel from the library's HIP events, in steps of
    # their own (reading the events waits for the step, which the timed loop
    # above does not)
    eval_ms = []
    for _ in range(args.steps):
        learned_step()
        eval_ms.append(ctx.last_kernel_ms())
    learned_eval_ms = float(np.mean(eval_ms))
    nn, depth, nleaves = T.info(ctx)
    log(f"[bench] learned FP32C treecode: {1e3 * learned_s:.1f} ms/step (low-order list kernel "
        f"{learned_eval_ms:.2f} ms), "
        f"rel-L2 {learned_err:.2e}; far {st['far']} near {st['near']} terms {st['terms']} model {st['model']}")

    gen = None if args.no_generalization else generalization(ctx, args, m, X, Y, Z, args.steps)

    if learned_err <= 1e-5:
        step_name, step_s, step_err = "learned", learned_s, learned_err
    elif batch_err <= 1e-5:
        step_name, step_s, step_err = "batch", batch_s, batch_err
    else:
        step_name, step_s, step_err = "direct_f32c", ms32 / 1e3, err32
    value = pairs / step_s

    def payoff(cost_s, other_s, own_s):
        """evaluations after which labels + training cost less than the time saved"""
        return None if other_s <= own_s else cost_s / (other_s - own_s)

    # roofline of the step's dominant evaluation kernel (FP32 VALU work); for
    # the learned step that is the low-order list kernel, so the far pairs and
    # terms of the order 8..12 list kernel are left out
    def tree_roof(stats, ms, kernel):
        return lists_roof(kernel, ms, stats)

    direct_flops = pairs * DIRECT_FLOPS
    direct_roof = {"bound": "valu", "kernel": "ds_f32c", "achieved": direct_flops / (ms32 / 1e3) / 1e12,
                   "peak": FP32_PEAK, "unit": "TFLOP/s", "traffic": None, "flops_per_launch": direct_flops,
                   "launch_ms": ms32}
    direct_roof["frac"] = direct_roof["achieved"] / FP32_PEAK
    direct_roof["frac_nonpacked"] = direct_roof["achieved"] / FP32_NONPACKED
    roof = {"learned": tree_roof(st, learned_eval_ms, lists_kernel(args)),
            "batch": tree_roof(batch_stats, batch_eval_ms, "tree_eval_batch"),
            "direct_f32c": direct_roof}[step_name]
    if step_name == "learned":
        roof["traffic"], roof["traffic_source"] = measured_traffic(roof["kernel"], args)

    # --- RK4 steps/s (4 treecode evaluations per step), learned and plain FP64
    rk4 = None
    if not args.no_rk4:
        za = dev(zeta + 2 * (2 * np.pi) * m.cz)
        A = dev(m.area)

        def rk4_rate(cfg, reps=2):
            set_stereo(ctx, args, cfg.method == L.LPM_METHOD_TREE_LEARNED)
            Xs, Ys, Zs = X.clone(), Y.clone(), Z.clone()
            ctx.rk4_step_dev(cfg, 0.01, 2 * np.pi, Xs, Ys, Zs, za, A)
            return 1.0 / wall(lambda: ctx.rk4_step_dev(cfg, 0.01, 2 * np.pi, Xs, Ys, Zs, za, A), reps)

        log("[bench] RK4 steps")
        cfg_l = L.SolverConfig.default(method=L.LPM_METHOD_TREE_LEARNED, prec=L.LPM_F32C, order=args.max_order,
                                       theta=args.theta_max, leaf_max=args.leaf, margin=args.margin, u_ref=u_ref)
        cfg_b = L.SolverConfig.default(method=L.LPM_METHOD_TREE, prec=L.LPM_F64F, order=args.base_order,
                                       theta=args.base_theta, leaf_max=args.base_leaf)
        rl, rb = rk4_rate(cfg_l), rk4_rate(cfg_b)

        # one step from the same start by each solver against the FP32C direct sum's step
        def one_step(cfg):
            set_stereo(ctx, args, cfg.method == L.LPM_METHOD_TREE_LEARNED)
            P = [X.clone(), Y.clone(), Z.clone()]
            ctx.rk4_step_dev(cfg, 0.01, 2 * np.pi, *P, za, A)
            ctx.synchronize()
            return torch.stack(P)

        P0 = torch.stack([X, Y, Z])
        Pd = one_step(L.SolverConfig.default(method=L.LPM_METHOD_DIRECT, prec=L.LPM_F32C))
        disp_err = lambda P: float(torch.linalg.norm(P - Pd) / torch.linalg.norm(Pd - P0))
        el, eb = disp_err(one_step(cfg_l)), disp_err(one_step(cfg_b))
        log(f"[bench] RK4 one-step displacement error vs the FP32C direct sum: learned {el:.2e}, plain FP64 {eb:.2e}")
        rk4 = {"rk4_steps_per_s": rl, "rk4_ms_per_step": 1e3 / rl, "rk4_plain_fp64_steps_per_s": rb,
               "rk4_speedup_vs_plain_fp64": rl / rb, "rk4_step_displacement_rel_l2_err": el,
               "rk4_plain_fp64_step_displacement_rel_l2_err": eb,
               "rk4_solver": f"learned FP32C treecode (theta_max {args.theta_max}, max order {args.max_order}, "
                             f"leaf {args.leaf}), dt 0.01, Omega 2 pi, FP64 update and renormalization"}
        log(f"[bench] RK4: learned {1e3 / rl:.1f} ms/step, plain FP64 treecode {1e3 / rb:.1f} ms/step")

    rk4_l10 = None
    if not args.no_rk4_l10:
        rk4_l10 = rk4_level10(ctx, args)

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline")
        cpu = cpu_baseline((m.cx, m.cy, m.cz), g, args.cpu_seconds)

    out = {
        "metric": METRIC, "value": value, "unit": "interactions/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * step_s, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32c+f64", "data": "synthetic",
        "config": {"workload": f"velocity of all {n} particles of a cubed-sphere level-{args.level} mesh, "
                               "8 Gaussian vortices", "particles": n, "pairs_per_step": pairs,
                   "step": {"learned": f"learned-order FP32C treecode (build + moments + evaluation), leaf "
                                       f"{args.leaf}, theta_max {args.theta_max}, tau {args.tau}, max order "
                                       f"{args.max_order}, margin {args.margin}, far field: "
                                       + ("Cartesian" if args.no_stereo else
                                          f"stereographic where the plane ratio <= {args.stereo_ratio} x the "
                                          f"sphere ratio (order p + {args.stereo_extra}), else Cartesian"),
                            "batch": f"fixed-order FP32C batch treecode p={args.batch_order} theta="
                                     f"{args.batch_theta} leaf {args.batch_leaf}",
                            "direct_f32c": "FP32C direct sum"}[step_name],
                   "rel_l2_err_vs_fp64_direct": step_err},
        "roofline": roof,
        "cpu_baseline": cpu,
        "pairs_evaluated_per_step": st["near"] + st["far"],
        "pairs_evaluated_per_s": (st["near"] + st["far"]) / learned_s,
        "learned": {"ms_per_step": 1e3 * learned_s, "low_list_kernel_ms": learned_eval_ms, "rel_l2_err": learned_err,
                    "speedup_vs_plain_fp64_tree": base_s / learned_s, "label_samples": int(len(ls.ids)),
                    "pairs_evaluated_per_s": (st["near"] + st["far"]) / learned_s,
                    "effective_interactions_per_s": pairs / learned_s,
                    "labels_and_training_s": t_label + t_train,
                    "payoff_evaluations_vs_batch_f32c_tree": payoff(t_label + t_train, batch_s, learned_s),
                    "payoff_evaluations_vs_plain_fp64_tree": payoff(t_label + t_train, base_s, learned_s),
                    "generalization": gen,
                    "label_s": t_label, "train_s": t_train, "train_loss": fr.final_loss,
                    "train_accuracy": fr.accuracy, "stats": st, "nodes": nn, "depth": depth, "leaves": nleaves},
        "plain_fp64_tree": {"form": "LPM_F64F (fused FP64 near field), Cartesian series",
                            "ms_per_step": 1e3 * base_s, "eval_ms": base_eval_ms, "rel_l2_err": base_err,
                            "order": args.base_order, "theta": args.base_theta, "leaf": args.base_leaf,
                            "effective_interactions_per_s": pairs / base_s, "roofline": base_roof,
                            "ieee_form_ms_per_step": 1e3 * base_ieee_s, "ieee_form_eval_ms": base_ieee_eval_ms},
        "batch_f32c_tree": {"ms_per_step": 1e3 * batch_s, "eval_ms": batch_eval_ms, "rel_l2_err": batch_err,
                            "order": args.batch_order, "theta": args.batch_theta, "leaf": args.batch_leaf,
                            "speedup_vs_plain_fp64_tree": base_s / batch_s, "stats": batch_stats},
        "direct_f32c": {"interactions_per_s": pairs / (ms32 / 1e3), "kernel_ms": ms32, "rel_l2_err": err32,
                        "roofline": direct_roof},
        "direct_f64": None if ms64 is None else
        {"kernel": "ds_f64 (LPM_F64F: fused FP64)", "interactions_per_s": pairs / (ms64f / 1e3),
         "kernel_ms": ms64f, "frac_of_fp64_peak": direct_flops / (ms64f / 1e3) / (FP64_PEAK * 1e12),
         "rel_l2_vs_ieee_form": err64f, "ieee_form_kernel_ms": ms64,
         "ieee_form_frac_of_fp64_peak": direct_flops / (ms64 / 1e3) / (FP64_PEAK * 1e12)},
    }
    if rk4:
        out.update(rk4)
    if rk4_l10:
        out["rk4_level10"] = rk4_l10
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
