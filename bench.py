"""Benchmark of the spherical Biot-Savart sum on one MI355X.

Workload: cubed-sphere level 9 -- 1,572,864 leaf-face particles (targets =
sources = leaf faces, the face self pair excluded); zeta = sum of 8 seeded
Gaussian vortices minus its mean; gamma = zeta * area.  Inputs are resident
in HBM before timing starts.

One step = one velocity evaluation of all particles by the treecode with the
FP32C near field (tree build + moments + evaluation, the work a time step
repeats at every stage), value = N (N - 1) / step time ("effective"
interactions/s), valid only if its relative L2 error against the FP64 direct
sum of the same run is <= 1e-5 (otherwise the FP32C direct sum is the step).

Also reported, each measured in this run:
  * FP64 direct sum (the error judge) and FP32C direct sum rates and errors,
    with the roofline of their pair kernels (ds_f64 / ds_f32c);
  * the plain FP64 treecode (same tree, FP64 near field) time and error;
  * RK4 steps/s at this size with the treecode velocity;
  * the roofline of the step's dominant kernel (tree_eval) from HIP events on
    the library's stream and its algorithmic flop count (near pairs x 18 +
    far pairs x (8 C(p+2,3) + 16));
  * a CPU baseline: the oracle (reference-order FP64 direct sum, OpenMP) on a
    bounded sample of targets.
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
FP64_PEAK = 78.6    # TFLOP/s, MI355X FP64 vector
DIRECT_FLOPS = 18   # velocity pair: d 3, |d|^2 5, q 2, rcp 1, w 1, acc 6
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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--level", type=int, default=9)
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--order", type=int, default=10)
    ap.add_argument("--leaf", type=int, default=128)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-rk4", action="store_true")
    ap.add_argument("--copy-probe", action="store_true", help="run a 256 MiB device copy (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1", "n_gpus": args.gpus}))
        return 2

    import numpy as np
    import torch
    import lpmcxx_amd as L
    from lpmcxx_amd import tree as T

    log(f"[bench] cubed-sphere level {args.level} mesh")
    m = L.Mesh(L.LPM_SEED_CUBED, args.level)
    n = m.n_leaves
    zeta = L.random_vortices(m.cx, m.cy, m.cz, m.area, n=8, seed=1234)
    g = zeta * m.area
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    X, Y, Z, G = dev(m.cx), dev(m.cy), dev(m.cz), dev(g)
    ctx = L.Context(0)
    stream = torch.cuda.ExternalStream(ctx.stream)
    torch.cuda.synchronize()
    pairs = n * (n - 1)

    if args.copy_probe:
        a = torch.empty(256 << 20, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log("[bench] copy probe: 256 MiB device-to-device")

    def timed(fn, reps=1):
        """(mean event ms per call on the library stream, wall s per call)"""
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
        ctx.synchronize()
        t0 = time.perf_counter()
        for e0, e1 in evs:
            e0.record(stream)
            fn()
            e1.record(stream)
        ctx.synchronize()
        wall = (time.perf_counter() - t0) / reps
        return float(np.mean([e0.elapsed_time(e1) for e0, e1 in evs])), wall

    def rel_l2(u, ref):
        return float(torch.linalg.norm(torch.stack(u) - torch.stack(ref)) / torch.linalg.norm(torch.stack(ref)))

    # --- FP64 direct sum: the error judge
    U64 = [torch.empty_like(X) for _ in range(3)]
    log("[bench] FP64 direct sum")
    ms64, _ = timed(lambda: ctx.direct_sum_dev(X, Y, Z, X, Y, Z, G, U64, prec=L.LPM_F64))

    # --- FP32C direct sum
    U32 = [torch.empty_like(X) for _ in range(3)]
    f32c = lambda: ctx.direct_sum_dev(X, Y, Z, X, Y, Z, G, U32, prec=L.LPM_F32C)
    f32c()
    ms32, wall32 = timed(f32c, reps=max(1, min(args.steps, 3)))
    err32 = rel_l2(U32, U64)
    log(f"[bench] direct: f64 {ms64:.1f} ms, f32c {ms32:.1f} ms, f32c rel-L2 {err32:.2e}")

    # --- treecode
    UT = [torch.empty_like(X) for _ in range(3)]
    p, theta, leaf = args.order, args.theta, args.leaf
    ev_ms = {}

    def tree_step(prec):
        T.build_dev(ctx, X, Y, Z, m.max_radius, leaf)
        T.moments_dev(ctx, G, p, 1)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        T.evaluate_dev(ctx, X, Y, Z, UT, theta, prec=prec, series=L.LPM_SERIES_CARTESIAN)
        e1.record(stream)
        ev_ms.setdefault(prec, []).append((e0, e1))

    log("[bench] FP64 treecode")
    _, wall_t64 = timed(lambda: tree_step(L.LPM_F64))
    err_t64 = rel_l2(UT, U64)
    far, near = T.last_stats(ctx)
    log(f"[bench] warmup {args.warmup}, {args.steps} timed FP32C treecode step(s)")
    for _ in range(args.warmup):
        tree_step(L.LPM_F32C)
    ctx.synchronize()
    ev_ms[L.LPM_F32C] = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        tree_step(L.LPM_F32C)
    ctx.synchronize()
    torch.cuda.synchronize()
    wall_tree = (time.perf_counter() - t0) / args.steps
    err_tree = rel_l2(UT, U64)
    eval_ms = float(np.mean([a.elapsed_time(b) for a, b in ev_ms[L.LPM_F32C]]))
    eval64_ms = float(np.mean([a.elapsed_time(b) for a, b in ev_ms[L.LPM_F64]]))
    far, near = T.last_stats(ctx)
    nn, depth, nleaves = T.info(ctx)
    log(f"[bench] tree f32c: {1e3 * wall_tree:.1f} ms/step (eval {eval_ms:.1f} ms), rel-L2 {err_tree:.2e}; "
        f"f64 tree {1e3 * wall_t64:.1f} ms, rel-L2 {err_t64:.2e}; far {far} near {near}")

    use_tree = err_tree <= 1e-5
    step_s = wall_tree if use_tree else wall32
    value = pairs / step_s
    tree_flops = near * DIRECT_FLOPS + far * (8 * nterms(p) + 16)
    tree_roof = {"bound": "valu", "kernel": "tree_eval", "achieved": tree_flops / (eval_ms / 1e3) / 1e12,
                 "peak": FP64_PEAK, "unit": "TFLOP/s", "traffic": None, "flops_per_launch": tree_flops,
                 "launch_ms": eval_ms, "far_pairs": far, "near_pairs": near}
    tree_roof["frac"] = tree_roof["achieved"] / FP64_PEAK
    direct_roof = {"bound": "valu", "kernel": "ds_f32c", "achieved": pairs * DIRECT_FLOPS / (ms32 / 1e3) / 1e12,
                   "peak": FP32_PEAK, "unit": "TFLOP/s", "traffic": None, "flops_per_launch": pairs * DIRECT_FLOPS,
                   "launch_ms": ms32}
    direct_roof["frac"] = direct_roof["achieved"] / FP32_PEAK

    # --- RK4 steps/s with the treecode velocity (4 evaluations per step)
    rk4 = None
    if not args.no_rk4:
        log("[bench] RK4 step (treecode, FP32C near field)")
        cfg = L.SolverConfig.default(method=L.LPM_METHOD_TREE, prec=L.LPM_F32C, order=p, theta=theta, leaf_max=leaf)
        za = dev(zeta + 2 * (2 * np.pi) * m.cz)
        A = dev(m.area)
        Xs, Ys, Zs = X.clone(), Y.clone(), Z.clone()
        ctx.rk4_step_dev(cfg, 0.01, 2 * np.pi, Xs, Ys, Zs, za, A)
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.rk4_step_dev(cfg, 0.01, 2 * np.pi, Xs, Ys, Zs, za, A)
        ctx.synchronize()
        rk4_s = time.perf_counter() - t0
        rk4 = {"rk4_steps_per_s": 1.0 / rk4_s, "rk4_ms_per_step": 1e3 * rk4_s,
               "rk4_solver": f"treecode p={p} theta={theta} leaf={leaf} FP32C near field, dt=0.01, Omega=2pi"}
        log(f"[bench] RK4: {1e3 * rk4_s:.1f} ms/step")

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
                   "step": ("treecode (build + moments + evaluation), Cartesian series "
                            f"p={p}, theta={theta}, leaf={leaf}, FP64 far field, FP32C near field")
                   if use_tree else "FP32C direct sum",
                   "rel_l2_err_vs_fp64_direct": err_tree if use_tree else err32},
        "roofline": tree_roof if use_tree else direct_roof,
        "cpu_baseline": cpu,
        "direct_f32c": {"interactions_per_s": pairs / (ms32 / 1e3), "ms": ms32, "rel_l2_err": err32,
                        "roofline": direct_roof},
        "direct_f64": {"interactions_per_s": pairs / (ms64 / 1e3), "ms": ms64,
                       "frac_of_fp64_peak": pairs * DIRECT_FLOPS / (ms64 / 1e3) / (FP64_PEAK * 1e12)},
        "tree_f64": {"ms_per_step": 1e3 * wall_t64, "eval_ms": eval64_ms, "rel_l2_err": err_t64,
                     "effective_interactions_per_s": pairs / wall_t64},
        "tree_f32c": {"ms_per_step": 1e3 * wall_tree, "eval_ms": eval_ms, "rel_l2_err": err_tree,
                      "effective_interactions_per_s": pairs / wall_tree, "nodes": nn, "depth": depth,
                      "leaves": nleaves, "speedup_vs_tree_f64": wall_t64 / wall_tree},
    }
    if rk4:
        out.update(rk4)
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
