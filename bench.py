"""Benchmark of the spherical Biot-Savart sum on one MI355X.

Workload: cubed-sphere level 9, 1,572,864 leaf-face particles (targets =
sources = leaf faces, the face self pair excluded), zeta = sum of 8 Gaussian
vortices (seeded), gamma = zeta * area.  One step = one full velocity
evaluation of all particles by the FP32-compensated direct sum, inputs
resident in HBM.  value = N (N - 1) pair interactions per second.

Also reported: the FP64 direct sum rate, the FP32C error against the FP64 sum
(relative L2 over all particles), the roofline of the dominant kernel
(ds_f32c) from HIP events on the library's stream, and a CPU baseline (the
oracle, OpenMP, on a bounded sample of targets).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

FLOPS_PER_PAIR = 18          # velocity pair: d 3, |d|^2 5, q 2, rcp 1, w 1, acc 6
FP32_PEAK_TFLOPS = 157.3     # MI355X published FP32 vector peak
FP64_PEAK_TFLOPS = 78.6
METRIC = "Biot-Savart interactions/s + RK4 steps/s at 1.57M particles, rel L2 err ≤1e-5"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(tgt, src, g, skip, target_seconds: float):
    """Oracle (reference-order FP64, OpenMP over targets) on a sample of targets."""
    import numpy as np
    from oracle import oracle as O
    n = len(tgt[0])
    rng = np.random.default_rng(5)
    k = 16
    while True:
        idx = np.sort(rng.choice(n, k, replace=False))
        sub = tuple(t[idx] for t in tgt)
        t0 = time.perf_counter()
        O.bs_direct(sub, src, g, skip=skip[idx])
        dt = time.perf_counter() - t0
        if dt >= target_seconds or k >= n:
            break
        k = min(n, int(k * max(2.0, 1.2 * target_seconds / max(dt, 1e-3))))
    pairs = k * (len(src[0]) - 1)
    return {"value": pairs / dt, "unit": "interactions/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"{k} random targets x all {len(src[0])} sources, FP64 reference-order oracle, {dt:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--level", type=int, default=9)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-f64", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1", "n_gpus": args.gpus}))
        return 2

    import numpy as np
    import torch
    import lpmcxx_amd as L

    log(f"[bench] building cubed-sphere level {args.level} mesh")
    m = L.Mesh(L.LPM_SEED_CUBED, args.level)
    n = m.n_leaves
    zeta = L.random_vortices(m.cx, m.cy, m.cz, m.area, n=8, seed=1234)
    g = zeta * m.area
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    X, Y, Z, G = dev(m.cx), dev(m.cy), dev(m.cz), dev(g)
    U = [torch.empty_like(X) for _ in range(3)]
    ctx = L.Context(0)
    stream = torch.cuda.ExternalStream(ctx.stream)
    torch.cuda.synchronize()

    def step(prec, out):
        ctx.direct_sum_dev(X, Y, Z, X, Y, Z, G, out, prec=prec)

    log(f"[bench] warmup {args.warmup} step(s), N = {n}")
    for _ in range(args.warmup):
        step(L.LPM_F32C, U)
    ctx.synchronize()

    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record(stream)
        step(L.LPM_F32C, U)
        ev[k][1].record(stream)
    ctx.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    ms_per_step = 1e3 * wall / args.steps
    pairs = n * (n - 1)
    value = pairs / (wall / args.steps)
    log(f"[bench] f32c: {ms_per_step:.2f} ms/step, {value:.4e} interactions/s, event {launch_ms:.2f} ms")

    extra = {}
    if not args.no_f64:
        U64 = [torch.empty_like(X) for _ in range(3)]
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        step(L.LPM_F64, U64)
        b.record(stream)
        ctx.synchronize()
        t64 = a.elapsed_time(b) / 1e3
        u32 = torch.stack(U).cpu().numpy()
        u64 = torch.stack(U64).cpu().numpy()
        err = float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64))
        emax = float(np.sqrt(((u32 - u64) ** 2).sum(0)).max() / np.sqrt((u64 ** 2).sum(0)).max())
        extra = {"f64_interactions_per_s": pairs / t64, "f64_ms": 1e3 * t64,
                 "f64_frac_of_fp64_peak": pairs * FLOPS_PER_PAIR / t64 / (FP64_PEAK_TFLOPS * 1e12),
                 "rel_l2_err_f32c_vs_f64": err, "rel_max_err_f32c_vs_f64": emax}
        log(f"[bench] f64: {1e3 * t64:.1f} ms, f32c rel-L2 err vs f64 {err:.3e}")

    achieved = pairs * FLOPS_PER_PAIR / (launch_ms / 1e3) / 1e12
    roofline = {"bound": "valu", "kernel": "ds_f32c", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
                "flops_per_launch": pairs * FLOPS_PER_PAIR, "launch_ms": launch_ms}

    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline")
        skip = np.arange(n, dtype=np.int64)
        cpu = cpu_baseline((m.cx, m.cy, m.cz), (m.cx, m.cy, m.cz), g, skip, args.cpu_seconds)

    out = {
        "metric": METRIC, "value": value, "unit": "interactions/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32c", "data": "synthetic",
        "config": {"workload": f"direct sum, cubed sphere level {args.level}, {n} particles, 8 Gaussian vortices",
                   "particles": n, "pairs_per_step": pairs, "precision": "FP32 compensated (FP64 final reduce)"},
        "roofline": roofline, "cpu_baseline": cpu, **extra,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
