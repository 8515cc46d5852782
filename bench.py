"""Benchmark: Mrays/s of the Kerr a=0.9 1920x1080 trace on one MI355X.

A step is one full-frame trace (2,073,600 rays) through libspacetime_hip's
C ABI, camera/sky/LUT already resident in HBM.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# Algorithmic FLOPs as written in the reference (src/metal.jl), counting
# + - * / sqrt as one each: kerr_rhs_mtl (:466-508) = 123, ks_rhs_mtl
# (:324-340) = 39, RK4 stage/combine arithmetic (:1382-1401) = 96, per-step
# radius/step-rule/test work (:1201-1260, :1421) = 27 (Kerr) / 16 (Schw).
F_STEP_KERR = 4 * 123 + 96 + 27      # 615
F_STEP_SCHW = 4 * 39 + 96 + 16       # 268
F_RAY_KERR = 330                     # ray gen + KS init + Bardeen + ISCO + escape direction/sky
F_RAY_SCHW = 180
FP32_PEAK_TFLOPS = 157.3             # MI355X FP32 vector, datasheet
FP64_PEAK_TFLOPS = 78.6


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "mixed", "fp64"])
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the tracer has no multi-GPU path in this build "
              "(rays shard by frame/tile with no exchange; see the design notes).", file=sys.stderr)
        sys.exit(2)

    from spacetime_amd import _lib as L
    from spacetime_amd import api, scenes

    prec = {"fp32": L.ST_PREC_FP32, "mixed": L.ST_PREC_MIXED, "fp64": L.ST_PREC_FP64}[args.precision]
    sc = scenes.kerr09_1080p(args.width, args.height)
    W, H = sc.width, sc.height
    ctx = api.HipPreviewContext(scenes.sky(), W, H, disc=sc.disc)
    c32, c64 = sc.cam_params()
    sp = sc.spacetime_params()
    targs = api.trace_args(W, H, sc.rule, precision=prec)

    def step():
        ctx.trace(c32, c64, sp, targs)

    for _ in range(args.warmup):
        step()
    ctx.sync()
    # per-launch device time (HIP events on the context's stream) and steps
    kern_ms = []
    for _ in range(3):
        step()
        st = ctx.stats()
        kern_ms.append(st["ms_total"])
    steps_fp32, steps_fp64, rays, promoted = st["steps_fp32"], st["steps_fp64"], st["rays"], st["promoted"]

    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.sync()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    mrays = W * H * args.steps / wall / 1e6

    kernel_ms = float(np.mean(kern_ms))
    flops = (steps_fp32 + steps_fp64) * F_STEP_KERR + rays * F_RAY_KERR
    peak = FP32_PEAK_TFLOPS if prec != L.ST_PREC_FP64 else FP64_PEAK_TFLOPS
    achieved_tf = flops / (kernel_ms * 1e-3) / 1e12
    traffic = None
    tfile = ROOT / "profiles" / "traffic_kerr09_1080p_fp32.json"
    if tfile.exists() and prec == L.ST_PREC_FP32 and (W, H) == (1920, 1080):
        try:
            traffic = json.loads(tfile.read_text()).get("bytes_per_launch")
        except Exception:
            traffic = None

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(sc)

    line = {
        "metric": "Mrays/s at 1080p Kerr a=0.9 (plain RK4 vs surrogate+fallback) at <=1e-3 rad error",
        "value": round(mrays, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "f32", "mixed": "f32+f64", "fp64": "f64"}[args.precision],
        "data": "synthetic (seeded 1024x512 star sky, analytic disc)",
        "config": {"workload": f"kerr a=0.9 {W}x{H} fixed-fraction RK4 h=0.05r, r_esc=100M, capture 1.01 r+, "
                               f"thin disc + sky, {args.precision}",
                   "rays_per_step": W * H, "parallelism": "single GPU"},
        "roofline": {"bound": "valu-fp32" if prec != L.ST_PREC_FP64 else "valu-fp64",
                     "achieved": round(achieved_tf, 3), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(achieved_tf / peak, 4), "traffic": traffic,
                     "kernel_ms": round(kernel_ms, 4), "flop_per_launch": flops,
                     "steps_per_ray": round((steps_fp32 + steps_fp64) / max(rays, 1), 3),
                     "f_step": F_STEP_KERR},
        "promoted_fraction": round(promoted / max(rays, 1), 5),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def cpu_baseline(sc):
    """The FP64 oracle (a C++ restatement of the reference CPU path) timed on
    this host: fixed-step RK4 (the same algorithm) over the full frame with
    every core, and adaptive Tsit5 (tol 1e-6, the reference `render`
    tolerance) on a bounded row sample."""
    try:
        sys.path.insert(0, str(ROOT / "tests"))
        import oracle  # noqa: F401
        from util import oracle_scene
    except Exception as e:  # the checker is absent: report it, do not guess
        return {"value": None, "unit": "Mrays/s", "cores": 0, "kind": "port", "sample": f"unavailable: {e}"}
    cores = len(os.sched_getaffinity(0))
    cores = min(cores, 16)
    o = oracle_scene(sc)
    rows = sc.height
    t0 = time.perf_counter()
    _, steps = o.render("rk4", row0=0, rows=rows, nthreads=cores)
    t_rk4 = time.perf_counter() - t0
    rk4_mrays = sc.width * rows / t_rk4 / 1e6
    trows = max(8, sc.height // 16)
    r0 = sc.height // 2 - trows // 2
    t0 = time.perf_counter()
    o.render("tsit5", tol=1e-6, row0=r0, rows=trows, nthreads=cores)
    t_ts = time.perf_counter() - t0
    ts_mrays = sc.width * trows / t_ts / 1e6
    return {"value": round(rk4_mrays, 4), "unit": "Mrays/s", "cores": cores, "kind": "port",
            "sample": f"full {sc.width}x{rows} frame, FP64 fixed-step RK4 (same algorithm), {t_rk4:.1f}s",
            "tsit5_1e-6": {"value": round(ts_mrays, 5), "sample": f"{trows} centre rows x {sc.width}, {t_ts:.1f}s"}}


if __name__ == "__main__":
    main()
