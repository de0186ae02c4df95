#!/usr/bin/env python3
"""Benchmark: Voronoi vertices/s of the device raycast walk at 5D, N=100k.

Workload: 100 000 generators in [0,1]^5 -- 90% from 8 isotropic Gaussians (centres
uniform in [0.2,0.8]^5, sigma 0.05, samples outside the cube redrawn) plus 10% uniform
background -- with the Dirichlet box cuboid(5, periodic=[]).  One step is one full
hv_build_mesh: tree build + the whole vertex walk, points already in HBM.

Prints ONE JSON line on stdout (diagnostics go to stderr).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "highvoronoi.jl_amd"))

import numpy as np  # noqa: E402

PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}
WORKLOAD = ("5D N=100000 Gaussian mixture (8 clusters, sigma=0.05, centres U[0.2,0.8]^5, "
            "10% uniform background, out-of-cube samples redrawn), cuboid(5, periodic=[])")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_leg(ctx, precision, steps, warmup):
    for _ in range(warmup):
        m = ctx.build_mesh(precision=precision)
        m.free()
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        m = ctx.build_mesh(precision=precision)
        stats.append(m.stats())
        m.free()
    t1 = time.perf_counter()
    return (t1 - t0) / steps, stats


def cpu_baseline(xs, cells):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from hvgpu import synth
    from oracle_py import Oracle
    b, nrm = synth.cube_planes(xs.shape[1])
    o = Oracle(xs, b, nrm)
    t0 = time.perf_counter()
    keys, _, st = o.voronoi(cells=(0, cells))
    t1 = time.perf_counter()
    return len(keys) / (t1 - t0), len(keys), t1 - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "bf16"])
    ap.add_argument("--cpu-cells", type=int, default=500, help="cells walked by the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported by this build "
              "(the box has a single MI355X; the multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    from hvgpu import synth
    from hvgpu._lib import Context

    xs = synth.gaussian_mixture_5d()
    d = xs.shape[1]
    base, normal = synth.cube_planes(d)
    ctx = Context(0)
    ctx.set_points(xs)
    ctx.set_boundary(base, normal)

    sec, stats = run_leg(ctx, args.precision, args.steps, args.warmup)
    st = stats[-1]
    nv = st["vertices"]
    value = nv / sec
    log(f"[{args.precision}] {nv} vertices, {st['rays']} rays, {st['levels']} levels, "
        f"{sec * 1e3:.2f} ms/step (device {st['ms_total']:.2f} ms, scan {st['ms_raycast']:.2f} ms)")

    extra = {}
    if args.precision != "fp64":
        sec64, stats64 = run_leg(ctx, "fp64", args.steps, args.warmup)
        extra["fp64_vertices_per_s"] = stats64[-1]["vertices"] / sec64
        extra["speedup_vs_fp64"] = value / extra["fp64_vertices_per_s"]
        extra["identical_vertex_count"] = stats64[-1]["vertices"] == nv
        extra["fallback_rate"] = st["fallbacks"] / max(1, st["rays"])
        extra["max_t_relerr"] = st["max_t_relerr"]

    # roofline of the candidate scan: 4d flops per scanned (ray, generator) pair
    pairs = st["pairs_scanned"]
    flops = 4.0 * d * pairs
    achieved = flops / (st["ms_raycast"] * 1e-3) / 1e12 if st["ms_raycast"] > 0 else 0.0
    peak = PEAK_TFLOPS[args.precision]
    roofline = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": None,
                "kernel": "k_raycast64", "launches": st["levels"],
                "avg_launch_ms": st["ms_raycast"] / max(1, st["levels"]),
                "flops_per_launch": flops / max(1, st["levels"]),
                "pairs_per_ray": pairs / max(1, st["rays"])}

    cpu = None
    if not args.no_cpu:
        vps, ncpu, tcpu = cpu_baseline(xs, args.cpu_cells)
        cpu = {"value": vps, "unit": "vertices/s", "cores": 1, "kind": "port",
               "sample": f"oracle (C restatement of the reference walk) over cells 1..{args.cpu_cells} of the "
                         f"same point set: {ncpu} vertices in {tcpu:.1f} s"}
        log(f"[cpu] {ncpu} vertices in {tcpu:.2f} s -> {vps:.0f} vertices/s")

    out = {
        "metric": "Voronoi vertices/s at 5D N=100k: mixed-precision vs FP64 GPU, identical vertex set",
        "value": value,
        "unit": "vertices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else f"{args.precision}+f64",
        "data": "synthetic",
        "config": {"workload": WORKLOAD, "n": int(xs.shape[0]), "d": d, "precision": args.precision,
                   "vertices": nv, "rays": st["rays"], "levels": st["levels"],
                   "duplicate_rays": st["duplicates"], "near_ties": st["near_ties"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
