#!/usr/bin/env python3
"""Benchmark: Voronoi vertices/s of the device raycast walk at 5D, N=100k.

Workload: 100 000 generators in [0,1]^5 -- 90% from 8 isotropic Gaussians (centres
uniform in [0.2,0.8]^5, sigma 0.05, samples outside the cube redrawn) plus 10% uniform
background -- with the Dirichlet box cuboid(5, periodic=[]).  One step is one full
hv_build_mesh: tree build + the whole vertex walk, points already in HBM.

The headline leg is the mixed-precision scan (BF16 MFMA filter + FP64 re-verification);
the plain FP64 scan runs beside it on the same points, and the two key tables are
compared in full.  Prints ONE JSON line on stdout (diagnostics go to stderr).
"""
import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "highvoronoi.jl_amd"))

import numpy as np  # noqa: E402

PEAK_TFLOPS = {"fp64": 78.6, "bf16": 2500.0}
WORKLOAD = ("5D N=100000 Gaussian mixture (8 clusters, sigma=0.05, centres U[0.2,0.8]^5, "
            "10% uniform background, out-of-cube samples redrawn, seed 2), cuboid(5, periodic=[])")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_leg(ctx, precision, steps, warmup):
    for _ in range(warmup):
        ctx.build_mesh(precision=precision).free()
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        m = ctx.build_mesh(precision=precision)
        stats.append(m.stats())
        m.free()
    t1 = time.perf_counter()
    return (t1 - t0) / steps, stats


def mesh_table(ctx, precision):
    m = ctx.build_mesh(precision=precision)
    keys, coords = m.export()
    m.free()
    return keys, coords


def compare(ctx, p1, p2):
    """(identical key tables, max |coordinate difference| or None when the keys differ)"""
    k1, c1 = mesh_table(ctx, p1)
    k2, c2 = mesh_table(ctx, p2)
    same = bool(k1.shape == k2.shape and np.array_equal(k1, k2))
    err = float(np.max(np.abs(c1 - c2))) if same and len(c1) else None
    return same, err


def cpu_baseline(xs, cells_per_thread, threads):
    """The oracle (C restatement of the reference walk), one context per thread over
    contiguous cell ranges; unique vertices of the union per second of wall time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from hvgpu import synth
    from oracle_py import Oracle
    b, nrm = synth.cube_planes(xs.shape[1])
    results = [None] * threads
    oracles = [Oracle(xs, b, nrm) for _ in range(threads)]

    def work(t):
        lo = t * cells_per_thread
        results[t] = oracles[t].voronoi(cells=(lo, lo + cells_per_thread))[0]

    ths = [threading.Thread(target=work, args=(t,)) for t in range(threads)]
    t0 = time.perf_counter()
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    t1 = time.perf_counter()
    allk = np.concatenate(results)
    nuniq = len(np.unique(allk, axis=0))
    return nuniq / (t1 - t0), nuniq, t1 - t0


def measured_traffic(precision):
    """HBM bytes per scan launch at this workload, from the committed counter summary
    (FETCH_SIZE and WRITE_SIZE runs scaled by a copy kernel of known size); None if absent."""
    path = os.path.join(ROOT, "profiles", "k_raycast64_5d_100k_mixture_traffic.txt")
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == precision:
                    return float(parts[2])
    except OSError:
        pass
    return None


def bench_batched(args):
    """512 independent meshes: sequential hv_build_mesh calls on one context, one
    hv_build_mesh_batch call on host workers (per-mesh walks on args.workers threads, one
    stream each) and one hv_build_mesh_batch call that walks the meshes of each dimension
    together (the headline); every mesh of the batches is compared with its sequential build."""
    from hvgpu import synth
    from hvgpu._lib import Context, build_mesh_batch

    meshes = synth.batched_meshes(count=512, seed=3)
    pts = [x for _, x in meshes]
    planes = [synth.cube_planes(x.shape[1]) for x in pts]
    prec = args.precision
    ctx = Context(0)

    def sequential():
        out = []
        for x, (b, nrm) in zip(pts, planes):
            ctx.set_points(x)
            ctx.set_boundary(b, nrm)
            out.append(ctx.build_mesh(precision=prec))
        return out

    def batched_with(nworkers):
        def run():
            out, status = build_mesh_batch(pts, planes, precision=prec, nworkers=nworkers)
            if any(status):
                raise RuntimeError(f"batched build failed: {sorted(set(status.tolist()))}")
            return out
        return run

    batched = batched_with(0)

    def timed(fn):
        for _ in range(args.warmup):
            for m in fn():
                m.free()
        best = None
        res = None
        for _ in range(args.steps):
            t0 = time.perf_counter()
            out = fn()
            dt = time.perf_counter() - t0
            if res is not None:
                for m in res:
                    m.free()
            res = out
            best = dt if best is None else min(best, dt)
        return best, res

    legs = set(args.legs.split(","))
    if legs != {"sequential", "workers", "fused"}:
        # a subset of the legs (profiling runs): timings only
        for name, fn in (("sequential", sequential), ("workers", batched_with(-args.workers)), ("fused", batched)):
            if name in legs:
                t, ms = timed(fn)
                log(f"[batched {prec}] {name}: {t * 1e3:.1f} ms, {sum(m.stats()['vertices'] for m in ms)} vertices")
                for m in ms:
                    m.free()
        return
    tseq, mseq = timed(sequential)
    twrk, mwrk = timed(batched_with(-args.workers))
    tbat, mbat = timed(batched)
    nv = sum(m.stats()["vertices"] for m in mbat)
    same = all(m.stats()["vertices"] == q.stats()["vertices"] == w.stats()["vertices"]
               for m, q, w in zip(mbat, mseq, mwrk))
    same = same and all(np.array_equal(m.export()[0], q.export()[0]) for m, q in zip(mbat[:64], mseq[:64]))
    for m in mbat + mseq + mwrk:
        m.free()
    dims = {d: sum(1 for dd, _ in meshes if dd == d) for d in (3, 4, 5, 6)}
    log(f"[batched {prec}] {nv} vertices in 512 meshes: sequential {tseq * 1e3:.1f} ms, "
        f"per-mesh walks on {args.workers} workers {twrk * 1e3:.1f} ms, one walk per dimension {tbat * 1e3:.1f} ms")
    out = {
        "metric": "Voronoi vertices/s, 512 independent meshes in one batched call (one walk per dimension)",
        "value": nv / tbat,
        "unit": "vertices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tbat * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16+f64" if prec == "bf16" else "f64",
        "data": "synthetic",
        "config": {"workload": "512 meshes, d uniform in {3,4,5,6}, N = floor(exp(U[ln 200, ln 5000])), "
                               "points U[0,1]^d (seed 3*2^32+k), cuboid(d, periodic=[])",
                   "meshes_per_dimension": dims, "vertices": nv, "precision": prec, "workers": args.workers},
        "sequential_vertices_per_s": nv / tseq,
        "sequential_ms": tseq * 1e3,
        "workers_ms": twrk * 1e3,
        "speedup_vs_sequential": tseq / tbat,
        "speedup_vs_workers": twrk / tbat,
        "identical_to_sequential": bool(same),
        "timing": "best of steps, host wall clock around the call",
    }
    print(json.dumps(out), flush=True)


def bench_mc_points(args):
    """Point-sample Monte-Carlo volumes and areas (1e8 stratified samples) at 6D N=50k in the
    periodic unit cube cuboid(6) (the mesh of the originals' cells with their periodic images),
    against the exact polytope integrator on the same mesh."""
    from hvgpu import synth
    from hvgpu.geometry import cuboid
    from hvgpu._lib import Context

    d, n, ns = 6, 50_000, 100_000_000
    xs = synth.uniform(d, n, 4)
    base, normal, bc = cuboid(d).arrays()
    ctx = Context(0)
    ctx.set_points(xs)
    ctx.set_boundary(base, normal, bc)
    if args.warmup > 0:   # the first build maps the context's pools
        ctx.build_mesh(precision=args.precision).free()
    t0 = time.perf_counter()
    m = ctx.build_mesh(precision=args.precision)
    mesh_ms = (time.perf_counter() - t0) * 1e3
    nv = m.stats()["vertices"]
    n_ext = len(m.images()[0])
    times, kms, pairs = [], [], []
    mc = None
    for k in range(args.warmup + args.steps):
        t0 = time.perf_counter()
        mc = ctx.integrate(m, "points", nsamples=ns, seed=17 + k)
        if k >= args.warmup:
            times.append(time.perf_counter() - t0)
            kms.append(mc["ms_kernel"])
            pairs.append(mc["pairs"])
    t0 = time.perf_counter()
    ex = ctx.integrate(m, "polytope")
    exact_ms = (time.perf_counter() - t0) * 1e3
    m.free()
    sec = min(times)
    vsum = float(np.sum(mc["volumes"]))
    vol, A, Am = ex["volumes"], ex["areas"], mc["areas"]
    same_rows = bool(np.array_equal(mc["nbr_ptr"], ex["nbr_ptr"]) and np.array_equal(mc["neighbors"], ex["neighbors"]))
    vol_l1 = float(np.sum(np.abs(mc["volumes"] - vol)) / np.sum(vol))
    sigma = np.sqrt(vol * (1.0 - vol) / ns)
    vol_z = float(np.max(np.abs(mc["volumes"] - vol) / sigma))
    area_l1 = float(np.sum(np.abs(Am - A)) / np.sum(A)) if same_rows else None
    area_bias = float((np.sum(Am) - np.sum(A)) / np.sum(A))
    # roofline of the sampling kernel: one (sample, generator) distance = 2d FP32 flops
    avg_ms = float(np.mean(kms))
    avg_pairs = float(np.mean(pairs))
    achieved = avg_pairs * 2 * d / (avg_ms * 1e-3) / 1e12
    peak = 157.3
    log(f"[mc-points] 6D N=50k periodic: mesh {mesh_ms:.0f} ms ({nv} vertices, {n_ext} generators with images); "
        f"{ns:.0e} samples: {sec * 1e3:.0f} ms (device {mc['ms']:.0f} ms, sampling kernel {avg_ms:.0f} ms, "
        f"{avg_pairs / ns:.0f} distances per sample), volume sum {vsum:.15f}")
    log(f"[mc-points] exact integrator on the same mesh {exact_ms:.0f} ms; volume L1 {vol_l1:.4f} (max {vol_z:.2f} sigma), "
        f"area L1 {area_l1}, area bias {area_bias:+.4f}")
    out = {
        "metric": "Monte-Carlo point samples/s (nearest and second-nearest generator per sample), 6D N=50k periodic",
        "value": ns / sec,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 distances, f64 decisions",
        "data": "synthetic",
        "config": {"workload": "6D N=50000 uniform in [0,1]^6 (seed 4), cuboid(6) periodic in every direction, "
                               "1e8 stratified samples, slab 0.03", "vertices": nv, "generators_with_images": n_ext,
                   "mesh_ms": mesh_ms},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": None, "kernel": f"k_mc_points<{d}>", "avg_launch_ms": avg_ms,
                     "pairs_per_launch": avg_pairs, "flops_per_pair": 2 * d,
                     "note": "peak = published FP32 vector rate; this build disables packed FP32, which halves the "
                             "reachable VALU rate to 78.6 TFLOP/s"},
        "volume_sum": vsum,
        "vs_exact_on_the_same_mesh": {"exact_integrator_ms": exact_ms, "same_neighbour_rows": same_rows,
                                      "volume_rel_l1": vol_l1, "volume_max_sigma": vol_z,
                                      "area_rel_l1": area_l1, "area_total_bias": area_bias},
    }
    print(json.dumps(out), flush=True)


def bench_integrators(args):
    """Wall time (best of steps, after warm-up) of one mesh build and of each integrator on it,
    and the Monte-Carlo errors (relative L1 of volumes and of areas) against the exact one."""
    from hvgpu import synth
    from hvgpu._lib import Context

    ctx = Context(0)

    def timed(fn):
        for _ in range(max(1, args.warmup)):
            r = fn()
            if hasattr(r, "free"):
                r.free()
        best, res = None, None
        for _ in range(max(1, args.steps)):
            t0 = time.perf_counter()
            r = fn()
            dt = time.perf_counter() - t0
            if res is not None and hasattr(res, "free"):
                res.free()
            res = r
            best = dt if best is None else min(best, dt)
        return res, best * 1e3

    rows = {}
    for name, d, n, seed in [("2D N=1000", 2, 1000, 0), ("4D N=20000", 4, 20000, 1), ("6D N=50000", 6, 50000, 4)]:
        xs = synth.uniform(d, n, seed)
        base, normal = synth.cube_planes(d)
        ctx.set_points(xs)
        ctx.set_boundary(base, normal)
        m, tm = timed(lambda: ctx.build_mesh(precision=args.precision))
        ex, tp = timed(lambda: ctx.integrate(m, "polytope"))
        f2 = lambda p: np.stack([p[:, 0], p[:, 1] ** 2], axis=1)
        ei, ti = timed(lambda: ctx.integrate(m, "polytope", integrand=f2, nf=2))
        ndir = 1000 if d < 6 else 200
        md, td = timed(lambda: ctx.integrate(m, "montecarlo", ndir=ndir, seed=1))
        mp, tq = timed(lambda: ctx.integrate(m, "points", nsamples=100_000_000, seed=1))
        nv = m.stats()["vertices"]
        m.free()
        A = ex["areas"]

        def err(r):
            return [float(np.sum(np.abs(r["volumes"] - ex["volumes"])) / np.sum(ex["volumes"])),
                    float(np.sum(np.abs(r["areas"] - A)) / np.sum(A))]

        rows[name] = {"vertices": nv, "mesh_ms": tm, "polytope_ms": tp, "polytope_integrand_ms": ti,
                      "integral_x0": float(ei["bulk"][:, 0].sum()), "integral_x1_sq": float(ei["bulk"][:, 1].sum()),
                      "directional_mc_ms": td,
                      "directional_mc_ndir": ndir, "directional_mc_l1_vol_area": err(md),
                      "points_mc_ms": tq, "points_mc_samples": 100_000_000, "points_mc_l1_vol_area": err(mp)}
        log(f"[integrators] {name}: {rows[name]}")
    out = {
        "metric": "integrator wall time per call (mesh, exact polytope, polytope with an integrand, directional MC, point-sample MC)",
        "value": rows["6D N=50000"]["points_mc_ms"],
        "unit": "ms (6D N=50k point-sample MC, 1e8 samples)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "higher_is_better": False,
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "uniform points (seeds 0, 1, 4), cuboid(d, periodic=[])", "cases": rows},
    }
    print(json.dumps(out), flush=True)


def bench_clipped(args):
    """The 5D N=100k mixture with out-of-cube cluster samples clipped onto the cube's faces
    instead of drawn again: generators on a boundary plane.  Such a generator coincides with
    its mirror image, the reference's walk of its cell reaches a vertex with more than d + 1
    generators (its non-general path), and the device build reports HV_ERR_DEGENERATE; the
    line gives that outcome per precision (value null when no mesh is built)."""
    from hvgpu import synth
    from hvgpu._lib import Context, HVError

    xs = synth.gaussian_mixture_5d(clip=True)
    on = int(np.sum(((xs == 0.0) | (xs == 1.0)).any(axis=1)))
    b, nrm = synth.cube_planes(5)
    ctx = Context(0)
    ctx.set_points(xs)
    ctx.set_boundary(b, nrm)
    res = {}
    for prec in ("bf16", "fp64"):
        t0 = time.perf_counter()
        try:
            m = ctx.build_mesh(precision=prec)
            res[prec] = {"status": "ok", "vertices": m.stats()["vertices"], "ms": (time.perf_counter() - t0) * 1e3}
            m.free()
        except HVError as e:
            res[prec] = {"status": f"error {e.code}", "message": str(e), "ms": (time.perf_counter() - t0) * 1e3}
        log(f"[clipped] {prec}: {res[prec]}")
    ok = res["bf16"]["status"] == "ok"
    out = {"metric": "Voronoi vertices/s at 5D N=100k, mixture clipped onto the cube (generators on boundary planes)",
           "value": res["bf16"]["vertices"] / (res["bf16"]["ms"] * 1e-3) if ok else None, "unit": "vertices/s",
           "n_gpus": 1, "steps": 1, "warmup": 0, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "bf16+f64", "data": "synthetic",
           "config": {"workload": "5D N=100000 Gaussian mixture (as the headline workload, seed 2) with out-of-cube "
                                  "samples clipped to [0,1]^5", "generators_on_a_plane": on},
           "outcome": res}
    print(json.dumps(out), flush=True)


def bench_scale(args):
    """Large uniform inputs in the unit cube: best-of-steps build time for BF16 and FP64 after a
    warm-up build, levels, and the key tables of the two precisions compared."""
    from hvgpu import synth
    from hvgpu._lib import Context

    ctx = Context(0)
    cases = {}
    for d, n in [(2, 10_000_000), (3, 5_000_000), (5, 1_000_000), (6, 200_000)]:
        xs = synth.uniform(d, n, 77)
        base, normal = synth.cube_planes(d)
        ctx.set_points(xs)
        ctx.set_boundary(base, normal)
        row, keys = {}, {}
        for prec in ("bf16", "fp64"):
            sec, stats = run_leg(ctx, prec, max(1, args.steps), max(1, args.warmup))
            st = stats[-1]
            row[prec] = {"ms": sec * 1e3, "vertices_per_s": st["vertices"] / sec, "levels": st["levels"]}
            row["vertices"] = st["vertices"]
            keys[prec] = mesh_table(ctx, prec)[0]
        row["identical_vertex_set"] = bool(np.array_equal(keys["bf16"], keys["fp64"]))
        del keys
        cases[f"{d}D N={n}"] = row
        log(f"[scale] {d}D N={n}: {row}")
    out = {
        "metric": "Voronoi vertices/s of large uniform inputs (BF16 build)",
        "value": cases["5D N=1000000"]["bf16"]["vertices_per_s"],
        "unit": "vertices/s (5D N=1M)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "higher_is_better": True,
        "vs_baseline": None,
        "dtype": "bf16+f64",
        "data": "synthetic",
        "config": {"workload": "uniform points in [0,1]^d (seed 77), cuboid(d, periodic=[])", "cases": cases},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="bf16", choices=["bf16", "fp64"],
                    help="headline leg (the other precision, and the FP32 filter, run beside it)")
    ap.add_argument("--cpu-cells", type=int, default=2000, help="cells per thread walked by the CPU baseline (about 15-20 s on 16 threads)")
    ap.add_argument("--cpu-cells1", type=int, default=2500, help="cells walked by the 1-thread CPU baseline (about 20 s)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(16, cores)")
    ap.add_argument("--tune", action="store_true",
                    help="run the autotuner (hv_tune) on the workload first and time with its scan tiles")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true", help="skip the FP64 leg and key-table comparison")
    ap.add_argument("--batched", action="store_true",
                    help="the 512-mesh workload instead (d in 3..6, N log-uniform in 200..5000): one "
                         "hv_build_mesh_batch call against 512 sequential hv_build_mesh calls")
    ap.add_argument("--workers", type=int, default=8, help="host workers (streams) of the batched call")
    ap.add_argument("--legs", default="sequential,workers,fused",
                    help="--batched: legs to run (a subset prints timings only, for profiling)")
    ap.add_argument("--clipped", action="store_true",
                    help="the headline mixture with out-of-cube samples clipped onto the faces (its own line)")
    ap.add_argument("--mc-points", action="store_true",
                    help="the point-sample Monte-Carlo integrator instead: 6D N=50k uniform, cuboid(6) periodic, "
                         "1e8 samples; against the exact integrator at 6D N=10k periodic")
    ap.add_argument("--integrators", action="store_true",
                    help="mesh + the three integrators (exact polytope, directional MC, point-sample MC) at "
                         "2D N=1000, 4D N=20k and 6D N=50k, with their errors against the exact integrator")
    ap.add_argument("--scale", action="store_true",
                    help="large uniform inputs (2D N=10M, 3D N=5M, 5D N=1M, 6D N=200k): vertices/s of the BF16 "
                         "and FP64 builds and whether their key tables agree")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 256 MiB device copy (3x): the known-size kernel that scales the "
                         "FETCH_SIZE / WRITE_SIZE counters into bytes")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        x = torch.ones(64 * 1024 * 1024, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        log("copied 268435456 bytes, 3 times")
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (single-GPU build; "
              "the multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    from hvgpu import synth
    from hvgpu._lib import Context

    if args.clipped:
        return bench_clipped(args)
    if args.batched:
        return bench_batched(args)
    if args.mc_points:
        return bench_mc_points(args)
    if args.integrators:
        return bench_integrators(args)
    if args.scale:
        return bench_scale(args)

    xs = synth.gaussian_mixture_5d()
    d = xs.shape[1]
    base, normal = synth.cube_planes(d)
    ctx = Context(0)
    ctx.set_points(xs)
    ctx.set_boundary(base, normal)

    tiles = None
    if args.tune:
        tiles = ctx.tune(args.precision)   # the autotuner's scan tiles for this workload
        log(f"[tune] rays_per_block {tiles[0]}, packet_floor {tiles[1]}: {tiles[2]:.2f} ms")
    sec, stats = run_leg(ctx, args.precision, args.steps, args.warmup)
    st = stats[-1]
    nv = st["vertices"]
    value = nv / sec
    log(f"[{args.precision}] {nv} vertices, {st['rays']} rays, {st['levels']} levels, "
        f"{sec * 1e3:.2f} ms/step (device {st['ms_total']:.2f} ms, scan {st['ms_raycast']:.2f} ms)")

    extra = {}
    if not args.no_compare:
        other = "fp64" if args.precision == "bf16" else "bf16"
        sec2, stats2 = run_leg(ctx, other, args.steps, args.warmup)
        s2 = stats2[-1]
        log(f"[{other}] {s2['vertices']} vertices, {sec2 * 1e3:.2f} ms/step (scan {s2['ms_raycast']:.2f} ms)")
        same, err = compare(ctx, args.precision, other)
        fp64_vps = (s2["vertices"] / sec2) if other == "fp64" else value
        mixed_vps = value if args.precision == "bf16" else s2["vertices"] / sec2
        mixed = st if args.precision == "bf16" else s2
        sec3, stats3 = run_leg(ctx, "fp32", args.steps, args.warmup)
        s3 = stats3[-1]
        log(f"[fp32] {s3['vertices']} vertices, {sec3 * 1e3:.2f} ms/step (scan {s3['ms_raycast']:.2f} ms)")
        same32, err32 = compare(ctx, "fp32", "fp64")
        extra = {"fp64_vertices_per_s": fp64_vps, "mixed_vertices_per_s": mixed_vps,
                 "speedup_vs_fp64": mixed_vps / fp64_vps, "identical_vertex_set": same,
                 "fp32_vertices_per_s": s3["vertices"] / sec3, "fp32_identical_vertex_set": same32,
                 "fp32_max_coord_err_vs_fp64": err32,
                 "fp32_fallback_rate": s3["fallbacks"] / max(1, s3["rays"]),
                 "fallback_rate": mixed["fallbacks"] / max(1, mixed["rays"]),
                 "fp64_rechecks_per_ray": mixed["fp64_rechecks"] / max(1, mixed["rays"]),
                 "max_coord_err_vs_fp64": err}

    # roofline of the candidate-scan kernel (k_raycast64): 4d algorithmic flops per scanned
    # (ray, generator) pair; kernel time from HIP events on the context's stream
    launches = st["levels"] - 1
    pairs = st["pairs_scanned"]
    flops = 4.0 * d * pairs
    achieved = flops / (st["ms_raycast"] * 1e-3) / 1e12 if st["ms_raycast"] > 0 else 0.0
    peak = PEAK_TFLOPS[args.precision]
    roofline = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": measured_traffic(args.precision),
                "traffic_source": "profiles/k_raycast64_5d_100k_mixture_traffic.txt: rocprofv3 --pmc FETCH_SIZE and "
                                  "WRITE_SIZE runs of this workload, scaled by a copy of known size (not measured "
                                  "in this run)",
                "kernel": f"k_raycast64<5,{1 if args.precision == 'bf16' else 0}>", "launches": launches,
                "avg_launch_ms": st["ms_raycast"] / max(1, launches),
                "flops_per_launch": flops / max(1, launches),
                "pairs_per_ray": pairs / max(1, st["rays"])}

    cpu = None
    if not args.no_cpu:
        # all host cores of the box's share (at most 16 per GPU), and one thread
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        vps, ncpu, tcpu = cpu_baseline(xs, args.cpu_cells, threads)
        vps1, ncpu1, tcpu1 = cpu_baseline(xs, args.cpu_cells1, 1)
        cpu = {"value": vps, "unit": "vertices/s", "cores": threads, "kind": "port",
               "sample": f"oracle (C restatement of the reference walk, FP64) over cells 1..{args.cpu_cells * threads} "
                         f"of the same point set, {threads} threads on contiguous cell ranges: "
                         f"{ncpu} unique vertices in {tcpu:.1f} s",
               "value_1_thread": vps1,
               "sample_1_thread": f"the same, 1 thread over cells 1..{args.cpu_cells1}: {ncpu1} vertices in {tcpu1:.1f} s",
               "host_cpus": os.cpu_count()}
        log(f"[cpu] {ncpu} vertices in {tcpu:.2f} s on {threads} threads -> {vps:.0f} vertices/s; "
            f"1 thread {vps1:.0f} vertices/s")

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
        "dtype": "bf16+f64" if args.precision == "bf16" else "f64",
        "data": "synthetic",
        "config": {"workload": WORKLOAD, "n": int(xs.shape[0]), "d": d, "precision": args.precision,
                   "vertices": nv, "rays": st["rays"], "levels": st["levels"],
                   "duplicate_rays": st["duplicates"], "near_ties": st["near_ties"],
                   "tiles": {"rays_per_block": tiles[0], "packet_floor": tiles[1], "source": "hv_tune"} if tiles else
                            {"rays_per_block": 64, "packet_floor": 2048, "source": "defaults"}},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
