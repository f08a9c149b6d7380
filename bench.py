"""Benchmark: ray-bounces/s of the stochastic split ray tracer on one MI355X.

Workload: the seeded polyhedral hall (fan-shaped plan extruded to 12 m,
~2,000 jittered triangles, per-surface random absorption and scattering per
band, 16 receivers r = 0.5 m), 1e6 geodesic source rays, 3 s cut-off, 1 ms
bins, 32^3 voxel grid.  One step is one complete trace (SplitRayTracer.Begin
-> Combine) of all rays; the inputs are resident in HBM before the timed
region starts.

Precision (--precision): "fp32" (default) walks the voxel grid and shades in
FP32 with FP64 histograms, the throughput arithmetic; "mixed" walks with FP32
triangle tests widened by their error bounds and decides every hit, and does
all shading, in FP64, so its histograms equal the FP64 tracer's bin for bin;
"fp64" walks in FP64 as well.

Prints one JSON line: `value` is ray-bounces per second over the timed steps.
`roofline` prices the octave kernel in algorithmic HBM bytes (128 B per octave
bounce + 40 B per spawned ray) per second of its own launch time against the
8 TB/s peak, gives the HBM bytes per launch rocprofv3 measured (`traffic`),
and names the resource that actually binds (`binding`: VALU issue, from the
profiled VALU instructions per bounce).  `cpu_baseline` times the C FP64
restatement on the host's cores over a bounded sample.  Other workloads
(--workload) report their own metric in the same form; `--workload jnd`
measures the second half of the metric: the wall time a plain and a
denoised trace of the 50k-triangle concert hall need to reach JND accuracy.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

BYTES_OCT_BOUNCE = 128      # read ray 40 B + write ray 40 B + hit triangle 48 B
BYTES_BROAD_BOUNCE = 184    # broadband ray 68 B x 2 + hit triangle 48 B
BYTES_SPAWN = 40            # each spawned octave ray written once
HBM_PEAK_GBS = 8000.0
# arithmetic of the path: walk test / hit decision and shading / histograms
DTYPES = {"fp32": "f32 (f64 histograms)", "mixed": "f32 walk screen + f64 hits, shading, histograms",
          "fp64": "f64"}
FP64_PEAK_TFLOPS = 78.6


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def traffic_file(precision):
    return os.path.join(ROOT, "profiles", f"k_octave_{precision}_hall_traffic.json")


def load_profile(args, n_rays):
    """The committed rocprofv3 counter summary of this precision's k_octave on
    the 1e6-ray hall, and whether this run is that same workload."""
    try:
        with open(traffic_file(args.precision)) as f:
            prof = json.load(f)
    except (OSError, ValueError):
        return None, False
    same = (args.workload == "hall" and prof.get("rays") == n_rays and prof.get("cutoff") == args.cutoff
            and prof.get("grid") == (args.grid or 32) and prof.get("precision") == args.precision
            and not args.tune and not args.autotune)
    return prof, same


def measured_traffic(args, tot, n_rays):
    """HBM bytes per k_octave launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes (each scaled by what the counter reads for a plain copy
    of known size): the profiled command's bytes per octave bounce times this
    run's octave bounces per launch (launch counts vary from run to run; the
    bounces of a workload do not).  None when no profile is committed."""
    prof, same = load_profile(args, n_rays)
    if prof is None or args.workload != "hall":
        return None, None
    per_bounce = prof["hbm_bytes_per_bounce"]
    where = "this workload" if same else "the 1e6-ray hall"
    return (per_bounce * tot["oct"] / max(tot["n_oct"], 1),
            f"{per_bounce:.2f} HBM B per octave bounce measured on {where} ('{prof['command']}', "
            f"{os.path.relpath(traffic_file(args.precision), ROOT)}) x this run's octave bounces per launch")


def valu_issue(args, tot, n_rays):
    """The resource that binds k_octave: VALU issue.  The profiled VALU
    instructions per octave bounce (rocprofv3 SQ_INSTS_VALU of the same
    workload), at 4 cycles per wave64 instruction, times this run's octave
    bounces per second of kernel time, over the cycles of 1,024 SIMDs at
    2.4 GHz.  None unless this run is the profiled workload."""
    prof, same = load_profile(args, n_rays)
    ipb = prof.get("valu_instr_per_oct_bounce") if prof else None
    if not same or not ipb or tot["ms_oct"] <= 0:
        return None
    rate = tot["oct"] / (tot["ms_oct"] * 1e-3)
    busy = ipb * 4 * rate / (1024 * 2.4e9)
    lanes = prof.get("valu_lane_utilisation")
    return {"resource": "valu_issue", "frac": busy, "instr_per_bounce": ipb, "lane_use": lanes,
            "lane_frac": busy * lanes if lanes else None,
            "note": "VALU issue busy fraction: profiled VALU instructions per octave bounce x 4 cycles x "
                    "this run's octave bounces per second of k_octave time / (1,024 SIMDs x 2.4 GHz); "
                    "lane_frac also weights it by the profiled active-lane share"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--rays", type=int, default=1_000_000)
    ap.add_argument("--cutoff", type=float, default=3.0)
    ap.add_argument("--precision", choices=["fp32", "fp64", "mixed"], default="fp32")
    ap.add_argument("--grid", type=int, default=None,
                    help="voxels per axis (default: 32 for the hall, automatic for the batch)")
    ap.add_argument("--cpu-rays", type=int, default=5000, help="source rays of the CPU sample")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--batch-rays", type=int, default=0, help="source rays per phase-1 launch (0: automatic)")
    ap.add_argument("--progress", type=float, default=0, metavar="SECONDS",
                    help="trace asynchronously and print the completed source rays every SECONDS")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="pach_tune setting, e.g. dist_lds=0")
    ap.add_argument("--autotune", type=int, default=0, metavar="RAYS",
                    help="run pach_autotune with RAYS sample source rays before the warm-up "
                         "(-1: the tuner sizes its sample to ~2e9 bounces per trial)")
    ap.add_argument("--is-order", type=int, default=3, help="image_sources workload: reflection order")
    ap.add_argument("--is-cpu-objects", type=int, default=16,
                    help="image_sources workload: first objects of the CPU sample's sequences")
    ap.add_argument("--map-increment", type=float, default=0.5,
                    help="map workload: receiver spacing and diameter (m)")
    ap.add_argument("--workload", choices=["hall", "batch", "image_sources", "filters", "map", "jnd"], default="hall",
                    help="hall: the 2k-triangle hall; batch: 256 packed rooms in one launch sequence; jnd: "
                         "time to JND-accurate T30/EDT/C80 on the concert hall, denoised vs plain")
    ap.add_argument("--jnd-ref-rays", type=int, default=1_000_000, help="jnd workload: rays of the reference trace")
    ap.add_argument("--jnd-ladder", default="20000,35000,50000,70000,100000,140000,200000",
                    help="jnd workload: ray counts traced plainly and denoised")
    ap.add_argument("--jnd-share", type=float, default=0.95,
                    help="jnd workload: share of (receiver, band) pairs that must be inside every JND")
    ap.add_argument("--batch-ray-scale", type=float, default=0.01,
                    help="batch workload: scale of each room's log-uniform [1e4, 1e6] ray count")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.calib_copy:
        import torch

        n = args.calib_copy * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy_bytes": n * 4}))
        return
    if args.gpus != 1:
        print("bench.py: this build traces on one GPU only; the multi-GPU split (disjoint "
              "source-ray ranges plus one histogram all-reduce) is not implemented", file=sys.stderr)
        sys.exit(2)

    from pachydermacoustic_universal_amd import _abi as A
    from pachydermacoustic_universal_amd import rooms

    if args.workload == "image_sources":
        return bench_image_sources(args, A, rooms)
    if args.workload == "filters":
        return bench_filters(args, A, rooms)
    if args.workload == "jnd":
        return bench_jnd(args, A, rooms)
    at, c, rho = A.uniform_medium()
    prec = {"fp64": A.PREC_FP64, "mixed": A.PREC_MIXED}.get(args.precision, A.PREC_FP32)
    tr = A.Tracer(0)
    for kv in args.tune:
        k, v = kv.split("=")
        tr.tune(k, int(v))
    sr = 1000
    sct = int(np.floor(args.cutoff * 1000 * sr / 1000))
    if args.workload == "batch":
        members, base_rays = rooms.batch_rooms(seed=3, n=256)
        packed = rooms.pack_rooms(members)
        room = packed.room
        room_rays = np.maximum(1, np.round(np.array(base_rays) * args.batch_ray_scale)).astype(np.int64)
        n_rays = int(room_rays.sum())
        grid = (0, 0, 0) if not args.grid else (args.grid,) * 3
    else:
        room = rooms.fan_hall()
        n_rays = args.rays
        grid = (args.grid or 32,) * 3
    tr.upload_scene(room.verts, room.surf, room.alpha, room.scat, room.trans, grid)
    tr.set_source(room.source, np.ones(8), A.DIR_GEODESIC)
    tr.set_receivers(room.receivers, room.radius, at, c, c * rho, sr, sct)
    map_ctr = None
    if args.workload == "map":
        map_ctr = hall_map(room, args.map_increment)
        tr.set_map(map_ctr, args.map_increment, True, sct)
    if args.workload == "batch":
        tr.set_rooms(packed.sources, np.ones((len(members), 8)), room_rays, packed.surf_count, packed.rec_count)
    trace_rays = 0 if args.workload == "batch" else n_rays
    tuning = None
    if args.autotune:
        t_t = time.perf_counter()
        tuning = tr.autotune(rays=trace_rays, cutoff_m=args.cutoff * c, sample_rays=args.autotune,
                             seed=args.seed + 7, precision=prec)
        tuning["seconds"] = time.perf_counter() - t_t
        print(f"[bench] autotune: {tuning}", file=sys.stderr, flush=True)

    def step(seed):
        if not args.progress:
            tr.trace(rays=trace_rays, cutoff_m=args.cutoff * c, seed=seed, precision=prec, batch_rays=args.batch_rays)
            return tr.stats()
        # a long trace: run it asynchronously and report completed source rays
        tr.trace(rays=trace_rays, cutoff_m=args.cutoff * c, seed=seed, precision=prec, batch_rays=args.batch_rays,
                 async_=True)
        # polled every 20 ms so the timed region ends when the trace does,
        # reported every `progress` seconds
        t_rep = time.perf_counter()
        while tr.poll()["running"]:
            time.sleep(0.02)
            if time.perf_counter() - t_rep >= args.progress:
                t_rep = time.perf_counter()
                p = tr.poll()
                print(f"[bench] {p['rays_done']} of {p['rays_total']} source rays done", file=sys.stderr, flush=True)
        rc = tr.wait()
        if rc != 0:
            raise A.PachError(rc, A.lib().pach_last_error(tr.h).decode())
        return tr.stats()

    for w in range(args.warmup):
        step(args.seed + 1000 + w)
        print(f"[bench] warmup {w} done", file=sys.stderr, flush=True)
    tot = dict(bounces=0, oct=0, broad=0, spawn=0, ms_oct=0.0, n_oct=0, ms_p1=0.0, ms_exact=0.0)
    t0 = time.perf_counter()
    for k in range(args.steps):
        st = step(args.seed + k)
        last = st
        tot["bounces"] += st["bounces"]
        tot["oct"] += st["bounces_late"] + st["bounces_verylate"]
        tot["broad"] += st["bounces_broad"]
        tot["spawn"] += st["spawned"]
        tot["ms_oct"] += st["ms_octave"]
        tot["n_oct"] += st["launches_octave"]
        tot["ms_p1"] += st["ms_phase1"]
        tot["ms_exact"] += st["ms_exact"]
        print(f"[bench] step {k}: {st['bounces']:.3e} bounces in {st['ms_trace']:.1f} ms",
              file=sys.stderr, flush=True)
    dt = time.perf_counter() - t0
    value = tot["bounces"] / dt
    # dominant kernel: the persistent octave-ray kernel
    oct_bytes = tot["oct"] * BYTES_OCT_BOUNCE + tot["spawn"] * BYTES_SPAWN
    achieved = oct_bytes / (tot["ms_oct"] / 1e3) / 1e9 if tot["ms_oct"] > 0 else 0.0
    traffic, traffic_note = measured_traffic(args, tot, n_rays)
    avg_ms = tot["ms_oct"] / max(tot["n_oct"], 1)

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O

        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        med = O.Medium(c=c, atten=at, rho=rho)
        t1 = time.perf_counter()
        if args.workload == "batch":
            # every room traced alone from its own first ray index, its share of the sample
            cb, nr = 0, 0
            base = np.concatenate([[0], np.cumsum(room_rays)])
            for i in range(len(members)):
                k = max(1, int(round(room_rays[i] * args.cpu_rays / n_rays)))
                _, cnt = O.trace(packed.member(i), rays=k, cutoff_s=args.cutoff, seed=args.seed, medium=med,
                                 threads=threads, ray_begin=int(base[i]))
                cb += cnt["bounces_broad"] + cnt["bounces_late"] + cnt["bounces_verylate"]
                nr += k
            what = f"{nr} source rays spread over the 256 rooms in proportion"
        else:
            omap = O.Map(map_ctr, args.map_increment, room.source, med, True, sct) if map_ctr is not None else None
            _, cnt = O.trace(room, rays=args.cpu_rays, cutoff_s=args.cutoff, seed=args.seed,
                             medium=med, grid=grid, threads=threads, map_rx=omap)
            cb = cnt["bounces_broad"] + cnt["bounces_late"] + cnt["bounces_verylate"]
            what = f"{args.cpu_rays} source rays of the same {'map' if omap else 'hall'} workload"
        cdt = time.perf_counter() - t1
        cpu = {"value": cb / cdt, "unit": "ray-bounces/s", "cores": threads, "kind": "port",
               "sample": f"{what} ({cb:.3e} bounces, {cdt:.1f} s), C FP64 restatement, "
                         f"{threads} threads on {cpu_model()} (nproc={os.cpu_count()})"}
    tr_grid = list(tuning["grid"]) if tuning else list(grid)
    if args.workload == "batch":
        workload_name = (f"256 packed rooms (shoeboxes 3-30 m and convex hulls of 12-500 tris), "
                         f"4 receivers each, {n_rays} rays (log-uniform [1e4, 1e6] x {args.batch_ray_scale}), "
                         f"{args.cutoff} s cut-off")
    elif args.workload == "map":
        workload_name = (f"seeded polyhedral hall, ~2000 tris, {len(map_ctr)} mapping receivers "
                         f"({args.map_increment} m grid over the audience area, 1.2 m up), {n_rays} rays, "
                         f"{args.cutoff} s cut-off, {tr_grid[0]}x{tr_grid[1]}x{tr_grid[2]} grid")
    else:
        workload_name = (f"seeded polyhedral hall, ~2000 tris, 16 receivers, {n_rays} rays, "
                         f"{args.cutoff} s cut-off, {tr_grid[0]}x{tr_grid[1]}x{tr_grid[2]} grid")
    out = {
        "metric": "ray-bounces/s", "value": value, "unit": "ray-bounces/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPES[args.precision], "data": "synthetic",
        "config": {"workload": workload_name,
                   "rays": n_rays, "triangles": room.n_tris, "receivers": int(room.receivers.shape[0]),
                   "grid": tr_grid, "tuning": tuning,
                   "bounces_per_step": tot["bounces"] / args.steps,
                   "walk": {k: last[k] for k in ("loop_steps", "voxels", "tri_tests", "shade_passes",
                                                 "lanes_shaded", "deposits", "spills", "exact_rewalks",
                                                 "exact_walks")},
                   "rays_per_s": n_rays * args.steps / dt},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_note": traffic_note,
                     "achieved_note": f"algorithmic bytes: {BYTES_OCT_BOUNCE} B per octave bounce (ray read "
                                      f"+ written, hit triangle) + {BYTES_SPAWN} B per spawned ray",
                     "measured_gbs": (traffic / (avg_ms * 1e-3) / 1e9) if traffic and avg_ms > 0 else None,
                     "kernel": "k_octave", "avg_launch_ms": avg_ms,
                     "launches": tot["n_oct"], "binding": valu_issue(args, tot, n_rays),
                     "ms_fp64_fallback": tot["ms_exact"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    tr.close()


def hall_map(room, inc):
    """Mapping receivers of the hall: a square grid of spacing inc at the
    height of its receivers, over their plan extent."""
    R = np.asarray(room.receivers)
    lo, hi = R.min(0), R.max(0)
    xs = np.arange(lo[0], hi[0] + 1e-9, inc)
    ys = np.arange(lo[1], hi[1] + 1e-9, inc)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), np.full(X.size, R[:, 2].mean())], 1)


def bench_image_sources(args, A, rooms):
    """Image-source path checks per second: every sequence of 1..is_order
    polygons of the hall (1,016 polygons, each its own object) for each of
    its 16 receivers, validated by k_image_sources (FP64)."""
    at, c, rho = A.uniform_medium()
    room = rooms.fan_hall()
    grid = (args.grid or 32,) * 3
    tr = A.Tracer(0)
    tr.upload_scene(room.verts, room.surf, room.alpha, room.scat, room.trans, grid)
    tr.set_source(room.source, np.ones(8), A.DIR_GEODESIC)
    tr.set_receivers(room.receivers, room.radius, at, c, c * rho, 1000, 1000)
    S, R, n = room.n_surfaces, int(room.receivers.shape[0]), args.is_order
    for w in range(args.warmup):
        tr.image_sources(n)
        print(f"[bench] warmup {w} done", file=sys.stderr, flush=True)
    tot = dict(checks=0, ms=0.0, launches=0, paths=0)
    t0 = time.perf_counter()
    for k in range(args.steps):
        P = tr.image_sources(n)
        st = tr.image_source_stats()
        for key in ("checks", "launches", "paths"):
            tot[key] += st[key]
        tot["ms"] += st["ms_kernel"]
        print(f"[bench] step {k}: {st['checks']:.3e} checks, {len(P)} paths, kernel {st['ms_kernel']:.1f} ms",
              file=sys.stderr, flush=True)
    dt = time.perf_counter() - t0
    # The tables a check reads (planes, receivers, one object's triangles) stay
    # in cache, so the kernel is bound by FP64 issue, not HBM.  Algorithmic
    # FP64 flops per check of a k-reflection sequence: k mirror images (17
    # each), the ray direction (3) and the division-free triangle test (52)
    # on each of the last object's triangles.
    tris_per_obj = room.n_tris / S
    per_step = sum(R * S ** k * (17 * k + 3 + 52 * tris_per_obj) for k in range(1, n + 1))
    achieved = per_step * args.steps / (tot["ms"] / 1e3) / 1e12
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O

        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        med = O.Medium(c=c, atten=at, rho=rho)
        q1 = min(args.is_cpu_objects, S)
        sc = O.Scene(room, grid=grid)
        t1 = time.perf_counter()
        found = sc.image_source_count(n, room.source, np.ones(8), room.receivers, med, 0, q1, threads)
        cdt = time.perf_counter() - t1
        checks = q1 * R * sum(S ** k for k in range(0, n))
        cpu = {"value": checks / cdt, "unit": "path checks/s", "cores": threads, "kind": "port",
               "sample": f"sequences starting with polygons 0..{q1 - 1} ({checks:.3e} checks, {found} paths, "
                         f"{cdt:.1f} s), C FP64 restatement, {threads} threads on {cpu_model()} "
                         f"(nproc={os.cpu_count()})"}
    out = {
        "metric": "image-source path checks/s", "value": tot["checks"] / dt, "unit": "path checks/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"image sources of the seeded polyhedral hall: every sequence of 1..{n} of its "
                               f"{S} polygons x {R} receivers, {grid[0]}^3 grid",
                   "checks_per_step": tot["checks"] / args.steps, "paths_per_step": tot["paths"] / args.steps},
        "roofline": {"bound": "fp64", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None, "kernel": "k_is_screen + k_image_sources",
                     "avg_launch_ms": tot["ms"] / max(tot["launches"], 1), "launches": tot["launches"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    tr.close()


def bench_filters(args, A, rooms):
    """Auralisation filters per second: the 7 filters (omni and 6 directional,
    Directional_Histogram.Create_Filter) of each of the hall's 16 receivers
    from a traced 3 s, 1 kHz histogram (SampleCT 3,000), one call of
    pach_receiver_filters per step (k_etc_conv, FP64)."""
    at, c, rho = A.uniform_medium()
    room = rooms.fan_hall()
    tr = A.Tracer(0)
    A.trace_room(tr, room, rays=20000, cutoff_s=args.cutoff, seed=args.seed, grid=(32, 32, 32))
    R, T = int(room.receivers.shape[0]), tr.sample_ct
    swl = np.full(8, 120.0)
    for w in range(args.warmup):
        tr.receiver_filters(swl)
        print(f"[bench] warmup {w} done", file=sys.stderr, flush=True)
    ms_k, t0 = 0.0, time.perf_counter()
    for k in range(args.steps):
        tr.receiver_filters(swl)
        st = tr.filter_stats()
        ms_k += st["ms_kernel"]
        print(f"[bench] step {k}: {st['channels']} channels, kernel {st['ms_kernel']:.2f} ms, "
              f"call {st['ms_total']:.1f} ms", file=sys.stderr, flush=True)
    dt = time.perf_counter() - t0
    # algorithmic FP64 flops: every envelope sample meets every carrier tap
    # once per band (9 bands incl. the 16 kHz one), one multiply-add each
    flops = 2.0 * 9 * T * A.CARRIER_LEN * 7 * R
    achieved = flops * args.steps / (ms_k / 1e3) / 1e12
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import filters_ref as FR

        pl = tr.planes(scaled=True)
        n_rec = 2
        t1 = time.perf_counter()
        for r in range(n_rec):
            FR.receiver_filters(pl[0, r], pl[1, r], pl[2:5, r], pl[5:8, r], swl, 1000)
        cdt = time.perf_counter() - t1
        cpu = {"value": n_rec / cdt, "unit": "receiver filter sets/s", "cores": 1, "kind": "port",
               "sample": f"{n_rec} receivers' 7 filters, numpy FP64 FFT convolution as the reference computes "
                         f"them ({cdt:.2f} s) on {cpu_model()}"}
    out = {
        "metric": "receiver filter sets/s", "value": R * args.steps / dt, "unit": "receiver filter sets/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"auralisation filters (omni + 6 directional) of the hall's {R} receivers from "
                               f"{T}-sample 1 kHz histograms of a 20,000-ray trace",
                   "channels_per_step": 7 * R, "kernel_ms_per_step": ms_k / args.steps},
        "roofline": {"bound": "fp64", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None, "kernel": "k_etc_conv",
                     "avg_launch_ms": ms_k / args.steps, "launches": args.steps},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    tr.close()


def bench_jnd(args, A, rooms):
    """Time to JND accuracy on the procedural concert hall (50,120 triangles,
    64 receivers, 3 s): a plain reference trace of --jnd-ref-rays, then every
    ladder count traced plainly and denoised by the shipped network (IS_Order
    0, the direct sound added as ETCurve does, C80 from its arrival).  A trace
    is JND-accurate when T30, EDT and C80 each have at least --jnd-share of the
    64 x 8 (receiver, band) pairs inside 5 % / 5 % / 1 dB of the reference.
    `value` is the wall time of the cheapest JND-accurate denoised trace
    (trace + network); `plain_s` that of the cheapest plain one; both also as
    a fraction of the 1e7-ray trace's time, extrapolated linearly in rays from
    the reference trace."""
    from pachydermacoustic_universal_amd import denoise as DN

    at, c, rho = A.uniform_medium()
    room = rooms.concert_hall()
    tr = A.Tracer(0)
    tr.upload_scene(room.verts, room.surf, room.alpha, room.scat, room.trans, (0, 0, 0))
    tr.set_source(room.source, np.ones(8), A.DIR_GEODESIC)
    tr.set_receivers(room.receivers, room.radius, at, c, c * rho, 1000, int(round(args.cutoff * 1000)))
    dE, dbin = DN.direct_part(room)
    desc = DN.descriptors(room)

    def etc(E):
        return DN.with_direct(E, dE, dbin)

    def trace(rays, seed):
        t0 = time.perf_counter()
        tr.trace(rays=rays, cutoff_m=args.cutoff * c, seed=seed, is_order=0)
        wall = time.perf_counter() - t0
        return tr.histogram(0, scaled=True), wall, tr.stats()

    for w in range(args.warmup):
        trace(20000, 99)
        print(f"[bench] warmup {w} done", file=sys.stderr, flush=True)
    E_ref, ref_wall, st = trace(args.jnd_ref_rays, 1)
    print(f"[bench] reference {args.jnd_ref_rays} rays: {ref_wall:.1f} s", file=sys.stderr, flush=True)
    oct_b = st["bounces_late"] + st["bounces_verylate"]
    achieved = (oct_b * BYTES_OCT_BOUNCE + st["spawned"] * BYTES_SPAWN) / (st["ms_octave"] / 1e3) / 1e9
    P_ref = DN.decay_params(etc(E_ref), start_bin=dbin)
    dn = DN.Denoiser.load(max_batch=64)
    rows = []
    t_all = time.perf_counter()
    for rays in sorted(int(float(r)) for r in args.jnd_ladder.split(",") if r):
        E, wall, _ = trace(rays, 1000 + rays)
        t0 = time.perf_counter()
        E_dn = dn.denoise(E, desc, rays=rays)
        t_dn = time.perf_counter() - t0
        row = {"rays": rays, "plain_s": wall, "denoised_s": wall + t_dn}
        for kind, X in (("plain", E), ("denoised", E_dn)):
            e = DN.summarize(DN.errors(DN.decay_params(etc(X), start_bin=dbin), P_ref))
            row[kind] = {k: {"mean": e[k]["mean"], "within": e[k]["within"]} for k in e}
        rows.append(row)
        print(f"[bench] {rays} rays: plain {wall:.2f} s, shares within JND "
              f"{[round(row['plain'][k]['within'], 3) for k in ('t30', 'edt', 'c80')]}; denoised "
              f"{[round(row['denoised'][k]['within'], 3) for k in ('t30', 'edt', 'c80')]}", file=sys.stderr, flush=True)
    dn.close()
    tr.close()
    dt = time.perf_counter() - t_all
    t_1e7 = ref_wall * 1e7 / args.jnd_ref_rays

    def first(kind):
        for r in rows:
            if min(r[kind][k]["within"] for k in ("t30", "edt", "c80")) >= args.jnd_share:
                return r
        return None

    fp, fd = first("plain"), first("denoised")
    out = {
        "metric": f"time to JND-accurate T30/EDT/C80 ({100 * args.jnd_share:.0f} % of pairs), denoised",
        "value": fd["denoised_s"] if fd else None, "unit": "s", "n_gpus": 1, "steps": 1, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 trace (f64 histograms), bf16 network", "data": "synthetic",
        "config": {"workload": "procedural concert hall, 50,120 tris, 64 receivers, 3 s cut-off, IS_Order 0 "
                               "(direct sound added), JND 5 % T30 / 5 % EDT / 1 dB C80",
                   "reference_rays": args.jnd_ref_rays, "reference_s": ref_wall,
                   "ladder": [r["rays"] for r in rows]},
        "plain_s": fp["plain_s"] if fp else None, "plain_rays": fp["rays"] if fp else None,
        "denoised_rays": fd["rays"] if fd else None,
        "speedup_vs_plain": (fp["plain_s"] / fd["denoised_s"]) if (fp and fd) else None,
        "fraction_of_1e7_trace": {"plain": fp["plain_s"] / t_1e7 if fp else None,
                                  "denoised": fd["denoised_s"] / t_1e7 if fd else None,
                                  "t_1e7_s": t_1e7, "note": "1e7-ray time extrapolated linearly from the reference"},
        "ladder": rows,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_octave",
                     "avg_launch_ms": st["ms_octave"] / max(st["launches_octave"], 1),
                     "launches": st["launches_octave"],
                     "achieved_note": "the reference trace's octave kernel, algorithmic bytes as the hall workload"},
        "cpu_baseline": None,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
