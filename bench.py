"""Benchmark: Gaussian-substeps/s of the XPBD step on one MI355X.

Default workload ("full", the 1M roofline configuration): 1M Gaussians uniform
in a 0.5 x 0.3 x 0.2 m ellipsoid on the z=0 plane (PCG64 seed 3), symmetric
kNN (k=12) distance constraints (compliance 1e-6), one volume cluster per node
(the node and its 3 nearest neighbours), k-means shape clusters of ~64 nodes
(shape blend 0.1), self-collision within the body (2^23 hash cells, the
reference's cell rule), gravity and ground, two capsule grippers (r 0.02 m,
half-length 0.04 m) that attach 500 nodes each, pull 0.1 m and twist 30 deg
over 60 frames then release, and the per-Gaussian scale/orientation update
once per frame (outputs left in device memory).  dt = 1/60 s per frame as 10
substeps x 4 iterations; one bench step = one frame.  "--workload distance"
keeps only distance + ground + grippers.

The default run also measures the batched 64-object mix ("--workload
batched": its Gaussian-substeps/s over one object per call, both sides
autotuned) and records it under "batched_mix" in the same line
(--no-batched skips it).

Prints one JSON line (see the repository contract).  --gpus N>1 is refused:
this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "3dgs-isaac-sim-xpbd_amd")
sys.path.insert(0, PKG)

import scenes  # noqa: E402
from xpbd_bridge import XPBD  # noqa: E402

LIB = os.path.join(PKG, "libxpbd_isaac.so")
ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle_xpbd.so")
PROFILES = {"full": os.path.join(ROOT, "profiles", "full_recipe_1M.md"),
            "distance": os.path.join(ROOT, "profiles", "distance_jacobi_1M.md")}
HBM_PEAK_GBS = 8000.0
SUBSTEPS = 10
ITERS = 4


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of the workload (scaled by the streaming
    calibration)."""
    try:
        for line in open(PROFILES[workload]):
            if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def configure(sim, workload, recipe):
    if workload == "full":
        scenes.configure_full(sim, recipe, iters=ITERS, substeps=SUBSTEPS)
    else:
        sim.set_solver(iters=ITERS, dt=1.0 / (60.0 * SUBSTEPS), under_relax=0.6, vel_damping=0.01,
                       dist_compliance=1e-6)
        sim.set_constraints(distance=True, shape=False, angle=False, volume=False)
        sim.set_ground(True, (0.0, 0.0, 1.0), 0.0, friction=0.5, restitution=0.0, gravity=9.81)
        sim.set_self_collision(False)


def run_frame(sim, script, frame, first, shapes, eager=False):
    """One frame of the gripper script.  eager: the substeps run through the
    library's event profile (ordinary launches instead of the captured graph)
    with the same parameters, for counter collection."""
    t, p, d = script.colliders(frame)
    sim.set_colliders(t, p, d)
    att = script.attached(frame)
    s = 0
    while s < SUBSTEPS:
        if s == (1 if first else 0):
            if att is not None:
                sim.set_attached(att[0], att[1], 1.0)
            elif frame == script.ramp:
                sim.clear_attached()
        if eager and (s > 0 or not first):
            sim.ext_profile(SUBSTEPS - s)   # the rest of the frame, no parameter change inside it
            s = SUBSTEPS
        else:
            sim.step()
            s += 1
    if shapes:
        sim.shapes_device(1e-2)


def summarize_pmc(pattern, last):
    """Per kernel, the mean of each counter over its dispatches among the last
    `last` dispatches of rocprofv3 --pmc CSV output (counter_collection.csv)."""
    import csv
    import glob
    import collections
    for path in sorted(glob.glob(pattern, recursive=True)):
        rows = list(csv.DictReader(open(path)))
        if not rows:
            continue
        disp = sorted({int(r["Dispatch_Id"]) for r in rows})
        keep = set(disp[-last:])
        acc = collections.defaultdict(lambda: collections.defaultdict(float))
        seen = collections.defaultdict(set)
        for r in rows:
            d = int(r["Dispatch_Id"])
            if d in keep:
                acc[r["Kernel_Name"]][r["Counter_Name"]] += float(r["Counter_Value"])
                seen[r["Kernel_Name"]].add(d)
        print(f"# {path}: last {len(keep)} dispatches")
        for k in sorted(acc, key=lambda k: -sum(acc[k].values())):
            n = len(seen[k])
            vals = " ".join(f"{c}={v / n:.6g}" for c, v in sorted(acc[k].items()))
            print(f"{k.split('(')[0]} | dispatches {n} | {vals}")


def cpu_baseline(scene, script, workload, recipe, seconds_budget=12.0, full_sample=100_000):
    """Oracle (serial restatement) on a bounded sample of the same workload,
    timed twice: one thread, then its order-independent per-node and
    per-cluster loops on every host core this process may use (at most 16).
    Each leg runs the first substeps until its time budget is spent.  The full
    recipe is sampled on a smaller scene from the same generator, recipe and
    gripper script: at 1M the oracle's serial self-collision walk over the
    capped cell (search + 16 d_c, reached while the grippers bite) takes
    minutes per substep."""
    import ctypes
    if workload == "full":
        scene = scenes.make_ellipsoid_scene(full_sample, 12, seed=3)
        script = scenes.GripperScript(scene, seed=7)
        recipe = scenes.full_recipe(scene, seed=3)
    n = scene["pos"].shape[0]
    cores = min(len(os.sched_getaffinity(0)), 16)

    def leg(threads):
        ctypes.CDLL(ORACLE_LIB).xpbd_oracle_set_threads(threads)
        log(f"cpu baseline: oracle on {n} Gaussians ({workload}), {threads} thread(s)")
        sim = XPBD(ORACLE_LIB)
        sim.create(scene)
        configure(sim, workload, recipe)
        # one untimed substep builds the oracle's topology caches
        t, p, d = script.colliders(0)
        sim.set_colliders(t, p, d)
        sim.step()
        t0 = time.perf_counter()
        substeps = 0
        frame = 0
        done = False
        while not done:
            t, p, d = script.colliders(frame)
            sim.set_colliders(t, p, d)
            att = script.attached(frame)
            for s in range(SUBSTEPS):
                if s == (1 if frame == 0 else 0) and att is not None:
                    sim.set_attached(att[0], att[1], 1.0)
                sim.step()
                substeps += 1
                log(f"cpu baseline: {threads} thread(s), substep {substeps} at {time.perf_counter() - t0:.1f}s")
                if time.perf_counter() - t0 > seconds_budget:
                    done = True
                    break
            frame += 1
            if frame >= 3:
                break
        dt = time.perf_counter() - t0
        sim.destroy()
        ctypes.CDLL(ORACLE_LIB).xpbd_oracle_set_threads(1)
        return n * substeps / dt, substeps, dt

    one, s1, d1 = leg(1)
    many, sm, dm = leg(cores)
    return {"value": many, "unit": "Gaussian-substeps/s", "cores": cores, "kind": "port",
            "sample": f"oracle, {n}-Gaussian scene from the same generator, recipe and gripper script, first "
                      f"substeps after one untimed substep: {sm} substeps in {dm:.1f} s on {cores} threads "
                      f"(OpenMP over the order-independent per-node and per-cluster loops; sorting, colouring, "
                      f"atomics-order sums and the candidate-list scan order stay serial); {s1} substeps in "
                      f"{d1:.1f} s on one thread",
            "single_thread": {"value": one, "cores": 1}}


def kernel_bytes(workload, n, m, sc_table, sc_passes, vol, sc_cnt=0, sc_touched=0):
    """Algorithmic bytes per launch (compulsory traffic) of each profiled kernel
    label: N nodes, M directed distance entries, L leaders, S members,
    R = S + L reverse entries, T hash cells, cnt self-collision candidates,
    touched nodes with at least one candidate."""
    L, S = vol
    R = S + L
    out = {
        "predict": 52 * n,
        "distance_jacobi_first": 36 * n + 20 * m,
        "distance_jacobi_mid": 36 * n + 20 * m,
        "distance_jacobi_last": 36 * n + 20 * m,
    }
    if workload == "full":
        out.update({
            "volume_solve_thread": 100 * L + 4 * S,
            "volume_solve_wave": 100 * L + 4 * S,
            "volume_apply": 56 * n + 4 * R,
            "shape_match": 60 * n,
            "selfcol_build": 150 * n + 4 * sc_table + 4 * sc_cnt + 16 * n * sc_passes,
            "selfcol_iteration": 4 * n + 4 * sc_cnt + 60 * sc_touched,
            "contact_push_commit": 64 * n,
        })
    return out


def plain_frame(single, script, f):
    """One frame of one object the way the singleton ABI runs it: every
    substep a synchronous call (native/xpbd_capi.cu:218-225), attachments
    after the first step on the first frame, then the shape update."""
    t, p, d = script.colliders(f)
    single.set_colliders(0, t, p, d)
    att = script.attached(f)
    for s in range(SUBSTEPS):
        if s == (1 if f == 0 else 0):
            if att is not None:
                single.set_attached([(0, att[0], att[1])])
            elif f == script.ramp:
                single.set_attached([(0, (), ())])
        single.step(1)
    # xpbd_compute_shapes with the outputs left in device memory (null host
    # pointers): one object, the deformed count read back
    single.lib.xpbd_batch_compute_shapes(single.h, 0, 1e-2, None, None, single.counts[0])


def bench_batched(args):
    """Many objects in one batch against the same objects one per call."""
    import serving
    log(f"generating {args.objects} objects, sizes log-uniform in [5k, 200k]")
    t = time.perf_counter()
    objs = scenes.object_mix(args.objects, 5_000, 200_000, seed=11)
    reqs = [serving.Request(j, *o, life=10 ** 6) for j, o in enumerate(objs)]
    total = sum(r.n for r in reqs)
    log(f"objects ready in {time.perf_counter() - t:.1f}s: {len(reqs)} objects, {total} Gaussians")

    b = serving.build_batch(reqs)
    frame = 0
    for _ in range(max(args.warmup, 1)):
        serving.step_frame(b, [(r.script, frame) for r in reqs])
        b.sync()
        frame += 1
    tune = b.autotune(substeps=2, hash_tables=args.tune_hash)
    log("autotune: " + json.dumps(tune))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        serving.step_frame(b, [(r.script, frame) for r in reqs])
        b.sync()
        frame += 1
    dt_b = time.perf_counter() - t0
    rate_b = total * SUBSTEPS * args.steps / dt_b
    log(f"batched: {args.steps} frames in {dt_b:.3f}s -> {rate_b / 1e6:.1f} M Gaussian-substeps/s")
    prof = b.profile(SUBSTEPS)
    tuning = b.tuning()
    b.close()

    singles = []
    for r in reqs:
        s = serving.build_batch([r])
        singles.append(s)
    frame = 0
    for _ in range(max(args.warmup, 1)):
        for s, r in zip(singles, reqs):
            plain_frame(s, r.script, frame)
        frame += 1
    # the same autotuner on every single-object batch (launch shapes and hash
    # table sizes as on the batched side), so that both sides are tuned
    t = time.perf_counter()
    plain_tunes = [s.autotune(substeps=2, hash_tables=args.tune_hash) for s in singles] if not args.no_tune_plain else []
    if plain_tunes:
        log(f"plain side autotuned ({len(singles)} objects) in {time.perf_counter() - t:.1f}s")
    plain_steps = max(1, args.steps // 2)
    t0 = time.perf_counter()
    for _ in range(plain_steps):
        for s, r in zip(singles, reqs):
            plain_frame(s, r.script, frame)
        frame += 1
    dt_p = time.perf_counter() - t0
    rate_p = total * SUBSTEPS * plain_steps / dt_p
    log(f"plain: {plain_steps} frames in {dt_p:.3f}s -> {rate_p / 1e6:.1f} M Gaussian-substeps/s")
    for s in singles:
        s.close()
    out = {
        "metric": "batched vs plain GPU speedup",
        "value": rate_b / rate_p,
        "unit": "x",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt_b / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{len(reqs)} objects, N log-uniform in [5k, 200k] (seed 11), ellipsoids with semi-axes "
                               f"U[0.05, 0.30] m, full recipe per object (kNN k=12 distance, node+3NN volume, k-means "
                               f"shape clusters of 64, self-collision within the body, nextpow2(2N) cells), own "
                               f"gripper pair per object, shapes every frame; {SUBSTEPS} substeps x {ITERS} iters",
                   "objects": len(reqs), "gaussians": total},
        "batched": {"gaussian_substeps_per_s": rate_b, "ms_per_frame": 1e3 * dt_b / args.steps, "tuning": tuning,
                    "autotune": tune, "kernels_ms": prof},
        "plain": {"gaussian_substeps_per_s": rate_p, "ms_per_frame": 1e3 * dt_p / plain_steps, "frames": plain_steps,
                  "path": "one single-object batch per object, one synchronous call per substep, objects in turn",
                  "autotuned": bool(plain_tunes)},
    }
    return out


def bench_served(args):
    """The request loop: Poisson arrivals, HBM-budget admission, one frame of
    every live object per tick."""
    import serving
    sys.setswitchinterval(0.0005)   # the builder thread must not hold the tick loop back
    log(f"preparing {args.templates} object templates, N log-uniform in [2k, 500k]")
    t = time.perf_counter()
    pool = serving.TemplatePool(args.templates, 2_000, 500_000, seed=5)
    reqs = serving.poisson_stream(pool, 20.0, args.ticks, seed=3, steady_start=not args.cold)
    log(f"templates ready in {time.perf_counter() - t:.1f}s; {len(reqs)} requests "
        f"({sum(r.arrival == 0 for r in reqs)} at tick 0)")
    budget_ms = args.tick_budget if args.tick_budget > 0 else None
    srv = serving.Server(hbm_budget=200e9, tick_budget_ms=budget_ms, max_segments=args.max_segments)
    last = [time.perf_counter()]

    def progress(rec):
        if time.perf_counter() - last[0] > 20:
            last[0] = time.perf_counter()
            log(f"tick {rec['tick']}: {rec['ms']:.1f} ms, {rec['live_objects']} live objects, "
                f"{rec['live_gaussians']} Gaussians, {rec['segments']} batches, {rec['queued']} queued")

    t0 = time.perf_counter()
    srv.run([r for r in reqs], max_ticks=args.ticks, on_tick=progress)
    wall = time.perf_counter() - t0
    srv.close()
    recs = srv.log[args.skip:]
    ms = np.array([r["ms"] for r in recs])
    gs = np.array([r["live_gaussians"] for r in recs], np.float64)
    started = [r for r in reqs if r.started is not None]
    waits = np.array([(r.started - r.arrival) for r in started], np.float64)
    out = {
        "metric": "p99 frame latency",
        "value": float(np.percentile(ms, 99)),
        "unit": "ms",
        "n_gpus": 1,
        "steps": len(recs),
        "warmup": args.skip,
        "ms_per_step": float(ms.mean()),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "served stream: Poisson arrivals 20/s, N log-uniform [2k, 500k] "
                               f"({args.templates} templates), life U[60, 600] frames, full recipe, own gripper "
                               "script per request, 200 GB HBM budget, one frame of every live object per tick"
                               + (f", admission also capped by a {budget_ms} ms tick budget" if budget_ms else ""),
                   "ticks": args.ticks, "requests": len(reqs), "steady_start": not args.cold},
        "p50_ms": float(np.percentile(ms, 50)),
        "p99_ms": float(np.percentile(ms, 99)),
        "max_ms": float(ms.max()),
        "gaussian_substeps_per_s": float(gs.sum() * SUBSTEPS / (ms.sum() / 1e3)),
        "mean_live_objects": float(np.mean([r["live_objects"] for r in recs])),
        "mean_live_gaussians": float(gs.mean()),
        "admission_wait_ticks_p50": float(np.percentile(waits, 50)) if waits.size else None,
        "admission_wait_ticks_p99": float(np.percentile(waits, 99)) if waits.size else None,
        "merges": srv.merges,
        "wall_s": wall,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--workload", choices=("full", "distance", "batched", "served"), default="full")
    ap.add_argument("--objects", type=int, default=64, help="batched: object count")
    ap.add_argument("--tune-hash", action="store_true", help="batched: let the autotuner size the hash tables")
    ap.add_argument("--no-tune-plain", action="store_true",
                    help="batched: leave the plain side at the default launch shapes (the batched side is autotuned)")
    ap.add_argument("--templates", type=int, default=40, help="served: prepared objects requests draw from")
    ap.add_argument("--ticks", type=int, default=150, help="served: ticks to run")
    ap.add_argument("--skip", type=int, default=10, help="served: ticks left out of the latency figures")
    ap.add_argument("--tick-budget", type=float, default=0.0, help="served: admit only while ticks stay under this (ms)")
    ap.add_argument("--max-segments", type=int, default=6, help="served: batches before merging")
    ap.add_argument("--cold", action="store_true", help="served: start empty instead of at the steady population")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true",
                    help="full: skip the batched 64-object mix that the default run also measures")
    ap.add_argument("--tune", action="store_true",
                    help="full/distance: autotune the launch shapes after the warmup frames (the defaults are "
                         "the ones measured best at the compressed frames of the gripper script, which the "
                         "light warmup state does not represent)")
    ap.add_argument("--group-lanes", type=int, default=None,
                    help="full: fine candidate scan (1 = one lane per node, 0 = one wave per node; "
                         "default: the library's automatic choice)")
    ap.add_argument("--eager", action="store_true",
                    help="full/distance: run the substeps as ordinary launches (the event profile) instead of the "
                         "captured graph, for rocprofv3 --pmc runs (the timed figure is then not the headline)")
    ap.add_argument("--pmc-summary", default=None, metavar="GLOB",
                    help="only summarise rocprofv3 --pmc CSV files (counter_collection.csv) matching GLOB")
    ap.add_argument("--last-dispatches", type=int, default=800,
                    help="with --pmc-summary: dispatches at the end of the run to average over")
    ap.add_argument("--calibrate", action="store_true",
                    help="only run a streaming kernel of known traffic (256 MiB read + 256 MiB written, x3) "
                         "to scale the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.pmc_summary:
        summarize_pmc(args.pmc_summary, args.last_dispatches)
        return
    if args.calibrate:
        import torch
        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.mul(a, 2.0, out=b)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "mul", "bytes_read": a.numel() * 4, "bytes_written": b.numel() * 4}))
        return
    if args.workload == "batched":
        print(json.dumps(bench_batched(args)), flush=True)
        return
    if args.workload == "served":
        return bench_served(args)

    log(f"generating {args.n}-Gaussian scene ({args.workload})")
    t = time.perf_counter()
    scene = scenes.make_ellipsoid_scene(args.n, 12, seed=3)
    script = scenes.GripperScript(scene, seed=7)
    recipe = scenes.full_recipe(scene, seed=3) if args.workload == "full" else None
    n = scene["pos"].shape[0]
    e = scene["edges"].shape[0]
    m = 2 * e
    log(f"scene ready in {time.perf_counter() - t:.1f}s: N={n} E={e} M={m}")

    sim = XPBD(LIB)
    if sim.backend != "hip-gfx950":
        raise SystemExit("bench.py needs the HIP library")
    sim.create(scene)
    configure(sim, args.workload, recipe)
    shapes = args.workload == "full"
    frame = 0
    for _ in range(args.warmup):
        run_frame(sim, script, frame, frame == 0, shapes, args.eager)
        if frame == 0 and args.group_lanes is not None:
            sim.ext_set_group_lanes(args.group_lanes)
        frame += 1
    log(f"warmup {args.warmup} frames done")
    tune = sim.ext_autotune(2) if args.tune else None
    if tune is not None:
        log("autotune: " + json.dumps(tune))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run_frame(sim, script, frame, frame == 0, shapes, args.eager)
        frame += 1
    t1 = time.perf_counter()
    elapsed = t1 - t0
    value = n * SUBSTEPS * args.steps / elapsed
    log(f"timed {args.steps} frames in {elapsed:.3f}s -> {value / 1e9:.3f} G Gaussian-substeps/s")
    stats = sim.self_collision_stats() if shapes else (0, 0, 0)
    grid = sim.ext_self_collision_grid_info() if shapes else (0.0, 0.0, 0.0)
    touched = int((sim.ext_self_collision_lists()[2] > 0).sum()) if shapes else 0
    log(f"self-collision: candidates/active/overflow {stats}, nodes with candidates {touched}, "
        f"cell/d_c/search {grid}")

    # per-kernel HIP-event timing on the library stream, same state and script
    phases = sim.ext_profile(SUBSTEPS)
    log("kernel ms (total, launches): " + json.dumps(phases))
    # the self-collision build's phases (selfcol_build_*) form one launch set
    # priced by one formula
    prof = {}
    for k, (tot, cnt) in phases.items():
        base = "selfcol_build" if k.startswith("selfcol_build") else k
        if base in prof:
            prof[base] = (prof[base][0] + tot, prof[base][1])
        else:
            prof[base] = (tot, cnt)
    vol = (n, 3 * n) if shapes else (0, 0)
    table = recipe["table"] if shapes else 0
    passes = (int(np.ceil(np.log2(table + 1))) + 7) // 8 if shapes else 0
    formulas = kernel_bytes(args.workload, n, m, table, passes, vol, stats[0], touched)
    kernels = {}
    for k, (tot, cnt) in prof.items():
        if k in formulas and cnt > 0:
            avg = tot / cnt / 1e3
            kernels[k] = {"ms_total": tot, "launches": cnt, "avg_us": avg * 1e6, "alg_bytes": formulas[k],
                          "GBps": formulas[k] / avg / 1e9}
    dominant = max(kernels, key=lambda k: kernels[k]["ms_total"])
    dk = kernels[dominant]
    achieved = dk["GBps"]
    sim.destroy()

    cpu = None if args.no_cpu else cpu_baseline(scene, script, args.workload, recipe)
    # the batched many-object mix (its own speedup over one object per call),
    # measured in the default run too so that it is on record beside the 1M line
    batched = None
    if args.workload == "full" and not args.no_batched:
        b = bench_batched(args)
        batched = {"metric": b["metric"], "value": b["value"], "unit": b["unit"], "steps": b["steps"],
                   "config": b["config"], "batched_gaussian_substeps_per_s": b["batched"]["gaussian_substeps_per_s"],
                   "plain_gaussian_substeps_per_s": b["plain"]["gaussian_substeps_per_s"],
                   "both_sides_autotuned": b["plain"]["autotuned"]}
    if args.workload == "full":
        wl = (f"1M-Gaussian ellipsoid full solver: kNN k=12 distance, node+3NN volume clusters, k-means shape "
              f"clusters of 64, self-collision within the body (2^23 cells; cell edge by the reference's rule, "
              f"search + 2 x largest predicted move, not 2 x mean scale), ground+gravity, 2 capsule grippers "
              f"pull/twist, per-frame shape update; {SUBSTEPS} substeps x {ITERS} iters per step; plain C-ABI path")
    else:
        wl = (f"1M-Gaussian ellipsoid, kNN k=12 distance, ground+gravity, 2 capsule grippers pull/twist; "
              f"{SUBSTEPS} substeps x {ITERS} iters per step; plain C-ABI path")
    out = {
        "metric": "Gaussian-substeps/s at 1M",
        "value": value,
        "unit": "Gaussian-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl, "gaussians": n, "edges": e, "substeps_per_step": SUBSTEPS, "iters": ITERS,
                   "shape_clusters": int(recipe["clusters"].max() + 1) if shapes else 0,
                   "selfcol_stats_last": list(stats), "selfcol_touched": touched,
                   "selfcol_cell_dc_search": list(grid), "autotune": tune},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dominant, args.workload),
                     "traffic_source": os.path.relpath(PROFILES[args.workload], ROOT),
                     "alg_bytes_per_launch": dk["alg_bytes"], "avg_launch_us": dk["avg_us"]},
        "cpu_baseline": cpu,
        "kernels": kernels,
        "batched_mix": batched,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
