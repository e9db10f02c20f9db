"""Benchmark: Gaussian-substeps/s of the XPBD step on one MI355X.

Workload (the 1M-Gaussian roofline recipe): 1M Gaussians uniform in a
0.5 x 0.3 x 0.2 m ellipsoid on the z=0 plane, symmetric kNN (k=12) distance
constraints (compliance 1e-6), gravity and ground, two capsule grippers
(r 0.02 m, half-length 0.04 m) that attach 500 nodes each, pull 0.1 m and
twist 30 deg over 60 frames then release; dt = 1/60 s per frame as 10
substeps x 4 iterations.  One bench step = one frame = 10 substeps.

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
HBM_PEAK_GBS = 8000.0
SUBSTEPS = 10
ITERS = 4


PROFILE = os.path.join(ROOT, "profiles", "distance_jacobi_1M.md")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic():
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (scaled by the streaming calibration)."""
    try:
        for line in open(PROFILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def configure(sim):
    sim.set_solver(iters=ITERS, dt=1.0 / (60.0 * SUBSTEPS), under_relax=0.6, vel_damping=0.01,
                   dist_compliance=1e-6)
    sim.set_constraints(distance=True, shape=False, angle=False, volume=False)
    sim.set_ground(True, (0.0, 0.0, 1.0), 0.0, friction=0.5, restitution=0.0, gravity=9.81)


def run_frame(sim, script, frame, first):
    t, p, d = script.colliders(frame)
    sim.set_colliders(t, p, d)
    att = script.attached(frame)
    for s in range(SUBSTEPS):
        if s == (1 if first else 0):
            if att is not None:
                sim.set_attached(att[0], att[1], 1.0)
            elif frame == script.ramp:
                sim.clear_attached()
        sim.step()


def cpu_baseline(scene, script, seconds_budget=20.0):
    """Oracle (serial restatement) on a bounded sample of the same workload."""
    import multiprocessing
    sim = XPBD(ORACLE_LIB)
    sim.create(scene)
    configure(sim)
    t0 = time.perf_counter()
    substeps = 0
    frame = 0
    while True:
        run_frame(sim, script, frame, frame == 0)
        substeps += SUBSTEPS
        frame += 1
        if time.perf_counter() - t0 > seconds_budget or frame >= 3:
            break
    dt = time.perf_counter() - t0
    sim.destroy()
    n = scene["pos"].shape[0]
    return {"value": n * substeps / dt, "unit": "Gaussian-substeps/s", "cores": 1, "kind": "port",
            "sample": f"oracle, same {n}-Gaussian scene and script, first {frame} frame(s) = {substeps} substeps, "
                      f"single thread, {dt:.1f} s; host has {multiprocessing.cpu_count()} logical CPUs"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate", action="store_true",
                    help="only run a streaming kernel of known traffic (256 MiB read + 256 MiB written, x3) "
                         "to scale the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.calibrate:
        import torch
        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.mul(a, 2.0, out=b)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "mul", "bytes_read": a.numel() * 4, "bytes_written": b.numel() * 4}))
        return

    log(f"generating {args.n}-Gaussian scene")
    t = time.perf_counter()
    scene = scenes.make_ellipsoid_scene(args.n, 12, seed=3)
    script = scenes.GripperScript(scene, seed=7)
    n = scene["pos"].shape[0]
    e = scene["edges"].shape[0]
    log(f"scene ready in {time.perf_counter() - t:.1f}s: N={n} E={e} M={2 * e}")

    sim = XPBD(LIB)
    if sim.backend != "hip-gfx950":
        raise SystemExit("bench.py needs the HIP library")
    sim.create(scene)
    configure(sim)
    frame = 0
    for _ in range(args.warmup):
        run_frame(sim, script, frame, frame == 0)
        frame += 1
    log(f"warmup {args.warmup} frames done")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run_frame(sim, script, frame, frame == 0)
        frame += 1
    t1 = time.perf_counter()
    elapsed = t1 - t0
    value = n * SUBSTEPS * args.steps / elapsed
    log(f"timed {args.steps} frames in {elapsed:.3f}s -> {value / 1e9:.3f} G Gaussian-substeps/s")

    # per-kernel HIP-event timing on the library stream, same state and script
    prof = sim.ext_profile(SUBSTEPS)
    log("kernel ms (total, launches): " + json.dumps(prof))
    key = "distance_jacobi_mid"
    tot, cnt = prof[key]
    avg_s = tot / cnt / 1e3
    m = 2 * e
    alg_bytes = 36 * n + 20 * m      # predicted pos in/out + row offsets + 5 streams per directed entry
    achieved = alg_bytes / avg_s / 1e9
    sim.destroy()

    cpu = None if args.no_cpu else cpu_baseline(scene, script)
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
        "config": {"workload": f"1M-Gaussian ellipsoid, kNN k=12 distance, ground+gravity, 2 capsule grippers "
                               f"pull/twist; {SUBSTEPS} substeps x {ITERS} iters per step; plain C-ABI path "
                               f"(xpbd_step per substep)",
                   "gaussians": n, "edges": e, "substeps_per_step": SUBSTEPS, "iters": ITERS},
        "roofline": {"bound": "hbm", "kernel": key, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "traffic_source": os.path.relpath(PROFILE, ROOT),
                     "alg_bytes_per_launch": alg_bytes, "avg_launch_us": avg_s * 1e6},
        "cpu_baseline": cpu,
        "kernel_ms": prof,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
