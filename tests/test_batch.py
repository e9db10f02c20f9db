"""GPU tests of the batched API and the serving loop.

  * objects stepped together in one batch advance bit for bit as each does
    alone through the singleton C ABI (same recipe, script and frames);
  * a batch object against the oracle, per frame from an identical state,
    within 1e-5 of the object's extent;
  * copy_object hands an object to another batch mid-run and it continues
    bit for bit;
  * the autotuner changes launch shapes only: results are bit-identical;
  * the serving loop (arrivals at different ticks, admission, background
    builds, departures, merges of batches) gives every object the trajectory
    it has alone, bit for bit.
"""
import numpy as np
import pytest

from harness import XPBD, max_abs
import scenes
import serving
from batch import XPBDBatch

pytestmark = pytest.mark.gpu

TOL_REL = 1e-5


def small_mix(sizes, seed=5):
    return [scenes.object_request(n, seed * 100 + j) for j, n in enumerate(sizes)]


def singleton_run(lib_path, scene, recipe, script, frames):
    sim = XPBD(lib_path)
    sim.create(scene)
    scenes.configure_full(sim, recipe, iters=serving.ITERS, substeps=serving.SUBSTEPS)
    out = []
    for f in range(frames):
        t, p, d = script.colliders(f)
        sim.set_colliders(t, p, d)
        att = script.attached(f)
        for s in range(serving.SUBSTEPS):
            if s == (1 if f == 0 else 0):
                if att is not None:
                    sim.set_attached(att[0], att[1], 1.0)
                elif f == script.ramp:
                    sim.clear_attached()
            sim.step()
        sim.shapes_device(1e-2)
        out.append(sim.positions())
    sim.destroy()
    return out


def batch_run(objs, frames, tuning=None, autotune=False):
    reqs = [serving.Request(j, *o, life=frames) for j, o in enumerate(objs)]
    b = serving.build_batch(reqs, tuning=tuning)
    out = [[] for _ in objs]
    report = None
    for f in range(frames):
        if autotune and f == 1:
            report = b.autotune(substeps=2)
        serving.step_frame(b, [(r.script, f) for r in reqs])
        b.sync()
        for k in range(len(objs)):
            out[k].append(b.positions(k))
    b.close()
    return out, report


def test_batch_equals_singleton(product_path):
    objs = small_mix([3000, 4600, 6100])
    frames = 4
    got, _ = batch_run(objs, frames)
    for k, (scene, recipe, script) in enumerate(objs):
        ref = singleton_run(product_path, scene, recipe, script, frames)
        for f in range(frames):
            assert np.array_equal(got[k][f], ref[f]), f"object {k} frame {f}: max |dx| {max_abs(got[k][f], ref[f])}"
    print(f"batch of {len(objs)} objects == singleton runs, {frames} frames, bit for bit")


def oracle_frame(sim, script, f):
    t, p, d = script.colliders(f)
    sim.set_colliders(t, p, d)
    att = script.attached(f)
    for s in range(serving.SUBSTEPS):
        if s == (1 if f == 0 else 0):
            if att is not None:
                sim.set_attached(att[0], att[1], 1.0)
            elif f == script.ramp:
                sim.clear_attached()
        sim.step()


def test_batch_object_vs_oracle(oracle_path):
    objs = small_mix([1500, 2300], seed=9)
    frames = 3
    reqs = [serving.Request(j, *o, life=frames) for j, o in enumerate(objs)]
    b = serving.build_batch(reqs)
    start = [[None] * frames for _ in objs]     # state at the start of each frame
    after = [[None] * frames for _ in objs]
    for f in range(frames):
        for k in range(len(objs)):
            start[k][f] = b.get_state(k)
        serving.step_frame(b, [(r.script, f) for r in reqs])
        b.sync()
        for k in range(len(objs)):
            after[k][f] = b.positions(k)
    b.close()
    # the oracle is one global instance: one object at a time, each frame
    # from the batch's state at the frame's start
    worst = 0.0
    for k, (scene, recipe, script) in enumerate(objs):
        ext = scenes.scene_extent(scene)
        cpu = XPBD(oracle_path)
        cpu.create(scene)
        scenes.configure_full(cpu, recipe, iters=serving.ITERS, substeps=serving.SUBSTEPS)
        for f in range(frames):
            if f > 0:
                cpu.ext_set_state(*start[k][f])
            oracle_frame(cpu, script, f)
            err = max_abs(after[k][f], cpu.positions()) / ext
            worst = max(worst, err)
            assert err < TOL_REL, f"object {k} frame {f}: {err:.3e} of extent"
        cpu.destroy()
    print(f"batch objects vs oracle: worst per-frame |dx|/extent {worst:.3e}")


def test_copy_object_continues(product_path):
    objs = small_mix([2500, 3300, 4100], seed=13)
    frames = 6
    ref, _ = batch_run(objs, frames)
    reqs = [serving.Request(j, *o, life=frames) for j, o in enumerate(objs)]
    a = serving.build_batch(reqs)
    for f in range(3):
        serving.step_frame(a, [(r.script, f) for r in reqs])
        a.sync()
    # objects 2 and 0 continue in a new batch, in that order
    moved = [reqs[2], reqs[0]]
    b = serving.build_batch(moved)
    b.copy_object(0, a, 2)
    b.copy_object(1, a, 0)
    for f in range(3, frames):
        serving.step_frame(b, [(r.script, f) for r in moved])
        b.sync()
        for slot, k in ((0, 2), (1, 0)):
            got = b.positions(slot)
            assert np.array_equal(got, ref[k][f]), f"object {k} frame {f}: max |dx| {max_abs(got, ref[k][f])}"
    a.close()
    b.close()


def test_autotune_keeps_results(product_path):
    objs = small_mix([3500, 5200], seed=17)
    frames = 3
    ref, _ = batch_run(objs, frames)
    got, report = batch_run(objs, frames, autotune=True)
    assert report is not None and "start_ms" in report
    for k in range(len(objs)):
        for f in range(frames):
            assert np.array_equal(got[k][f], ref[k][f]), f"object {k} frame {f}"
    # every tuned shape also gives the same results (the one-wave-per-node
    # fine scan and the lane scan, other block shapes)
    for alt_tuning in ([1, 1, 1, 1, 16, 0, 0], [1, 1, 1, 1, 16, 0, 1]):
        alt, _ = batch_run(objs, frames, tuning=alt_tuning)
        for k in range(len(objs)):
            assert np.array_equal(alt[k][-1], ref[k][-1]), f"tuning {alt_tuning} object {k}"
    print("autotune report", report)


def test_server_alone_vs_mix(product_path):
    sizes = [2100, 3300, 2600, 4000, 2900, 3700]
    lives = [9, 5, 12, 6, 8, 7]
    arrive = [0, 0, 1, 3, 4, 6]
    objs = small_mix(sizes, seed=21)
    reqs = [serving.Request(j, *objs[j], life=lives[j], arrival_tick=arrive[j]) for j in range(len(objs))]
    srv = serving.Server(record=True, max_segments=2, waste_limit=0.3)
    srv.run(reqs, max_ticks=200)
    ticks = srv.tick_no
    merges = srv.merges
    srv.close()
    assert all(r.done is not None for r in reqs)
    assert merges >= 1
    for r in reqs:
        alone = serving.run_alone(r)
        mix = srv.trajectories[r.id]
        assert len(mix) == r.life
        for f in range(r.life):
            assert np.array_equal(mix[f], alone[f]), f"request {r.id} frame {f}: max |dx| {max_abs(mix[f], alone[f])}"
    print(f"served {len(reqs)} requests over {ticks} ticks with {merges} batch merges: "
          f"every trajectory equals its run alone, bit for bit")
