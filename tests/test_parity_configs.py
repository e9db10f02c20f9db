"""GPU tests at the workload sizes: the 200k and 1M full-recipe solvers, the
64-object batched mix and a served stream of requests from 2k to 500k
Gaussians.

  * 200k full recipe (kNN k=12 distance, node + 3NN volume clusters, k-means
    shape clusters of 64, self-collision within the body in 2^21 cells, the
    gripper script): per frame from an identical state against the oracle.
    Positions within 1e-5 of the scene extent after every frame; deformed
    shapes within 1e-4 rad through their covariances; the first substep's
    self-collision artefacts (hash keys, stable sorted order, candidate counts
    and lists) bit for bit.
  * 1M full recipe (2^23 cells): the same checks on the first frame and on
    one substep from a state 12 frames into the script, then 20 frames of
    properties (finite positions, no node below the ground by more than the
    self-collision radius, the volume ratio J near 1).
  * 64 objects, 5k-200k Gaussians: the batch equals each sampled object's
    run alone bit for bit, and the two smallest objects match the oracle per
    frame.
  * a served stream (Poisson arrivals at 20 per second, 2k-500k Gaussians):
    every sampled request's trajectory in the mix equals its run alone.
"""
import ctypes
import os
import time

import numpy as np
import pytest

from conftest import ORACLE_LIB
from harness import XPBD, max_abs, ulp_diff
import scenes
import serving

pytestmark = pytest.mark.gpu

TOL_REL = 1e-5
SUBSTEPS = 10


def oracle_threads():
    n = min(len(os.sched_getaffinity(0)), 16)
    ctypes.CDLL(ORACLE_LIB).xpbd_oracle_set_threads(n)
    return n


def cov(scales, quats):
    w, x, y, z = quats.T.astype(np.float64)
    R = np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], 1)
    return np.einsum("nij,nj,nkj->nik", R, scales.astype(np.float64) ** 2, R)


def frame_script(sim, script, f, first_substep_hook=None):
    t, p, d = script.colliders(f)
    sim.set_colliders(t, p, d)
    att = script.attached(f)
    for s in range(SUBSTEPS):
        if s == (1 if f == 0 else 0):
            if att is not None:
                sim.set_attached(att[0], att[1], 1.0)
            elif f == script.ramp:
                sim.clear_attached()
        sim.step()
        if s == 0 and first_substep_hook is not None:
            first_substep_hook()


def compare_lists(gpu, cpu, tag):
    kg, og, cg, lg = gpu.ext_self_collision_lists()
    kc, oc, cc, lc = cpu.ext_self_collision_lists()
    np.testing.assert_array_equal(kg, kc, err_msg=f"{tag}: hash keys")
    np.testing.assert_array_equal(og, oc, err_msg=f"{tag}: sorted order")
    np.testing.assert_array_equal(cg, cc, err_msg=f"{tag}: candidate counts")
    mask = np.arange(32)[None, :] < cc[:, None]
    np.testing.assert_array_equal(np.where(mask, lg, 0), np.where(mask, lc, 0), err_msg=f"{tag}: candidate lists")
    return int(cc.sum())


def compare_shapes(gpu, cpu, tag):
    sg, qg, kg = gpu.shapes()
    sc, qc, kc = cpu.shapes()
    assert kg == kc, f"{tag}: deformed count {kg} vs {kc}"
    np.testing.assert_allclose(sg, sc, rtol=1e-5, err_msg=f"{tag}: deformed scales")
    Cg, Cc = cov(sg, qg), cov(sc, qc)
    err = np.linalg.norm(Cg - Cc, axis=(1, 2)) / np.linalg.norm(Cc, axis=(1, 2))
    assert err.max() < 2e-4, f"{tag}: covariance rel err {err.max():.2e}"
    return float(err.max())


def full_object(n, seed=3):
    scene = scenes.make_ellipsoid_scene(n, 12, seed=seed)
    recipe = scenes.full_recipe(scene, seed=seed)
    script = scenes.GripperScript(scene, seed=7)
    return scene, recipe, script


@pytest.mark.timeout(900)
def test_full_recipe_200k_per_frame(product_path, oracle_path):
    t0 = time.perf_counter()
    scene, recipe, script = full_object(200_000)
    assert recipe["table"] == 1 << 21
    threads = oracle_threads()
    extent = scenes.scene_extent(scene)
    gpu, cpu = XPBD(product_path), XPBD(oracle_path)
    for sim in (gpu, cpu):
        sim.create(scene)
        scenes.configure_full(sim, recipe)
    worst, worst_cov, cand = 0.0, 0.0, []
    for f in range(6):
        if f > 0:
            gpu.ext_set_state(*cpu.ext_get_state())
        frame_script(cpu, script, f)
        frame_script(gpu, script, f, first_substep_hook=None)
        err = max_abs(gpu.positions(), cpu.positions()) / extent
        worst = max(worst, err)
        assert err <= TOL_REL, f"frame {f}: |dx|/extent {err:.3e}"
        worst_cov = max(worst_cov, compare_shapes(gpu, cpu, f"frame {f}"))
    # self-collision artefacts of one substep from an identical state
    gpu.ext_set_state(*cpu.ext_get_state())
    for sim in (gpu, cpu):
        t, p, d = script.colliders(6)
        sim.set_colliders(t, p, d)
        att = script.attached(6)
        sim.set_attached(att[0], att[1], 1.0)
        sim.step()
    cand.append(compare_lists(gpu, cpu, "200k substep"))
    print(f"200k full recipe: 6 frames, worst |dx|/extent {worst:.3e}, covariance rel err {worst_cov:.2e}, "
          f"lists bit-exact ({cand[0]} candidates); oracle on {threads} threads, {time.perf_counter() - t0:.1f}s")
    gpu.destroy()
    cpu.destroy()


@pytest.mark.timeout(1200)
def test_full_recipe_1m(product_path, oracle_path):
    t0 = time.perf_counter()
    scene, recipe, script = full_object(1_000_000)
    assert recipe["table"] == 1 << 23
    threads = oracle_threads()
    extent = scenes.scene_extent(scene)
    gpu, cpu = XPBD(product_path), XPBD(oracle_path)
    for sim in (gpu, cpu):
        sim.create(scene)
        scenes.configure_full(sim, recipe)
    # frame 0 from rest on both: lists after the first substep, positions after the frame
    hooks = {}
    frame_script(cpu, script, 0, first_substep_hook=lambda: hooks.setdefault("c", cpu.ext_self_collision_lists()))
    frame_script(gpu, script, 0, first_substep_hook=lambda: hooks.setdefault("g", gpu.ext_self_collision_lists()))
    for a, b, what in zip(hooks["g"], hooks["c"], ("keys", "order", "counts")):
        np.testing.assert_array_equal(a, b, err_msg=f"1M frame 0 substep 0: {what}")
    err0 = max_abs(gpu.positions(), cpu.positions()) / extent
    assert err0 <= TOL_REL, f"1M frame 0: {err0:.3e}"
    compare_shapes(gpu, cpu, "1M frame 0")
    t1 = time.perf_counter()
    print(f"1M full recipe: frame 0 compared ({t1 - t0:.1f}s)", flush=True)
    # GPU alone to frame 12, then one substep from that state on both
    for f in range(1, 12):
        frame_script(gpu, script, f)
    pos, vel = gpu.ext_get_state()
    cpu.ext_set_state(pos, vel)
    t, p, d = script.colliders(11)       # the colliders' previous pose, as the GPU saw it
    cpu.set_colliders(t, p, d)
    for sim in (gpu, cpu):
        t, p, d = script.colliders(12)
        sim.set_colliders(t, p, d)
        att = script.attached(12)
        sim.set_attached(att[0], att[1], 1.0)
        sim.step()
    ncand = compare_lists(gpu, cpu, "1M frame 12 substep 0")
    err1 = max_abs(gpu.positions(), cpu.positions()) / extent
    assert err1 <= TOL_REL, f"1M frame 12 substep: {err1:.3e}"
    t2 = time.perf_counter()
    print(f"1M full recipe: frame-12 substep compared ({t2 - t1:.1f}s)", flush=True)
    cpu.destroy()
    # 20 more frames: properties
    jmin = []
    search = None
    for f in range(12, 32):
        frame_script(gpu, script, f)
        if f == 12:
            search = gpu.ext_self_collision_grid_info()[2]
        p = gpu.positions()
        assert np.isfinite(p).all(), f"frame {f}: non-finite positions"
        assert p[:, 2].min() >= -search, f"frame {f}: a node {p[:, 2].min():.4g} below the ground"
        j, jstd = gpu.volume_stats(True)
        jmin.append(j)
        assert 0.8 < j < 1.2, f"frame {f}: volume-weighted J {j}"
    print(f"1M full recipe: frame 0 |dx|/extent {err0:.3e} (lists bit-exact), frame-12 substep {err1:.3e} "
          f"(lists bit-exact, {ncand} candidates), 20 frames finite, J in [{min(jmin):.4f}, {max(jmin):.4f}]; "
          f"oracle on {threads} threads {t2 - t1 + (t1 - t0):.1f}s total")
    gpu.destroy()


def _obj_frame(sim, script, f):
    frame_script(sim, script, f)


@pytest.mark.timeout(1200)
def test_batched_mix_64_objects(product_path, oracle_path):
    from batch import XPBDBatch  # noqa: F401
    t0 = time.perf_counter()
    objs = scenes.object_mix(64, 5_000, 200_000, seed=11)
    reqs = [serving.Request(j, *o, life=10 ** 6) for j, o in enumerate(objs)]
    total = sum(r.n for r in reqs)
    t1 = time.perf_counter()
    frames = 3
    b = serving.build_batch(reqs)
    starts = {}
    out = [[] for _ in reqs]
    sizes = np.array([r.n for r in reqs])
    small = [int(k) for k in np.argsort(sizes)[:2]]
    rng = np.random.default_rng(4)
    sample = sorted({int(np.argmax(sizes)), small[0]} | {int(k) for k in rng.choice(64, 3, replace=False)})
    for f in range(frames):
        for k in small:
            starts[(k, f)] = b.get_state(k)
        serving.step_frame(b, [(r.script, f) for r in reqs])
        b.sync()
        for k in range(len(reqs)):
            out[k].append(b.positions(k))
    b.close()
    t2 = time.perf_counter()
    # sampled objects alone (one-object batches, same path as a singleton)
    for k in sample:
        alone = serving.run_alone(serving.Request(k, *objs[k], life=frames))
        for f in range(frames):
            assert np.array_equal(out[k][f], alone[f]), f"object {k} (n={reqs[k].n}) frame {f}"
    # the two smallest against the oracle, per frame from the batch's state
    oracle_threads()
    worst = 0.0
    for k in small:
        scene, recipe, script = objs[k]
        ext = scenes.scene_extent(scene)
        cpu = XPBD(oracle_path)
        cpu.create(scene)
        scenes.configure_full(cpu, recipe)
        for f in range(frames):
            if f > 0:
                cpu.ext_set_state(*starts[(k, f)])
            frame_script(cpu, script, f)
            err = max_abs(out[k][f], cpu.positions()) / ext
            worst = max(worst, err)
            assert err <= TOL_REL, f"object {k} frame {f}: {err:.3e}"
        cpu.destroy()
    print(f"64-object mix ({total} Gaussians, built in {t1 - t0:.1f}s, {frames} frames in {t2 - t1:.1f}s): "
          f"objects {sample} equal their runs alone bit for bit; smallest two vs oracle {worst:.3e}")


@pytest.mark.timeout(1200)
def test_served_stream_alone_vs_mix(product_path):
    t0 = time.perf_counter()
    pool = serving.TemplatePool(6, 2_000, 500_000, seed=8)
    sizes = sorted(int(it[0]["pos"].shape[0]) for it in pool.items)
    reqs = serving.poisson_stream(pool, 20.0, 45, life=(15, 40), seed=6, steady_start=False)
    t1 = time.perf_counter()
    srv = serving.Server(record=True, max_segments=3)
    srv.run(reqs, max_ticks=400)
    ticks, merges = srv.tick_no, srv.merges
    ms = np.array([r["ms"] for r in srv.log])
    srv.close()
    assert all(r.done is not None for r in reqs)
    rng = np.random.default_rng(2)
    big = max(reqs, key=lambda r: r.n)
    sample = {big.id} | {int(i) for i in rng.choice(len(reqs), 4, replace=False)}
    for r in reqs:
        if r.id not in sample:
            continue
        alone = serving.run_alone(r)
        mix = srv.trajectories[r.id]
        assert len(mix) == r.life
        for f in range(r.life):
            assert np.array_equal(mix[f], alone[f]), f"request {r.id} (n={r.n}) frame {f}"
    print(f"served stream: {len(reqs)} requests (templates {sizes}), {ticks} ticks, {merges} merges, "
          f"tick p50 {np.percentile(ms, 50):.1f} ms p99 {np.percentile(ms, 99):.1f} ms; sampled requests "
          f"{sorted(sample)} equal their runs alone bit for bit; {time.perf_counter() - t0:.1f}s "
          f"(templates {t1 - t0:.1f}s)")
