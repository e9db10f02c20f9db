"""The oracle's OpenMP loops (used only for the multi-core CPU baseline) give
the same bits as its serial order: every parallel loop writes only its own
node's or cluster's entries (CPU only)."""
import ctypes

import numpy as np

from harness import XPBD
import scenes


def run(path, threads, scene, recipe, script, frames=2):
    ctypes.CDLL(path).xpbd_oracle_set_threads(threads)
    sim = XPBD(path)
    sim.create(scene)
    scenes.configure_full(sim, recipe, iters=2, substeps=5)
    sim.step()   # attachments are refused before the first step
    out = []
    for f in range(frames):
        script.apply(sim, f)
        for s in range(5):
            sim.step()
        out.append(sim.ext_get_state())
    stats = sim.self_collision_stats()
    shapes = sim.shapes()
    sim.destroy()
    ctypes.CDLL(path).xpbd_oracle_set_threads(1)
    return out, stats, shapes


def test_threads_give_the_serial_bits(oracle_path):
    scene = scenes.make_ellipsoid_scene(3000, 12, semi=(0.06, 0.05, 0.04), seed=8)
    recipe = scenes.full_recipe(scene, cluster_size=64, seed=2)
    script = scenes.GripperScript(scene, seed=3, attach=60)
    a, sa, ha = run(oracle_path, 1, scene, recipe, script)
    b, sb, hb = run(oracle_path, 4, scene, recipe, script)
    assert sa == sb
    for (xa, va), (xb, vb) in zip(a, b):
        np.testing.assert_array_equal(xa, xb)
        np.testing.assert_array_equal(va, vb)
    for u, w in zip(ha, hb):
        np.testing.assert_array_equal(u, w)
    assert sa[0] > 0, "the scene should produce self-collision candidates"
