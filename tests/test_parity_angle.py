"""GPU parity of the umbrella angle constraint (forward.cu:4990-5130,
10708-10819) against the oracle:

  * the pairing (host code on both sides) is identical entry for entry;
  * distance + angle + ground + a moving sphere: positions from an identical
    state agree per frame within 1e-5 of the extent (the GPU gathers each
    node's corrections in the oracle's serial order, so they are expected to
    be bit-identical; the reference's own float atomics are not).
"""
import numpy as np
import pytest

from harness import XPBD, max_abs
import scenes

pytestmark = pytest.mark.gpu

TOL_REL = 1e-5


def setup(sim, scene):
    sim.create(scene)
    sim.set_solver(iters=4, dt=1 / 600, under_relax=0.6, vel_damping=0.01, dist_compliance=1e-6,
                   angle_compliance=1e-4, angle_blend=0.2)
    sim.set_constraints(distance=True, shape=False, angle=True, volume=False)
    sim.set_ground(True, (0, 0, 1), 0.0, friction=0.5, restitution=0.0, gravity=9.81)


def test_angle_pairing_identical(product_path, oracle_path):
    scene = scenes.make_ellipsoid_scene(5000, 8, semi=(0.08, 0.06, 0.05), seed=31)
    gpu, cpu = XPBD(product_path), XPBD(oracle_path)
    gpu.create(scene)
    cpu.create(scene)
    pg, cg = gpu.ext_angle_pairs()
    pc, cc = cpu.ext_angle_pairs()
    np.testing.assert_array_equal(pg, pc)
    np.testing.assert_array_equal(cg.view(np.int32), cc.view(np.int32))
    print(f"angle pairing: {np.count_nonzero(pg >= 0)} of {len(pg)} entries paired, identical")
    gpu.destroy()
    cpu.destroy()


def test_angle_frames(product_path, oracle_path):
    scene = scenes.make_ellipsoid_scene(6000, 8, semi=(0.08, 0.06, 0.05), seed=32)
    gpu, cpu = XPBD(product_path), XPBD(oracle_path)
    for sim in (gpu, cpu):
        setup(sim, scene)
    extent = scenes.scene_extent(scene)
    worst, exact = 0.0, 0
    frames = 12
    for f in range(frames):
        pose = scenes.sphere_pose(f, frames=frames, start=(-0.12, 0.0, 0.05), travel=0.1)
        for sim in (gpu, cpu):
            sim.set_colliders([0], [pose], [[0.03, 0.0, 0.0]])
        if f > 0:
            pos, vel = cpu.ext_get_state()
            gpu.ext_set_state(pos, vel)
        for _ in range(10):
            gpu.step()
            cpu.step()
        a, b = gpu.positions(), cpu.positions()
        worst = max(worst, max_abs(a, b) / extent)
        exact += int(np.array_equal(a, b))
    print(f"angle frames: per-frame max |dx|/extent {worst:.3e}; bit-identical frames {exact}/{frames}")
    assert worst <= TOL_REL
    gpu.destroy()
    cpu.destroy()
