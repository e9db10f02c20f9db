"""GPU parity of the graph-coloured Gauss-Seidel passes against the oracle.

  * colour assignment bit-exact: the distance edge list, rest/stiffness and
    colour offsets, and the volume leader order and colour offsets, equal the
    oracle's;
  * positions within 1e-5 of the scene extent per frame from an identical
    state (expected: bit-exact, colours run in the same order and no two
    threads of a colour share a node), for distance GS, volume GS (k-ring
    clusters and tetrahedra) and both together with ground and a collider.
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree

from harness import XPBD, max_abs
import scenes

pytestmark = pytest.mark.gpu

TOL_REL = 1e-5


@pytest.fixture(scope="module")
def scene():
    return scenes.make_ellipsoid_scene(6000, 10, seed=23)


def configure(sim, scene, dist_gs, vol_gs, volume, tetra):
    sim.create(scene)
    sim.set_solver(iters=4, dt=1 / 600, under_relax=0.6, vel_damping=0.01, dist_compliance=1e-6)
    sim.set_constraints(distance=True, shape=False, angle=False, volume=volume)
    sim.set_volume(1e-6, 2, 64, 2)
    sim.set_ground(True, (0, 0, 1), 0.0, friction=0.5, restitution=0.2, gravity=9.81)
    sim.set_colliders([0], np.array([[1, 0, 0, 0, 1, 0, 0, 0, 1, 0.0, 0.0, 0.12]], np.float32),
                      np.array([[0.05, 0, 0]], np.float32))
    sim.set_object_shape(0.0)
    sim.set_self_collision(False)
    sim.ext_set_gauss_seidel(dist_gs, vol_gs)
    if tetra:
        _, nn = cKDTree(scene["pos"]).query(scene["pos"], 4)
        off = np.arange(0, 3 * len(nn) + 1, 3, dtype=np.int32)
        sim.ext_set_volume_clusters(off, np.ascontiguousarray(nn[:, 1:].astype(np.int32).ravel()), 3)


@pytest.mark.parametrize("dist_gs,vol_gs,volume,tetra", [
    (True, False, False, False),
    (False, True, True, False),
    (True, True, True, True),
])
def test_gs_parity(scene, product_path, oracle_path, dist_gs, vol_gs, volume, tetra):
    gpu, cpu = XPBD(product_path), XPBD(oracle_path)
    for sim in (gpu, cpu):
        configure(sim, scene, dist_gs, vol_gs, volume, tetra)
    extent = scenes.scene_extent(scene)
    worst = 0.0
    for f in range(20):
        if f > 0:
            pos, vel = cpu.ext_get_state()
            gpu.ext_set_state(pos, vel)
        for _ in range(5):
            gpu.step()
            cpu.step()
        worst = max(worst, max_abs(gpu.positions(), cpu.positions()) / extent)
    ig, ic = gpu.ext_gs_info(), cpu.ext_gs_info()
    assert ig == ic
    if dist_gs:
        for a, b in zip(gpu.ext_distance_gs_edges(), cpu.ext_distance_gs_edges()):
            np.testing.assert_array_equal(a, b)
    if vol_gs:
        for a, b in zip(gpu.ext_volume_gs_order(), cpu.ext_volume_gs_order()):
            np.testing.assert_array_equal(a, b)
    print(f"GS dist={dist_gs} vol={vol_gs} tetra={tetra}: {ig}; per-frame max |dx|/extent {worst:.3e}")
    assert worst <= TOL_REL
    gpu.destroy()
    cpu.destroy()
