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


def _with_duplicates(scene, seed=5):
    """The scene with 5% of its edges listed again (other rest length and
    stiffness, some reversed): duplicate pairs are summed in CSR order."""
    rng = np.random.Generator(np.random.PCG64(seed))
    e = scene["edges"]
    pick = rng.choice(len(e), size=len(e) // 20, replace=False)
    extra = e[pick].copy()
    flip = rng.random(len(pick)) < 0.5
    extra[flip] = extra[flip][:, ::-1]
    out = dict(scene)
    out["edges"] = np.ascontiguousarray(np.concatenate([e, extra]).astype(np.int32))
    out["rest"] = np.ascontiguousarray(np.concatenate([scene["rest"], scene["rest"][pick] * 1.01]).astype(np.float32))
    out["stiff"] = np.ascontiguousarray(np.concatenate([scene["stiff"], np.full(len(pick), 0.5, np.float32)]))
    return out


@pytest.mark.parametrize("tetra", [False, True])
def test_gs_device_colourings_equal(product_path, oracle_path, tetra):
    """Both Gauss-Seidel colourings built on the GPU equal the host's (and the
    oracle's) edge for edge and leader for leader, duplicate edges included,
    and the solver then steps bit for bit the same."""
    scene = _with_duplicates(scenes.make_ellipsoid_scene(40000, 10, seed=29))
    got = {}
    try:
        for mode in (0, 1):
            sim = XPBD(product_path)
            sim.ext_set_device_builders(mode)
            configure(sim, scene, True, True, True, tetra)
            for _ in range(10):
                sim.step()
            got[mode] = (sim.positions(), sim.ext_distance_gs_edges(), sim.ext_volume_gs_order(), sim.ext_gs_info())
            sim.destroy()
    finally:
        XPBD(product_path).ext_set_device_builders(-1)
    (ph, dh, vh, ih), (pd, dd, vd, idv) = got[0], got[1]
    assert ih == idv
    for name, a, b in zip(("ab", "rest", "stiff", "colorOff"), dd, dh):
        np.testing.assert_array_equal(a, b, err_msg=f"distance {name}")
    for name, a, b in zip(("order", "colorOff"), vd, vh):
        np.testing.assert_array_equal(a, b, err_msg=f"volume {name}")
    assert np.array_equal(pd, ph), f"positions after 10 steps: max |dx| {max_abs(pd, ph)}"
    cpu = XPBD(oracle_path)
    configure(cpu, scene, True, True, True, tetra)
    cpu.step()
    dc, vc = cpu.ext_distance_gs_edges(), cpu.ext_volume_gs_order()
    cpu.destroy()
    for name, a, b in zip(("ab", "rest", "stiff", "colorOff"), dd, dc):
        np.testing.assert_array_equal(a, b, err_msg=f"oracle distance {name}")
    for name, a, b in zip(("order", "colorOff"), vd, vc):
        np.testing.assert_array_equal(a, b, err_msg=f"oracle volume {name}")
    print(f"device GS colourings (tetra={tetra}): {idv}; equal to the host's and the oracle's, positions bit-identical")
