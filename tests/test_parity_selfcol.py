"""GPU parity for self-collision and object shape matching against the oracle.

  * build artefacts from an identical state are bit-exact: cell keys, the
    stable key order (the device radix sort against std::stable_sort, with an
    even and an odd number of 8-bit passes), candidate counts and candidate
    lists, statistics;
  * two spheres falling onto each other and the ground with object shape
    matching and self-collision: positions within 1e-5 of the scene extent per
    frame from an identical state.  The object sums are fp64 atomics on the
    GPU (an order the oracle does not share), so this case is toleranced, not
    bit-exact; the 100-frame drift is printed.
"""
import numpy as np
import pytest

from harness import XPBD, max_abs
import scenes

pytestmark = pytest.mark.gpu

TOL_REL = 1e-5


def still(sim, scene, within, cell, table=0, obj=0.0):
    sim.create(scene)
    sim.set_solver(iters=2, dt=1 / 600, vel_damping=0.0)
    sim.set_constraints(distance=False, shape=False, angle=False, volume=False)
    sim.set_ground(False, (0, 0, 1), 0.0, gravity=0.0)
    sim.set_colliders([], [], [])
    sim.set_object_shape(obj)
    sim.ext_set_self_collision_grid(cell, table)
    sim.set_self_collision(True, 1.5, 2.0, within)


@pytest.mark.parametrize("n_each,within,cell,table", [
    (1200, False, 0.0, 0),        # 14-bit keys: two sort passes
    (1200, True, 0.05, 0),
    (1200, False, 0.0, 1 << 17),  # 18-bit keys: three passes, result in the alternate buffers
    (60000, False, 0.0, 0),       # many sort tiles
])
def test_build_artefacts_bit_exact(product_path, oracle_path, n_each, within, cell, table):
    scene = scenes.stacked_bodies(n_each, 8, semi=(0.08, 0.08, 0.08), gap=0.004, seed=11)
    gpu, cpu = XPBD(product_path), XPBD(oracle_path)
    for sim in (gpu, cpu):
        still(sim, scene, within, cell, table, obj=0.3)
        sim.step()
    pos, _ = cpu.ext_get_state()
    vel = np.zeros_like(pos)
    vel[scene["pos"][:, 2] > 0.16, 2] = -0.5
    for sim in (gpu, cpu):
        sim.ext_set_state(pos, vel)
    for _ in range(3):
        pos, vel = cpu.ext_get_state()
        gpu.ext_set_state(pos, vel)
        gpu.step()
        cpu.step()
        kg, og, cg, lg = gpu.ext_self_collision_lists()
        kc, oc, cc, lc = cpu.ext_self_collision_lists()
        np.testing.assert_array_equal(kg, kc)
        np.testing.assert_array_equal(og, oc)
        np.testing.assert_array_equal(cg, cc)
        np.testing.assert_array_equal(lg, lc)
        assert gpu.self_collision_stats() == cpu.self_collision_stats()
    print(f"selfcol n={2 * n_each} within={within} cell={cell} table={table}: stats {cpu.self_collision_stats()}")
    assert cpu.self_collision_stats()[0] > 0
    gpu.destroy()
    cpu.destroy()


def falling(sim, scene):
    sim.create(scene)
    sim.set_solver(iters=6, dt=1 / 600, under_relax=0.6, vel_damping=0.01, dist_compliance=1e-6)
    sim.set_constraints(distance=True, shape=False, angle=False, volume=False)
    sim.set_ground(True, (0, 0, 1), 0.0, friction=0.5, restitution=0.1, gravity=9.81)
    sim.set_colliders([], [], [])
    sim.set_object_shape(0.3)
    sim.ext_set_self_collision_grid(0.0, 0)
    sim.set_self_collision(True, 1.5, 2.0, False)


def test_falling_bodies_parity(product_path, oracle_path):
    scene = scenes.stacked_bodies(2500, 8, semi=(0.08, 0.06, 0.05), gap=0.03, count=3, seed=41)
    gpu, cpu = XPBD(product_path), XPBD(oracle_path)
    for sim in (gpu, cpu):
        falling(sim, scene)
    extent = scenes.scene_extent(scene)
    worst, contacts = 0.0, 0
    for f in range(40):
        if f > 0:
            pos, vel = cpu.ext_get_state()
            gpu.ext_set_state(pos, vel)
        for _ in range(10):
            gpu.step()
            cpu.step()
        worst = max(worst, max_abs(gpu.positions(), cpu.positions()) / extent)
        contacts += cpu.self_collision_stats()[0]
    print(f"falling bodies: per-frame max |dx|/extent {worst:.3e}; candidates seen {contacts}")
    assert contacts > 0
    assert worst <= TOL_REL
    gpu.destroy()
    cpu.destroy()


def test_falling_bodies_drift_report(product_path, oracle_path):
    scene = scenes.stacked_bodies(2500, 8, semi=(0.08, 0.06, 0.05), gap=0.03, count=3, seed=41)
    gpu, cpu = XPBD(product_path), XPBD(oracle_path)
    for sim in (gpu, cpu):
        falling(sim, scene)
    extent = scenes.scene_extent(scene)
    drift = []
    for f in range(100):
        for _ in range(10):
            gpu.step()
            cpu.step()
        drift.append(max_abs(gpu.positions(), cpu.positions()) / extent)
    print(f"falling bodies 100-frame drift: max {max(drift):.3e} (frame {int(np.argmax(drift))}), "
          f"final {drift[-1]:.3e}")
    assert np.all(np.isfinite(gpu.positions()))
    gpu.destroy()
    cpu.destroy()


def squeezed(sim, scene, cell, table, fine=True):
    """One object with self-contact inside it, squeezed to 25 % along z (many
    candidates, lists that overflow their 32 slots) and moving."""
    sim.create(scene)
    sim.set_solver(iters=2, dt=1 / 600, vel_damping=0.0)
    sim.set_constraints(distance=False, shape=False, angle=False, volume=False)
    sim.set_ground(False, (0, 0, 1), 0.0, gravity=0.0)
    sim.set_colliders([], [], [])
    sim.set_object_shape(0.0)
    sim.ext_set_self_collision_grid(cell, table)
    if sim.backend == "hip-gfx950":
        sim.ext_set_self_collision_fine(fine)
    sim.set_self_collision(True, 1.5, 2.0, True)


@pytest.mark.parametrize("cell,table", [
    (0.0, 0),          # the reference rule: cell = search + 2 max move
    (0.0, 1024),       # a small table: stencil buckets collide
    (0.012, 4096),     # a fixed cell shorter than the reach, colliding buckets
])
def test_fine_scan_lists_bit_exact(product_path, oracle_path, cell, table):
    """The predicted-position grid path gives the reference's lists: against
    the oracle and against the GPU's own coarse walk, from identical states."""
    scene = scenes.make_ellipsoid_scene(12000, 8, semi=(0.08, 0.06, 0.05), seed=53)
    gpu, cpu = XPBD(product_path), XPBD(oracle_path)
    for sim in (gpu, cpu):
        squeezed(sim, scene, cell, table)
        sim.step()
    rng = np.random.Generator(np.random.PCG64(3))
    pos = scene["pos"].copy()
    pos[:, 2] *= 0.25
    vel = rng.normal(0.0, 0.4, size=pos.shape).astype(np.float32)
    over = 0
    for it in range(3):
        for sim in (gpu, cpu):
            sim.ext_set_state(pos, vel)
        gpu.ext_set_self_collision_fine(True)
        gpu.step()
        cpu.step()
        kf, of, cf, lf = gpu.ext_self_collision_lists()
        kc, oc, cc, lc = cpu.ext_self_collision_lists()
        np.testing.assert_array_equal(kf, kc)
        np.testing.assert_array_equal(of, oc)
        np.testing.assert_array_equal(cf, cc)
        np.testing.assert_array_equal(lf, lc)
        assert gpu.self_collision_stats() == cpu.self_collision_stats()
        over += cpu.self_collision_stats()[2]
        # the same state through the one-wave-per-node fine scan for every
        # node (the default is the lane scan, which hands crowded nodes and
        # lists past 64 to the wave scan)
        gpu.ext_set_state(pos, vel)
        gpu.ext_set_group_lanes(0)
        gpu.step()
        _, _, cgr, lgr = gpu.ext_self_collision_lists()
        np.testing.assert_array_equal(cgr, cf, err_msg="wave fine scan")
        np.testing.assert_array_equal(lgr, lf, err_msg="wave fine scan")
        assert gpu.self_collision_stats() == cpu.self_collision_stats()
        gpu.ext_set_group_lanes(-1)
        # the same state through the coarse walk
        gpu.ext_set_state(pos, vel)
        gpu.ext_set_self_collision_fine(False)
        gpu.step()
        _, _, cg, lg = gpu.ext_self_collision_lists()
        np.testing.assert_array_equal(cg, cf)
        np.testing.assert_array_equal(lg, lf)
        pos, vel = cpu.ext_get_state()
    print(f"fine scan cell={cell} table={table}: lists equal the oracle's and the coarse walk's; "
          f"stats {cpu.self_collision_stats()}, overflow seen {over}")
    assert over > 0
    gpu.destroy()
    cpu.destroy()
