"""The device node order (Morton by default) is invisible at the C ABI.

The same scripted runs with the caller's order (mode 0) and the Morton order
(mode 1) must give bit-identical results through every per-node output:
positions (host copy and the device pointer of xpbd_positions_device),
velocities, inverse masses after attaching, deformed shapes, volume topology,
self-collision artefacts.
"""
import ctypes

import numpy as np
import pytest

from harness import XPBD
import scenes

pytestmark = pytest.mark.gpu


def device_positions(sim):
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemcpy.restype = ctypes.c_int
    out = np.empty((sim.n, 3), np.float32)
    ptr = sim.positions_device_ptr()
    assert ptr != 0
    assert hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0   # device to host
    return out


def run(product_path, scene, mode, volume):
    sim = XPBD(product_path)
    sim.ext_set_node_order(mode)
    sim.create(scene)
    sim.set_solver(iters=4, dt=1 / 600, under_relax=0.6, vel_damping=0.01, dist_compliance=1e-6,
                   shape_compliance=1e-7, shape_blend=0.3)
    sim.set_constraints(distance=True, shape=True, angle=False, volume=volume)
    sim.set_volume(1e-6, 2, 64, 2)
    sim.set_ground(True, (0, 0, 1), 0.0, friction=0.5, restitution=0.2, gravity=9.81)
    sim.set_colliders([], [], [])
    # object shape matching sums in fp64 atomics (order-dependent): off here
    sim.set_object_shape(0.0)
    sim.ext_set_self_collision_grid(0.0, 0)
    sim.set_self_collision(True, 1.5, 2.0, False)
    out = {}
    for f in range(12):
        if f == 4:
            idx = np.arange(0, scene["pos"].shape[0], 97, dtype=np.int32)
            sim.set_attached(idx, sim.positions()[idx] + np.float32(0.01), 2.0)
        for _ in range(5):
            sim.step()
    out["pos"] = sim.positions()
    out["dev"] = device_positions(sim)
    out["state"] = sim.ext_get_state()
    out["inv_mass"] = sim.ext_inv_mass()
    out["shapes"] = sim.shapes(1e-3)
    out["lists"] = sim.ext_self_collision_lists()
    out["stats"] = sim.self_collision_stats()
    if volume:
        out["topo"] = sim.ext_volume_topology()
        out["J"] = sim.volume_stats(True)
    sim.destroy()
    return out


@pytest.mark.parametrize("volume", [False, True])
def test_node_order_is_invisible(product_path, volume):
    scene = scenes.stacked_bodies(1500, 8, semi=(0.07, 0.06, 0.05), gap=0.01, count=2, seed=51)
    a = run(product_path, scene, 0, volume)
    b = run(product_path, scene, 1, volume)
    np.testing.assert_array_equal(a["pos"], b["pos"])
    np.testing.assert_array_equal(a["dev"], a["pos"])
    np.testing.assert_array_equal(b["dev"], b["pos"])
    for x, y in zip(a["state"], b["state"]):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(a["inv_mass"], b["inv_mass"])
    np.testing.assert_array_equal(a["shapes"][0], b["shapes"][0])
    np.testing.assert_array_equal(a["shapes"][1], b["shapes"][1])
    assert a["shapes"][2] == b["shapes"][2]
    for x, y in zip(a["lists"], b["lists"]):
        np.testing.assert_array_equal(x, y)
    assert a["stats"] == b["stats"]
    if volume:
        for x, y in zip(a["topo"], b["topo"]):
            np.testing.assert_array_equal(x, y)
        assert a["J"] == b["J"]
    print(f"node order invisible (volume={volume}): stats {a['stats']}")
