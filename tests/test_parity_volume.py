"""GPU parity of the volume constraint against the oracle.

Cluster topology, material types and rest determinants must match bit for bit
(integer artefacts, and the same float operations on both sides).  Positions
are compared per frame from an identical state within 1e-5 of the scene extent;
the wave64 cluster solve sums in a different order, so only that path is
expected to differ in the last bits.
"""
import numpy as np
import pytest

from harness import XPBD, max_abs
import scenes

pytestmark = pytest.mark.gpu

TOL_REL = 1e-5


@pytest.fixture(scope="module")
def scene():
    return scenes.make_ellipsoid_scene(4000, 10, seed=21)


def tetra(scene):
    from scipy.spatial import cKDTree
    p = scene["pos"].astype(np.float64)
    _, nb = cKDTree(p).query(p, k=4)
    return np.arange(0, 3 * (p.shape[0] + 1), 3, dtype=np.int32), nb[:, 1:].astype(np.int32).ravel(), 3


def setup(sim, scene, k, cap, r, explicit):
    sim.create(scene)
    sim.set_solver(iters=4, dt=1 / 600, under_relax=0.6, vel_damping=0.01, dist_compliance=1e-6)
    sim.set_constraints(distance=True, shape=False, angle=False, volume=True)
    sim.set_volume(1e-6, k, cap, r)
    sim.set_ground(True, (0, 0, 1), 0.0, friction=0.5, restitution=0.0, gravity=9.81)
    if explicit is not None:
        sim.ext_set_volume_clusters(*explicit)


@pytest.mark.parametrize("case", ["k1", "k3r2", "k2cap32", "tetra"])
def test_volume_parity(scene, product_path, oracle_path, case):
    k, cap, r, explicit = {"k1": (1, 64, 1, None), "k3r2": (3, 2048, 2, None), "k2cap32": (2, 32, 1, None),
                           "tetra": (1, 64, 1, tetra(scene))}[case]
    gpu, cpu = XPBD(product_path), XPBD(oracle_path)
    for sim in (gpu, cpu):
        setup(sim, scene, k, cap, r, explicit)
        sim.step()
    tg, tc = gpu.ext_volume_topology(), cpu.ext_volume_topology()
    for name, a, b in zip(("leader", "members", "rev", "matType"), tg[:4], tc[:4]):
        np.testing.assert_array_equal(a, b, err_msg=name)
    np.testing.assert_array_equal(tg[4].view(np.int32), tc[4].view(np.int32), err_msg="detRest bits")
    mg, mc = gpu.ext_volume_members(), cpu.ext_volume_members()
    for i, (a, b) in enumerate(zip(mg, mc)):
        np.testing.assert_array_equal(a, b, err_msg=f"members of cluster {i}")
    extent = scenes.scene_extent(scene)
    worst = 0.0
    for f in range(30):
        pos, vel = cpu.ext_get_state()
        gpu.ext_set_state(pos, vel)
        for _ in range(10):
            gpu.step()
            cpu.step()
        worst = max(worst, max_abs(gpu.positions(), cpu.positions()) / extent)
    jg, jc = gpu.volume_stats(True), cpu.volume_stats(True)
    print(f"volume {case}: per-frame max |dx|/extent {worst:.3e}; J_vw gpu {jg} oracle {jc}")
    assert worst <= TOL_REL
    assert jg[0] == pytest.approx(jc[0], abs=1e-4)
    gpu.destroy()
    cpu.destroy()


def _topology(sim):
    return sim.ext_volume_topology()[:3], sim.ext_volume_members()


@pytest.mark.parametrize("n,k,cap,r,gs", [
    (30000, 1, 64, 1, False),      # own rows, duplicates kept
    (30000, 3, 2048, 2, False),    # the reference rule: greedy cover over the device balls
    (30000, 2, 32, 1, False),      # farthest-point sampling on the device
    (30000, 3, 48, 3, True),       # sampling, a wider leader spacing, the Gauss-Seidel colouring
    (6000, 6, 96, 1, False),       # balls past the device limit: built on the host
])
def test_device_builders_equal_host(product_path, oracle_path, n, k, cap, r, gs):
    """k-ring member lists built on the GPU equal the host's (and the oracle's)
    list for list, leader flags and reverse-list sizes included; the solver
    then steps bit for bit the same."""
    scene = scenes.make_ellipsoid_scene(n, 12, seed=23)
    got = {}
    try:
        for mode in (0, 1):
            sim = XPBD(product_path)
            sim.ext_set_device_builders(mode)
            setup(sim, scene, k, cap, r, None)
            if gs:
                sim.ext_set_gauss_seidel(False, True)
            for _ in range(20):
                sim.step()
            # positions first: the topology read-back fetches the lists to the host
            got[mode] = (sim.positions(), _topology(sim))
            sim.destroy()
    finally:
        XPBD(product_path).ext_set_device_builders(-1)
    (ph, (th, mh)), (pd, (td, md)) = got[0], got[1]
    for name, a, b in zip(("leader", "members", "rev"), td, th):
        np.testing.assert_array_equal(a, b, err_msg=name)
    assert len(md) == len(mh)
    for i, (a, b) in enumerate(zip(md, mh)):
        np.testing.assert_array_equal(a, b, err_msg=f"members of cluster {i}")
    assert np.array_equal(pd, ph), f"positions after 20 steps: max |dx| {max_abs(pd, ph)}"
    sizes = np.array([len(m) for m in mh])
    if (k, cap, r) == (3, 2048, 2):
        cpu = XPBD(oracle_path)
        setup(cpu, scene, k, cap, r, None)
        cpu.step()
        tc, mc = _topology(cpu)
        cpu.destroy()
        for name, a, b in zip(("leader", "members", "rev"), td, tc):
            np.testing.assert_array_equal(a, b, err_msg=f"oracle {name}")
        for i, (a, b) in enumerate(zip(md, mc)):
            np.testing.assert_array_equal(a, b, err_msg=f"oracle members of cluster {i}")
    print(f"device builders n={n} k={k} cap={cap} r={r} gs={gs}: {int(td[0].sum())} clusters, members mean "
          f"{sizes[sizes > 0].mean():.1f} max {sizes.max()}; equal to the host's, positions bit-identical")
