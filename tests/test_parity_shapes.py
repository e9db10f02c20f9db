"""GPU parity for shape matching, deformed shapes (compute_shapes), contact
ellipsoids and a collider-heavy gripper scene, against the oracle.

Tolerances: positions within 1e-5 of the scene extent per frame from an
identical state; deformed scales within 1e-5 relative and orientations within
1e-4 rad (compared through the covariances R diag(s^2) R^T, since the eigen
solver may order the axes differently); deformed counts equal.
"""
import numpy as np
import pytest

from harness import XPBD, max_abs
import scenes

pytestmark = pytest.mark.gpu

TOL_REL = 1e-5


@pytest.fixture(scope="module")
def scene():
    return scenes.make_ellipsoid_scene(3000, 10, seed=31)


def base(sim, scene, shape=True, robust=True, volume=False):
    sim.create(scene)
    sim.set_solver(iters=4, dt=1 / 600, under_relax=0.6, vel_damping=0.01, dist_compliance=1e-6,
                   shape_compliance=1e-7, shape_blend=0.3)
    sim.set_constraints(distance=True, shape=shape, angle=False, volume=volume)
    sim.set_shape_robust(robust)
    sim.set_ground(True, (0, 0, 1), 0.0, friction=0.5, restitution=0.2, gravity=9.81)


def run_pair(scene, product_path, oracle_path, frames, configure, per_frame=None):
    gpu, cpu = XPBD(product_path), XPBD(oracle_path)
    for sim in (gpu, cpu):
        configure(sim)
    extent = scenes.scene_extent(scene)
    worst = 0.0
    for f in range(frames):
        if f > 0:
            pos, vel = cpu.ext_get_state()
            gpu.ext_set_state(pos, vel)
        for s in range(10):
            if per_frame:
                per_frame(gpu, f, s)
                per_frame(cpu, f, s)
            gpu.step()
            cpu.step()
        worst = max(worst, max_abs(gpu.positions(), cpu.positions()) / extent)
    return gpu, cpu, worst


@pytest.mark.parametrize("robust", [True, False])
def test_shape_matching_parity(scene, product_path, oracle_path, robust):
    gpu, cpu, worst = run_pair(scene, product_path, oracle_path, 20, lambda s: base(s, scene, robust=robust))
    print(f"shape matching robust={robust}: per-frame max |dx|/extent {worst:.3e}")
    assert worst <= TOL_REL
    gpu.destroy()
    cpu.destroy()


def cov(scales, quats):
    w, x, y, z = quats.T.astype(np.float64)
    R = np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], 1)
    return np.einsum("nij,nj,nkj->nik", R, scales.astype(np.float64) ** 2, R)


def test_compute_shapes_parity(scene, product_path, oracle_path):
    gpu, cpu, worst = run_pair(scene, product_path, oracle_path, 15, lambda s: base(s, scene))
    pos, vel = cpu.ext_get_state()
    gpu.ext_set_state(pos, vel)
    sg, qg, kg = gpu.shapes(1e-3)
    sc, qc, kc = cpu.shapes(1e-3)
    assert kg == kc and kc > 0
    np.testing.assert_allclose(sg, sc, rtol=1e-5, atol=1e-9)
    Cg, Cc = cov(sg, qg), cov(sc, qc)
    err = np.linalg.norm(Cg - Cc, axis=(1, 2)) / np.linalg.norm(Cc, axis=(1, 2))
    # 1e-4 rad of rotation moves a covariance by at most ~2e-4 of its norm
    assert err.max() < 2e-4
    same = np.abs(np.sum(qg * qc, 1))
    print(f"compute_shapes: deformed {kg}; max cov rel err {err.max():.2e}; min |q.q'| {same.min():.8f}")
    gpu.destroy()
    cpu.destroy()


def test_grippers_contact_shapes_parity(scene, product_path, oracle_path):
    script = scenes.GripperScript(scene, seed=5, attach=80, ramp_frames=12)

    def configure(sim):
        base(sim, scene, shape=False)
        sim.set_contact_shape(3, tau=0.2, radius_cap=0.01)

    def per_frame(sim, f, s):
        if s == 0:
            t, p, d = script.colliders(f)
            # add a box and a sphere so every collider type is exercised, with friction
            box = np.concatenate([np.eye(3).ravel(), [0.18, 0.0, 0.05]]).astype(np.float32)
            sph = np.concatenate([np.eye(3).ravel(), [-0.2, 0.05, 0.06 - 0.002 * f]]).astype(np.float32)
            sim.set_colliders(np.r_[t, 1, 0], np.vstack([p, box, sph]), np.vstack([d, [0.04, 0.05, 0.06],
                              [0.05, 0, 0]]), margin=0.002, friction=0.3)
        if s == 1:
            att = script.attached(f)
            if att is not None:
                sim.set_attached(att[0], att[1], 2.0)
            elif f == script.ramp:
                sim.clear_attached()

    gpu, cpu, worst = run_pair(scene, product_path, oracle_path, 16, configure, per_frame)
    hg, pg, cg = gpu.collider_stats(4)
    hc, pc, cc = cpu.collider_stats(4)
    print(f"grippers+contact: per-frame max |dx|/extent {worst:.3e}; hits gpu {hg} oracle {hc}")
    assert worst <= TOL_REL
    np.testing.assert_array_equal(hg, hc)
    np.testing.assert_allclose(pg, pc, rtol=1e-3, atol=1e-6)
    gpu.destroy()
    cpu.destroy()


@pytest.mark.parametrize("robust", [True, False])
def test_cluster_shape_matching_parity(scene, product_path, oracle_path, robust):
    lab = scenes.kmeans_clusters(scene["pos"], size=64, seed=5)

    def configure(sim):
        base(sim, scene, robust=robust)
        sim.ext_set_shape_clusters(lab)

    gpu, cpu, worst = run_pair(scene, product_path, oracle_path, 20, configure)
    print(f"cluster shape matching robust={robust} ({lab.max() + 1} clusters): per-frame max |dx|/extent {worst:.3e}")
    assert worst <= TOL_REL
    gpu.destroy()
    cpu.destroy()


@pytest.mark.parametrize("robust", [True, False])
@pytest.mark.parametrize("size", [64, 190])
def test_cluster_sums_bit_exact(scene, product_path, oracle_path, robust, size):
    """Cluster shape matching alone, one substep from an identical perturbed
    state: the wave's lane partials and butterfly tree are restated in the
    oracle, so positions agree bit for bit (190-member clusters take three
    lane rounds)."""
    lab = scenes.kmeans_clusters(scene["pos"], size=size, seed=5)
    rng = np.random.default_rng(1)
    pos = (scene["pos"] + rng.normal(0, 0.003, scene["pos"].shape)).astype(np.float32)
    out = []
    for path in (product_path, oracle_path):
        sim = XPBD(path)
        sim.create(scene)
        sim.set_solver(iters=1, dt=1 / 600, under_relax=0.6, vel_damping=0.01, dist_compliance=1e-6,
                       shape_compliance=1e-7, shape_blend=0.3)
        sim.set_constraints(distance=False, shape=True, angle=False, volume=False)
        sim.set_shape_robust(robust)
        sim.set_ground(False, (0, 0, 1), 0.0, gravity=0.0)
        sim.ext_set_shape_clusters(lab)
        sim.step()
        sim.ext_set_state(pos, np.zeros_like(pos))
        sim.step()
        out.append(sim.positions())
        sim.destroy()
    moved = np.abs(out[1] - pos).max()
    print(f"cluster sums size={size} robust={robust}: {lab.max() + 1} clusters, moved {moved:.3e}")
    assert moved > 1e-5
    np.testing.assert_array_equal(out[0], out[1])
