"""Pins the oracle's shape-matching, deformed-shape and contact-ellipsoid
restatements (CPU only) against closed forms and numpy.

  * fp64 polar rotation == U V^T of numpy's SVD (det-fixed) for well-conditioned
    matrices; the flat fix keeps det(R) = +1 for rank-2 inputs;
  * float Jacobi eigenvalues == numpy.linalg.eigvalsh, eigenvectors orthonormal;
  * shape matching leaves a rigidly rotated object in place;
  * deformed shapes: rest returns the input shape; a rigid rotation R of the
    whole object maps every covariance to R S R^T; a uniform scale s multiplies
    every scale by s (within the regulariser's 1e-4 bias);
  * contact ellipsoid M == R diag((k s)^2) R^T with k = sqrt(2 ln(alpha / tau)).
"""
import ctypes

import numpy as np
import pytest

from harness import XPBD
import scenes


def cm(a):
    """numpy row-major 3x3 -> column-major float32[9]"""
    return np.ascontiguousarray(np.asarray(a, np.float32).T.reshape(-1))


def from_cm(v):
    return np.asarray(v, np.float64).reshape(3, 3).T


@pytest.fixture(scope="module")
def lib(oracle_path):
    L = ctypes.CDLL(oracle_path)
    L.oracle_polar_double.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.oracle_polar_double.restype = ctypes.c_int
    L.oracle_eig_sym.argtypes = [ctypes.c_void_p] * 3
    return L


def test_polar_double_matches_svd(lib):
    rng = np.random.default_rng(3)
    for _ in range(200):
        A = rng.standard_normal((3, 3))
        A /= np.linalg.norm(A)
        U, s, Vt = np.linalg.svd(A)
        if s[-1] < 0.08 * s[0]:
            continue
        D = np.diag([1, 1, np.sign(np.linalg.det(U @ Vt))])
        R_ref = U @ D @ Vt
        a = cm(A)
        r = np.zeros(9, np.float32)
        assert lib.oracle_polar_double(a.ctypes.data, r.ctypes.data) == 1
        R = from_cm(r)
        if np.linalg.det(A) > 0:
            np.testing.assert_allclose(R, R_ref, atol=2e-5)
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=2e-6)
        assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-5)


def test_polar_flat_fix_keeps_rotation(lib):
    rng = np.random.default_rng(4)
    for _ in range(50):
        Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        A = Q @ np.diag([1.0, 0.5, 1e-4]) @ np.linalg.qr(rng.standard_normal((3, 3)))[0]
        r = np.zeros(9, np.float32)
        assert lib.oracle_polar_double(cm(A).ctypes.data, r.ctypes.data) == 1
        R = from_cm(r)
        assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-5)


def test_float_eigen_matches_numpy(lib):
    rng = np.random.default_rng(5)
    for _ in range(200):
        B = rng.standard_normal((3, 3))
        S = B @ B.T + 0.1 * np.eye(3)
        ev = np.zeros(3, np.float32)
        V = np.zeros(9, np.float32)
        lib.oracle_eig_sym(cm(S).ctypes.data, ev.ctypes.data, V.ctypes.data)
        np.testing.assert_allclose(ev, np.linalg.eigvalsh(S), rtol=1e-4, atol=1e-5)
        Vm = from_cm(V)
        np.testing.assert_allclose(Vm.T @ Vm, np.eye(3), atol=1e-4)


@pytest.fixture(scope="module")
def scene():
    return scenes.make_ellipsoid_scene(1200, 10, seed=13)


def rot(axis, th):
    return scenes._rot_axis_angle(axis, th)


def sim_at(oracle_path, scene, pos, shape=False):
    sim = XPBD(oracle_path)
    sim.create(scene)
    sim.set_solver(iters=2, dt=1 / 600, shape_compliance=0.0, shape_blend=1.0)
    sim.set_constraints(distance=False, shape=shape, angle=False, volume=False)
    sim.set_ground(False, (0, 0, 1), 0.0, gravity=0.0)
    sim.step()
    sim.ext_set_state(pos, np.zeros_like(pos))
    return sim


def test_shape_matching_keeps_rigid_rotation(oracle_path, scene):
    R = rot([0.3, -0.5, 0.8], 0.7)
    p = (scene["pos"].astype(np.float64) @ R.T + [0.02, 0.0, 0.05]).astype(np.float32)
    sim = sim_at(oracle_path, scene, p, shape=True)
    sim.step()
    ext = scenes.scene_extent(scene)
    assert np.max(np.abs(sim.positions() - p)) / ext < 2e-6
    sim.destroy()


def covariances(scales, quats):
    w, x, y, z = quats.T.astype(np.float64)
    R = np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], 1)
    S2 = scales.astype(np.float64) ** 2
    return np.einsum("nij,nj,nkj->nik", R, S2, R)


def test_deformed_shapes_rest_rotation_scale(oracle_path, scene):
    sim = sim_at(oracle_path, scene, scene["pos"])
    sc, q, k = sim.shapes(1e-2)
    assert k == 0
    np.testing.assert_array_equal(sc, scene["scales"])
    np.testing.assert_array_equal(q, scene["rots"])
    R = rot([0.2, 0.9, -0.4], 0.9)
    p = (scene["pos"].astype(np.float64) @ R.T).astype(np.float32)
    sim.ext_set_state(p, None)
    sc, q, k = sim.shapes(1e-3)
    moved = np.linalg.norm(p - scene["pos"], axis=1) > 1e-3
    deg = np.bincount(scene["edges"].ravel(), minlength=len(p)) >= 3
    sel = moved & deg
    assert k == sel.sum() and k > 100
    C0 = covariances(scene["scales"], scene["rots"])
    C1 = covariances(sc, q)
    expect = np.einsum("ij,njk,lk->nil", R, C0, R)
    rel = np.linalg.norm((C1 - expect)[sel], axis=(1, 2)) / np.linalg.norm(expect[sel], axis=(1, 2))
    assert np.median(rel) < 2e-3
    p2 = (scene["pos"].astype(np.float64) * 1.2).astype(np.float32)
    sim.ext_set_state(p2, None)
    sc2, q2, k2 = sim.shapes(1e-3)
    # the eigen solver may permute the axes, so compare sorted scales
    ratio = np.sort(sc2.astype(np.float64), 1) / np.sort(scene["scales"].astype(np.float64), 1)
    sel2 = (np.linalg.norm(p2 - scene["pos"], axis=1) > 1e-3) & deg
    assert np.median(np.abs(ratio[sel2] - 1.2)) < 2e-3
    np.testing.assert_allclose(covariances(sc2, q2)[sel2], 1.44 * C0[sel2], rtol=0, atol=1e-3 * np.abs(C0).max())
    sim.destroy()


def test_contact_pack_closed_form(oracle_path, scene):
    # the contact matrix is only observable through contact; check the closed
    # form through a ground test: with ellipsoid ground contact on, a Gaussian
    # resting on the plane sits at height sqrt(n^T M n) = k * |S R^T n|
    n = 1
    one = {"pos": np.array([[0.0, 0.0, 0.5]], np.float32), "scales": np.array([[0.02, 0.01, 0.005]], np.float32),
           "rots": np.array([[np.cos(0.3), np.sin(0.3), 0.0, 0.0]], np.float32),
           "opacity": np.array([0.8], np.float32), "edges": np.zeros((0, 2), np.int32),
           "rest": np.zeros(0, np.float32), "stiff": np.zeros(0, np.float32)}
    sim = XPBD(oracle_path)
    sim.create(one)
    sim.set_solver(iters=1, dt=1 / 60, vel_damping=0.0)
    sim.set_constraints(distance=False, shape=False, angle=False, volume=False)
    sim.set_ground(True, (0, 0, 1), 0.0, friction=0.0, restitution=0.0, gravity=9.81)
    sim.set_contact_shape(2, tau=0.2, radius_cap=0.0)
    sim.step()
    sim.ext_set_state(np.array([[0.0, 0.0, -0.1]], np.float32), np.zeros((1, 3), np.float32))
    sim.step()
    z = float(sim.positions()[0, 2])
    k = np.sqrt(2 * np.log(0.8 / 0.2))
    w, x, y, zq = one["rots"][0].astype(np.float64)
    Rm = np.array([[1 - 2 * (y * y + zq * zq), 2 * (x * y - w * zq), 2 * (x * zq + w * y)],
                   [2 * (x * y + w * zq), 1 - 2 * (x * x + zq * zq), 2 * (y * zq - w * x)],
                   [2 * (x * zq - w * y), 2 * (y * zq + w * x), 1 - 2 * (x * x + y * y)]])
    M = Rm @ np.diag((k * one["scales"][0].astype(np.float64)) ** 2) @ Rm.T
    assert z == pytest.approx(np.sqrt(M[2, 2]), rel=1e-5)
    sim.destroy()
    assert n == 1


def cluster_sim(oracle_path, scene, pos, labels, robust=True, iters=1):
    sim = XPBD(oracle_path)
    sim.create(scene)
    sim.ext_set_shape_clusters(labels)
    sim.set_solver(iters=iters, dt=1 / 600, vel_damping=0.0, shape_compliance=0.0, shape_blend=1.0)
    sim.set_constraints(distance=False, shape=True, angle=False, volume=False)
    sim.set_shape_robust(robust)
    sim.set_ground(False, (0, 0, 1), 0.0, gravity=0.0)
    sim.step()
    sim.ext_set_state(pos, np.zeros_like(pos))
    return sim


def test_kmeans_clusters_partition(scene):
    lab = scenes.kmeans_clusters(scene["pos"], size=64, seed=3)
    n = scene["pos"].shape[0]
    assert lab.shape == (n,) and lab.min() == 0
    sizes = np.bincount(lab)
    assert (sizes > 0).all() and len(sizes) <= n // 64
    # each node is nearer its own centroid than any other (Lloyd fixed point up to the last move)
    cent = np.stack([np.bincount(lab, weights=scene["pos"][:, a]) / sizes for a in range(3)], 1)
    d = np.linalg.norm(scene["pos"][:, None, :] - cent[None], axis=2)
    assert (np.argmin(d, 1) == lab).mean() > 0.97


@pytest.mark.parametrize("robust", [True, False])
def test_cluster_shape_matching_goals(oracle_path, scene, robust):
    lab = scenes.kmeans_clusters(scene["pos"], size=64, seed=3)
    rest = scene["pos"].astype(np.float64)
    # a rigid motion leaves every cluster where it is
    R = rot([0.3, -0.5, 0.8], 0.7)
    p = (rest @ R.T + [0.02, 0.0, 0.05]).astype(np.float32)
    sim = cluster_sim(oracle_path, scene, p, lab, robust)
    sim.step()
    ext = scenes.scene_extent(scene)
    assert np.max(np.abs(sim.positions() - p)) / ext < 2e-6
    sim.destroy()
    # an affine map moves every cluster onto its best rigid fit (blend 1)
    A = np.array([[1.15, 0.1, 0.0], [0.0, 0.9, 0.08], [0.05, 0.0, 1.05]])
    p = (rest @ A.T).astype(np.float32)
    sim = cluster_sim(oracle_path, scene, p, lab, robust)
    sim.step()
    got = sim.positions().astype(np.float64)
    worst = 0.0
    for k in range(lab.max() + 1):
        m = lab == k
        if m.sum() < 4:
            continue
        q = rest[m] - rest[m].mean(0)
        c = p[m].astype(np.float64).mean(0)
        U, _, Vt = np.linalg.svd((p[m] - c).T @ q)
        Rk = U @ np.diag([1, 1, np.sign(np.linalg.det(U @ Vt))]) @ Vt
        worst = max(worst, np.abs(got[m] - (c + q @ Rk.T)).max())
    assert worst / ext < (2e-5 if robust else 2e-4)
    sim.destroy()
