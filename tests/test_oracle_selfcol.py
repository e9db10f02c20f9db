"""Pins the oracle's self-collision and object-shape restatements (CPU only).

  * cell keys == the spatial hash of numpy float32 cell coordinates (fixed
    grid and the displacement rule at rest, where the cell is the search
    radius exactly);
  * node order == numpy's stable argsort of the keys;
  * candidate lists == an O(N^2) brute force of the pair rules (cross-object
    pairs closer than the search radius with rest distance >= d_c; same-object
    pairs only with withinBody, beyond the exclusion radius) when no list
    overflows: the 27-cell neighbourhood must find every such pair;
  * a resting single body produces no cross-object candidates;
  * self-collision separates two interpenetrating bodies (more than without);
  * object shape matching moves an affinely deformed object onto the
    best-fit rigid transform of its rest shape (Kabsch by numpy's SVD);
  * the active set after the first object-shape substep holds exactly the
    nodes inside the other object's box inflated by search + 2 d_c.
"""
import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

from harness import XPBD
import scenes

H1, H2, H3 = np.uint32(73856093), np.uint32(19349663), np.uint32(83492791)


def spacing_of(scene):
    """Median of every node's shortest positive rest length (float32)."""
    n = scene["pos"].shape[0]
    e, r = scene["edges"], scene["rest"]
    mins = np.full(n, np.inf, np.float32)
    ok = (r > 0) & np.isfinite(r)
    np.minimum.at(mins, e[ok, 0], r[ok])
    np.minimum.at(mins, e[ok, 1], r[ok])
    mins = np.sort(mins[np.isfinite(mins)])
    return np.float32(mins[len(mins) // 2])


def components(scene, min_size=50):
    n = scene["pos"].shape[0]
    e = scene["edges"]
    A = coo_matrix((np.ones(len(e)), (e[:, 0], e[:, 1])), shape=(n, n))
    _, lab = connected_components(A, directed=False)
    size = np.bincount(lab)
    return np.where(size[lab] >= min_size, lab, -1)


def cell_keys(x, inv_cell, table):
    c = np.floor(np.clip(x * np.float32(inv_cell), np.float32(-1e9), np.float32(1e9))).astype(np.int64)
    c = (c & 0xFFFFFFFF).astype(np.uint32)
    h = (c[:, 0] * H1) ^ (c[:, 1] * H2) ^ (c[:, 2] * H3)
    return h & np.uint32(table - 1)


def brute_candidates(x, rest, comp, dc, search, excl2, within):
    """Candidate sets of every node by the pair rules, in float32 like the solver."""
    n = len(x)
    out = []
    for i in range(n):
        d = x[i] - x
        d2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
        q = rest[i] - rest
        q2 = q[:, 0] * q[:, 0] + q[:, 1] * q[:, 1] + q[:, 2] * q[:, 2]
        cross = comp != comp[i]
        cand = cross & (np.sqrt(d2) < search) & (q2 >= dc * dc)
        if within:
            cand |= (~cross) & (d2 < search * search) & (q2 >= excl2)
        cand[i] = False
        out.append(set(np.nonzero(cand)[0].tolist()))
    return out


def still_sim(path, scene, sc=True, within=False, obj=0.0, cell=0.0):
    """Solver with no motion: predicted positions equal the committed ones."""
    sim = XPBD(path)
    sim.create(scene)
    sim.set_solver(iters=2, dt=1 / 600, vel_damping=0.0)
    sim.set_constraints(distance=False, shape=False, angle=False, volume=False)
    sim.set_ground(False, (0, 0, 1), 0.0, gravity=0.0)
    sim.set_colliders([], [], [])
    sim.set_object_shape(obj)
    sim.ext_set_self_collision_grid(cell, 0)
    sim.set_self_collision(sc, 1.5, 2.0, within)
    return sim


@pytest.fixture(scope="module")
def pair():
    # two spheres whose surfaces nearly touch
    return scenes.stacked_bodies(1200, 8, semi=(0.08, 0.08, 0.08), gap=0.004, seed=11)


@pytest.mark.parametrize("within,cell", [(False, 0.0), (False, 0.05), (True, 0.05)])
def test_keys_order_candidates_brute_force(oracle_path, pair, within, cell):
    sim = still_sim(oracle_path, pair, within=within, cell=cell)
    sim.step()
    x = sim.positions()
    sim.ext_set_state(x, np.zeros_like(x))
    sim.step()
    keys, order, counts, contacts = sim.ext_self_collision_lists()
    n = x.shape[0]
    table = 1024
    while table < 2 * n:
        table *= 2
    dc = np.float32(1.5) * spacing_of(pair)
    search = np.float32(1.5) * dc
    excl = np.float32(2.0) * dc
    reach = np.float32(cell) if cell > 0 else search
    inv = np.float32(1.0) / reach
    np.testing.assert_array_equal(keys, cell_keys(x, inv, table))
    np.testing.assert_array_equal(order, np.argsort(keys, kind="stable"))
    comp = components(pair)
    assert len(set(comp.tolist()) - {-1}) == 2
    ref = brute_candidates(x, pair["pos"], comp, dc, search, excl * excl, within)
    stats = sim.self_collision_stats()
    assert stats[2] == 0, "a list overflowed; the brute force only holds without overflow"
    got = [set(contacts[i, :counts[i]].tolist()) for i in range(n)]
    assert got == ref
    assert stats[0] == sum(len(s) for s in ref) > 0
    assert stats[1] == n
    sim.destroy()


def test_single_body_at_rest_has_no_cross_candidates(oracle_path):
    one = scenes.make_ellipsoid_scene(1500, 8, semi=(0.1, 0.08, 0.06), seed=4)
    sim = still_sim(oracle_path, one)
    sim.step()
    sim.step()
    _, _, counts, _ = sim.ext_self_collision_lists()
    assert counts.sum() == 0 and sim.self_collision_stats()[0] == 0
    sim.destroy()


def min_cross_distance(x, comp):
    a, b = x[comp == 0], x[comp == 1]
    from scipy.spatial import cKDTree
    return float(cKDTree(b).query(a)[0].min())


def test_self_collision_separates_bodies(oracle_path):
    # the upper sphere falls onto the lower one; rest shapes do not overlap
    scene = scenes.stacked_bodies(1000, 8, semi=(0.08, 0.08, 0.08), gap=0.004, seed=21)
    comp = components(scene)
    dc = float(np.float32(1.5) * spacing_of(scene))
    vel = np.zeros_like(scene["pos"])
    vel[comp == 1, 2] = -1.0
    out = {}
    for on in (False, True):
        sim = XPBD(oracle_path)
        sim.create(scene)
        sim.set_solver(iters=10, dt=1 / 600, vel_damping=0.0)
        sim.set_constraints(distance=True, shape=False, angle=False, volume=False)
        sim.set_ground(False, (0, 0, 1), 0.0, gravity=0.0)
        sim.set_colliders([], [], [])
        sim.set_object_shape(0.0)
        sim.ext_set_self_collision_grid(0.0, 0)
        sim.set_self_collision(on, 1.5, 2.0, False)
        sim.step()
        sim.ext_set_state(scene["pos"], vel)
        for _ in range(40):
            sim.step()
        out[on] = min_cross_distance(sim.positions(), comp)
        sim.destroy()
    assert out[False] < 0.2 * dc
    assert out[True] > 0.5 * dc


def test_object_shape_goes_to_best_rigid_fit(oracle_path):
    body = scenes.make_ellipsoid_scene(800, 8, semi=(0.1, 0.07, 0.05), seed=8)
    rest = body["pos"].astype(np.float64)
    A = np.array([[1.10, 0.08, -0.03], [0.02, 0.93, 0.05], [-0.04, 0.06, 1.04]])
    th = 0.4
    Rz = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    c0 = rest.mean(0)
    p = ((rest - c0) @ (Rz @ A).T + c0 + [0.01, -0.02, 0.03]).astype(np.float32)
    sim = still_sim(oracle_path, body, sc=False, obj=1.0)
    sim.step()
    sim.ext_set_state(p, np.zeros_like(p))
    sim.step()
    got = sim.positions().astype(np.float64)
    c0f = body["pos"].astype(np.float64).mean(0).astype(np.float32).astype(np.float64)
    q = body["pos"].astype(np.float64) - c0f
    c = p.astype(np.float64).mean(0)
    U, _, Vt = np.linalg.svd((p - c).T @ q)
    R = U @ np.diag([1, 1, np.sign(np.linalg.det(U @ Vt))]) @ Vt
    goal = c + q @ R.T
    ext = scenes.scene_extent(body)
    assert np.max(np.abs(got - goal)) / ext < 2e-5
    sim.destroy()


def test_active_set_is_the_inflated_other_box(oracle_path):
    scene = scenes.stacked_bodies(900, 8, semi=(0.08, 0.08, 0.08), gap=0.025, seed=31)
    comp = components(scene)
    sim = still_sim(oracle_path, scene, obj=0.5)
    sim.step()
    assert sim.self_collision_stats()[1] == scene["pos"].shape[0]   # no valid shift yet
    x = sim.positions()
    sim.ext_set_state(x, np.zeros_like(x))
    sim.step()
    dc = np.float32(1.5) * spacing_of(scene)
    inflate = float(np.float32(1.5) * dc + np.float32(2.0) * dc)
    counts = []
    for eps in (-1e-6, 1e-6):
        act = np.zeros(len(x), bool)
        for k in (0, 1):
            other = x[comp == 1 - k]
            lo, hi = other.min(0) - inflate - eps, other.max(0) + inflate + eps
            inside = np.all((x >= lo) & (x <= hi), axis=1)
            act |= (comp == k) & inside
        counts.append(int(act.sum()))
    active = sim.self_collision_stats()[1]
    assert counts[0] <= active <= counts[1]
    assert 0 < active < len(x)
    sim.destroy()
