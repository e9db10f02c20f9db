"""Oracle pins for the umbrella angle constraint (forward.cu:4990-5130,
10708-10819), on the CPU:

  * pairing of a planar star follows the angular order around the centre,
    and a gap wider than 120 degrees leaves the entry unpaired;
  * on a random graph every pair is (entry, next neighbour around the centre)
    with each neighbour following at most one other, and its rest cosine is
    the cosine of the two rest directions;
  * the rest shape and any rigid motion of it are left in place;
  * a bent three-node hinge opens back toward its rest angle (a sign error in
    the gradient or the update would close it further).
"""
import numpy as np
import pytest

from harness import XPBD
import scenes


def create(sim, pos, edges):
    n = pos.shape[0]
    e = np.asarray(edges, np.int32).reshape(-1, 2)
    d = np.linalg.norm(pos[e[:, 0]].astype(np.float64) - pos[e[:, 1]].astype(np.float64), axis=1).astype(np.float32)
    q = np.zeros((n, 4), np.float32)
    q[:, 0] = 1.0
    sim.create({"pos": np.ascontiguousarray(pos, np.float32), "scales": np.full((n, 3), 0.01, np.float32),
                "rots": q, "opacity": np.full(n, 0.9, np.float32), "edges": np.ascontiguousarray(e),
                "rest": d, "stiff": np.ones(len(e), np.float32)})


def angle_only(sim, iters=1, compliance=0.0, blend=1.0):
    sim.set_solver(iters=iters, dt=1 / 600, under_relax=0.6, vel_damping=0.0, angle_compliance=compliance,
                   angle_blend=blend)
    sim.set_constraints(distance=False, shape=False, angle=True, volume=False)
    sim.set_ground(False, (0, 0, 1), 0.0, gravity=0.0)
    sim.set_colliders([], [], [])


def test_star_pairing(oracle_path):
    sim = XPBD(oracle_path)
    # centre 0, neighbours at 0, 60, 150, 200, 280 degrees in the z=0 plane;
    # the longest (neighbour 1, angle 0) defines u
    angles = [0.0, 60.0, 150.0, 200.0, 280.0]
    radii = [1.0, 0.9, 0.8, 0.85, 0.7]
    pos = [[0.0, 0.0, 0.0]] + [[r * np.cos(np.deg2rad(a)), r * np.sin(np.deg2rad(a)), 0.0] for a, r in zip(angles, radii)]
    pos = np.array(pos, np.float32)
    create(sim, pos, [[0, j] for j in range(1, 6)])
    pair, cos = sim.ext_angle_pairs()
    # entries of centre 0 are its first five (input edge order)
    expected_next = {1: 2, 2: 3, 3: 4, 4: 5, 5: 1}
    for k in range(5):
        j = k + 1
        assert pair[k] == expected_next[j]
        a0, a1 = np.deg2rad(angles[k]), np.deg2rad(angles[expected_next[j] - 1])
        assert cos[k] == pytest.approx(np.cos(a1 - a0), abs=1e-6)
    # a gap of 150 degrees (between 150 and 300 with the 200 and 280 ones removed)
    sim.destroy()
    pos2 = pos[[0, 1, 2, 3]]
    create(sim, pos2, [[0, 1], [0, 2], [0, 3]])
    pair, cos = sim.ext_angle_pairs()
    # around: 0 -> 60 -> 150 -> (210 back to 0): the last gap is > 120 degrees
    assert pair[0] == 2 and pair[1] == 3
    assert pair[2] == -1 and cos[2] == 2.0
    sim.destroy()


def test_random_graph_pairing_is_angular_cycle(oracle_path):
    sim = XPBD(oracle_path)
    scene = scenes.make_ellipsoid_scene(400, 8, semi=(0.05, 0.04, 0.03), seed=21)
    sim.create(scene)
    pair, cos = sim.ext_angle_pairs()
    n = scene["pos"].shape[0]
    # rebuild the reference CSR (both directions, input edge order)
    rows = [[] for _ in range(n)]
    for a, b in scene["edges"]:
        rows[a].append(b)
        rows[b].append(a)
    pos = scene["pos"].astype(np.float64)
    e = 0
    paired = 0
    for i in range(n):
        row = rows[i]
        nxts = []
        for j in row:
            p = pair[e]
            if p >= 0:
                assert p in row and p != j
                v1, v2 = pos[j] - pos[i], pos[p] - pos[i]
                c = np.dot(v1, v2) / np.linalg.norm(v1) / np.linalg.norm(v2)
                assert cos[e] == pytest.approx(c, abs=2e-6)
                nxts.append(p)
                paired += 1
            else:
                assert cos[e] == 2.0
            e += 1
        assert len(nxts) == len(set(nxts))   # each neighbour follows at most one other
    assert e == len(pair) and paired > 0.8 * e
    sim.destroy()


@pytest.mark.parametrize("rigid", [False, True])
def test_rest_and_rigid_motion_stay(oracle_path, rigid):
    sim = XPBD(oracle_path)
    scene = scenes.make_ellipsoid_scene(300, 8, semi=(0.05, 0.04, 0.03), seed=22)
    sim.create(scene)
    angle_only(sim, iters=2)
    sim.step()
    pos = scene["pos"].astype(np.float64)
    if rigid:
        th = 0.7
        R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
        pos = pos @ R.T + np.array([0.1, -0.2, 0.05])
    pos = pos.astype(np.float32)
    sim.ext_set_state(pos, np.zeros_like(pos))
    sim.step()
    moved = np.abs(sim.positions().astype(np.float64) - pos).max()
    assert moved < 2e-6 * np.abs(pos).max() + 1e-7
    sim.destroy()


def test_bent_hinge_opens(oracle_path):
    sim = XPBD(oracle_path)
    # centre 0 with neighbours along x and y (rest 90 degrees); nodes 1 and 2
    # have one neighbour each, so only the pair around 0 exists
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    create(sim, pos, [[0, 1], [0, 2]])
    angle_only(sim, iters=1, blend=1.0)
    sim.step()
    pair, cos = sim.ext_angle_pairs()
    assert pair[0] == 2 and cos[0] == pytest.approx(0.0, abs=1e-7)
    bent = pos.copy()
    bent[2] = [np.cos(np.deg2rad(60.0)), np.sin(np.deg2rad(60.0)), 0.0]
    sim.ext_set_state(bent, np.zeros_like(bent))

    def ang():
        p = sim.positions().astype(np.float64)
        v1, v2 = p[1] - p[0], p[2] - p[0]
        return np.degrees(np.arccos(np.dot(v1, v2) / np.linalg.norm(v1) / np.linalg.norm(v2)))

    seen = [60.0]
    for _ in range(20):
        sim.ext_set_state(sim.positions(), np.zeros_like(bent))   # no velocity carried over
        sim.step()
        seen.append(ang())
    # the first pass opens the hinge; the passes settle on the rest angle
    assert seen[1] > seen[0] + 5.0
    assert abs(seen[-1] - 90.0) < 1e-3
    sim.destroy()
