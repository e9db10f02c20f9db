"""Pins the oracle's graph load (ChainMail::loadGraph, forward.cu:2702-2779)
(CPU only): invalid and self edges are dropped, every edge acts on both of
its ends, and a repeated edge is two constraints.  Checked through the
Jacobi pass: a list with dropped entries steps bit-identically to the same
list without them, a pair given as (0, 1) or as (1, 0) gives mirrored moves,
and a doubled edge doubles the summed correction before the per-node average
(so it moves the node like the single edge, with the average over two
entries).  The deduplicated Gauss-Seidel edge list is pinned against numpy in
tests/test_oracle_gs.py.
"""
import numpy as np

from harness import XPBD
import scenes


def run(path, scene, steps=3):
    sim = XPBD(path)
    sim.create(scene)
    sim.set_solver(iters=2, dt=1 / 600, under_relax=0.6, vel_damping=0.0, dist_compliance=1e-6)
    sim.set_constraints(distance=True, shape=False, angle=False, volume=False)
    sim.set_ground(True, (0, 0, 1), 0.0, gravity=9.81)
    for _ in range(steps):
        sim.step()
    x, v = sim.ext_get_state()
    sim.destroy()
    return x, v


def with_edges(scene, edges, rest):
    out = dict(scene)
    out["edges"] = np.ascontiguousarray(np.asarray(edges, np.int32).reshape(-1, 2))
    out["rest"] = np.ascontiguousarray(np.asarray(rest, np.float32))
    out["stiff"] = np.ones(out["edges"].shape[0], np.float32)
    return out


def test_invalid_and_self_edges_are_dropped(oracle_path):
    scene = scenes.make_ellipsoid_scene(600, 6, seed=9)
    e, r = scene["edges"], scene["rest"]
    n = scene["pos"].shape[0]
    rng = np.random.default_rng(4)
    bad = np.array([[-1, 3], [5, n], [7, 7], [n + 4, -2], [11, 11]], np.int32)
    rb = rng.uniform(0.01, 0.02, len(bad)).astype(np.float32)
    # interleave the bad entries at random places, keeping the good order
    at = np.sort(rng.choice(len(e) + len(bad), len(bad), replace=False))
    mixed_e, mixed_r, gi, bi = [], [], 0, 0
    for k in range(len(e) + len(bad)):
        if bi < len(bad) and k == at[bi]:
            mixed_e.append(bad[bi]); mixed_r.append(rb[bi]); bi += 1
        else:
            mixed_e.append(e[gi]); mixed_r.append(r[gi]); gi += 1
    xa, va = run(oracle_path, scene)
    xb, vb = run(oracle_path, with_edges(scene, mixed_e, mixed_r))
    np.testing.assert_array_equal(xa, xb)
    np.testing.assert_array_equal(va, vb)


def test_edge_acts_on_both_ends_in_either_direction(oracle_path):
    base = {
        "pos": np.array([[0.0, 0.0, 0.5], [0.13, 0.0, 0.5]], np.float32),
        "scales": np.full((2, 3), 0.01, np.float32),
        "rots": np.tile(np.array([1, 0, 0, 0], np.float32), (2, 1)),
        "opacity": np.full(2, 0.9, np.float32),
    }
    xa, _ = run(oracle_path, with_edges(base, [[0, 1]], [0.1]), steps=1)
    xb, _ = run(oracle_path, with_edges(base, [[1, 0]], [0.1]), steps=1)
    np.testing.assert_array_equal(xa, xb)
    # both ends moved toward each other by the same amount (mirror symmetry)
    d0 = xa[0, 0] - 0.0
    d1 = 0.13 - xa[1, 0]
    assert d0 > 0 and d1 > 0
    np.testing.assert_allclose(d0, d1, rtol=1e-6)


def test_repeated_edge_is_two_entries(oracle_path):
    """Doubled (0, 1): node 0 has two entries with the same correction, summed
    then divided by two, so the first iteration moves it as the single edge
    does; the multipliers of both entries are then updated separately, so the
    second iteration differs from the single edge's (two constraints)."""
    base = {
        "pos": np.array([[0.0, 0.0, 0.5], [0.13, 0.0, 0.5]], np.float32),
        "scales": np.full((2, 3), 0.01, np.float32),
        "rots": np.tile(np.array([1, 0, 0, 0], np.float32), (2, 1)),
        "opacity": np.full(2, 0.9, np.float32),
    }
    sims = []
    for edges, rest in (([[0, 1]], [0.1]), ([[0, 1], [0, 1]], [0.1, 0.1])):
        sim = XPBD(oracle_path)
        sim.create(with_edges(base, edges, rest))
        sim.set_solver(iters=1, dt=1 / 600, under_relax=1.0, vel_damping=0.0, dist_compliance=1e-3)
        sim.set_constraints(distance=True, shape=False, angle=False, volume=False)
        sim.set_ground(False, (0, 0, 1), 0.0, gravity=0.0)
        sim.step()
        sims.append(sim)
    xa, _ = sims[0].ext_get_state()
    xb, _ = sims[1].ext_get_state()
    np.testing.assert_allclose(xa, xb, atol=1e-7)
    for s in sims:
        s.destroy()
