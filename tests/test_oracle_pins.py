"""More pins of the oracle (CPU), each against a closed form or an
independent brute force:

  * candidate-list overflow (forward.cu:7079-7180): on a crowded toy every
    node's list is the 32 nearest accepted candidates (the replace-the-worst
    rule of 7158-7172 keeps the 32 smallest priorities), counts and the
    overflow total follow from an O(N^2) count;
  * farthest-point member selection under the cap (forward.cu:9995-10030):
    every cluster equals a numpy farthest-point sampling of the node's k-hop
    ball (breadth-first, CSR order), seeded at the leader;
  * self-collision separation (7205-7322): an overlapping same-object pair
    ends exactly d_c apart; two plates closing fast are pushed along their
    normal until the gap is d_c;
  * the contact-average push (11858-11890): nodes without a cross-object
    correction all move by the same vector, stiffness times the mean of the
    corrected nodes' moves;
  * one tetrahedron's volume step (8531-8700, 9187-9243): J = V / V0 from the
    signed volume, the XPBD multiplier with the analytic volume gradient, the
    0.25 rest-length step clamp;
  * one hinge's angle step (4990-5130): the XPBD multiplier with the
    finite-difference gradient of the cosine.
"""
import numpy as np
import pytest

from harness import XPBD
import scenes

DT = 1.0 / 600.0


def make_scene(pos, edges):
    pos = np.ascontiguousarray(pos, np.float32)
    n = pos.shape[0]
    e = np.ascontiguousarray(np.asarray(edges, np.int32).reshape(-1, 2))
    d = np.linalg.norm(pos[e[:, 0]].astype(np.float64) - pos[e[:, 1]].astype(np.float64), axis=1).astype(np.float32)
    q = np.zeros((n, 4), np.float32)
    q[:, 0] = 1.0
    return {"pos": pos, "scales": np.full((n, 3), 0.01, np.float32), "rots": q,
            "opacity": np.full(n, 0.9, np.float32), "edges": e, "rest": d, "stiff": np.ones(len(e), np.float32)}


def quiet(sim, iters=1, distance=False, under_relax=1.0):
    """No gravity, ground or colliders; one iteration per substep."""
    sim.set_solver(iters=iters, dt=DT, under_relax=under_relax, vel_damping=0.0, dist_compliance=1e-6,
                   angle_compliance=0.0, angle_blend=1.0)
    sim.set_constraints(distance=distance, shape=False, angle=False, volume=False)
    sim.set_ground(False, (0, 0, 1), 0.0, gravity=0.0)
    sim.set_colliders([], [], [])


def grid(nx, ny, nz, h, origin=(0.0, 0.0, 0.0)):
    ii = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    pos = ii * h + np.asarray(origin)
    idx = (ii[:, 0] * ny + ii[:, 1]) * nz + ii[:, 2]
    lut = np.full((nx, ny, nz), -1)
    lut[ii[:, 0], ii[:, 1], ii[:, 2]] = idx
    edges = []
    for a, (dx, dy, dz) in enumerate(((1, 0, 0), (0, 1, 0), (0, 0, 1))):
        m = (ii[:, 0] + dx < nx) & (ii[:, 1] + dy < ny) & (ii[:, 2] + dz < nz)
        src = idx[m]
        dst = lut[ii[m, 0] + dx, ii[m, 1] + dy, ii[m, 2] + dz]
        edges.append(np.stack([src, dst], 1))
    return pos.astype(np.float32), np.concatenate(edges).astype(np.int32)


def f32_d2(a, b):
    e = (a.astype(np.float32) - b.astype(np.float32)).astype(np.float32)
    return ((e[..., 0] * e[..., 0] + e[..., 1] * e[..., 1]).astype(np.float32) + e[..., 2] * e[..., 2]).astype(np.float32)


# ---------------------------------------------------------------------------
def test_overflowing_lists_keep_the_32_nearest(oracle_path):
    h = 0.01
    pos, edges = grid(6, 6, 6, h)
    n = pos.shape[0]
    sim = XPBD(oracle_path)
    sim.create(make_scene(pos, edges))
    quiet(sim)
    sim.set_self_collision(True, 1.5, 2.0, True)
    sim.step()
    rng = np.random.default_rng(3)
    u = rng.standard_normal((n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    crowd = (u * (1.2 * h) * rng.uniform(0, 1, (n, 1)) ** (1 / 3)).astype(np.float32)
    sim.ext_set_state(crowd, np.zeros_like(crowd))
    sim.step()
    keys, order, counts, contacts = sim.ext_self_collision_lists()
    _, _, overflow = sim.self_collision_stats()
    dc = np.float32(1.5) * np.float32(h)
    search, excl = np.float32(1.5) * dc, np.float32(2.0) * dc
    d2 = f32_d2(crowd[:, None, :], crowd[None, :, :])
    q2 = f32_d2(pos[:, None, :], pos[None, :, :])
    acc = (d2 < search * search) & (q2 >= excl * excl)
    np.fill_diagonal(acc, False)
    nacc = acc.sum(1)
    assert nacc.min() > 32, "the toy must overflow every list"
    pr = np.sqrt(d2.astype(np.float32))
    for i in range(n):
        cand = np.flatnonzero(acc[i])
        keep = cand[np.argsort(pr[i, cand], kind="stable")[:32]]
        assert np.sort(pr[i, keep])[-1] < np.sort(pr[i, cand])[32], "a tie at the 32nd place"
        assert set(contacts[i, :counts[i]].tolist()) == set(keep.tolist()), f"node {i}"
    np.testing.assert_array_equal(counts, np.minimum(nacc, 32))
    assert overflow == int((nacc - 32).clip(min=0).sum())
    sim.destroy()


def bfs_ball(rows, i, k):
    seen = {i}
    frontier, out = [i], []
    for _ in range(k):
        nxt = []
        for f in frontier:
            for j in rows[f]:
                if j not in seen:
                    seen.add(j)
                    out.append(j)
                    nxt.append(j)
        frontier = nxt
    return out


def numpy_fps(pos, leader, cand, cap):
    P = pos[cand]
    md = f32_d2(P, pos[leader][None, :])
    taken = np.zeros(len(cand), bool)
    pick = []
    cur = int(np.argmax(md))
    for _ in range(cap):
        taken[cur] = True
        pick.append(cand[cur])
        md = np.minimum(md, f32_d2(P, P[cur][None, :]))
        free = np.where(taken, -1.0, md)
        if free.max() < 0:
            break
        cur = int(np.argmax(free))
    return pick


@pytest.mark.parametrize("k,cap", [(2, 16), (3, 24)])
def test_member_cap_is_farthest_point_sampling(oracle_path, k, cap):
    scene = scenes.make_ellipsoid_scene(500, 8, seed=14)
    sim = XPBD(oracle_path)
    sim.create(scene)
    quiet(sim, distance=True)
    sim.set_constraints(distance=True, shape=False, angle=False, volume=True)
    sim.set_volume(1e-6, k, cap, 1)
    sim.step()
    members = sim.ext_volume_members()
    n = scene["pos"].shape[0]
    rows = [[] for _ in range(n)]
    for a, b in scene["edges"]:
        rows[a].append(int(b))
        rows[b].append(int(a))
    sampled = 0
    for i in range(n):
        ball = bfs_ball(rows, i, k)
        want = ball if len(ball) <= cap else numpy_fps(scene["pos"], i, np.array(ball), cap)
        sampled += len(ball) > cap
        np.testing.assert_array_equal(members[i], want, err_msg=f"cluster {i}")
    assert sampled > n // 2
    sim.destroy()


# ---------------------------------------------------------------------------
def test_overlapping_pair_separates_to_dc(oracle_path):
    h = 0.01
    pos = np.array([[0, 0, 0], [h, 0, 0], [-5 * h, 0, 0]], np.float32)
    sim = XPBD(oracle_path)
    sim.create(make_scene(pos, [[0, 1]]))
    quiet(sim)
    sim.set_self_collision(True, 1.5, 2.0, True)
    sim.step()
    dc = 1.5 * h
    cur = np.array([[0, 0, 0], [h, 0, 0], [-0.5 * dc, 0, 0]], np.float32)
    sim.ext_set_state(cur, np.zeros_like(cur))
    sim.step()
    p = sim.positions().astype(np.float64)
    assert np.linalg.norm(p[0] - p[2]) == pytest.approx(dc, rel=1e-6)
    np.testing.assert_allclose(p[0], [0.25 * dc, 0, 0], atol=1e-9)     # each moves half the overlap
    np.testing.assert_allclose(p[2], [-0.75 * dc, 0, 0], atol=1e-9)
    np.testing.assert_array_equal(p[1], cur[1])                        # within search, beyond d_c: untouched
    sim.destroy()


def plates(n_a, n_b, h, gap_rest):
    """Two flat square grids in z, A above B, centred on the z axis."""
    pa, ea = grid(n_a, n_a, 1, h, origin=(-(n_a - 1) * h / 2, -(n_a - 1) * h / 2, gap_rest / 2))
    pb, eb = grid(n_b, n_b, 1, h, origin=(-(n_b - 1) * h / 2, -(n_b - 1) * h / 2, -gap_rest / 2))
    return np.concatenate([pa, pb]), np.concatenate([ea, eb + len(pa)]), len(pa)


def closing_plates(sim, n_a, n_b, h, G, g, stiffness=0.0):
    pos, edges, na = plates(n_a, n_b, h, 40 * h)
    sim.create(make_scene(pos, edges))
    quiet(sim, distance=True)
    sim.set_object_shape(stiffness)
    sim.set_self_collision(True, 1.5, 2.0, False)
    sim.step()
    x = pos.copy()
    x[:na, 2] = G / 2
    x[na:, 2] = -G / 2
    u = (G - g) / 2 / DT
    v = np.zeros_like(x)
    v[:na, 2] = -u
    v[na:, 2] = u
    sim.ext_set_state(x, v)
    pred = (x.astype(np.float64) + v * DT)
    sim.step()
    return pred, sim.positions().astype(np.float64), na


def test_fast_plates_separate_to_dc(oracle_path):
    h = 0.01
    dc = 1.5 * h
    sim = XPBD(oracle_path)
    G, g = 1.6 * dc, 0.6 * dc
    pred, p, na = closing_plates(sim, 14, 14, h, G, g)
    ij = np.stack(np.meshgrid(np.arange(14), np.arange(14), indexing="ij"), -1).reshape(-1, 2)
    interior = np.flatnonzero(((ij >= 3) & (ij <= 10)).all(1))
    for part, sign in ((interior, 1.0), (interior + na, -1.0)):
        d = p[part] - pred[part]
        np.testing.assert_allclose(d[:, 2], sign * (dc - g) / 2, rtol=2e-5)
        assert np.abs(d[:, :2]).max() < 1e-6 * dc
    gap = p[interior, 2] - p[interior + na, 2]
    np.testing.assert_allclose(gap, dc, rtol=2e-5)
    sim.destroy()


def test_contact_push_moves_the_rest_by_the_mean(oracle_path):
    h = 0.01
    dc = 1.5 * h
    s = 0.5
    sim = XPBD(oracle_path)
    pred, p, na = closing_plates(sim, 16, 8, h, 1.6 * dc, 0.6 * dc, stiffness=s)   # both >= 50 nodes: objects
    # the large plate A: its rim beyond plate B gets no cross-object correction
    for part in (np.arange(na),):
        d = (p[part] - pred[part]).astype(np.float32)
        vals, cnt = np.unique(d, axis=0, return_counts=True)
        T = vals[np.argmax(cnt)]                      # the one vector every uncorrected node moved by
        moved = np.any(d != T, axis=1)
        assert cnt.max() >= 10 and moved.sum() >= 10
        np.testing.assert_allclose(T, s * d[moved].astype(np.float64).mean(0), rtol=1e-4, atol=1e-9)
    sim.destroy()


# ---------------------------------------------------------------------------
def tet_volume(x):
    return np.linalg.det(np.stack([x[1] - x[0], x[2] - x[0], x[3] - x[0]])) / 6.0


def tet_volume_grad(x):
    g1 = np.cross(x[2] - x[0], x[3] - x[0]) / 6.0
    g2 = np.cross(x[3] - x[0], x[1] - x[0]) / 6.0
    g3 = np.cross(x[1] - x[0], x[2] - x[0]) / 6.0
    return np.stack([-(g1 + g2 + g3), g1, g2, g3])


@pytest.mark.parametrize("squash,near", [(0.95, False), (0.6, True)])
def test_tetrahedron_volume_step(oracle_path, squash, near):
    a = 0.01
    rest = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], np.float64) * (a / (2 * np.sqrt(2)))
    rest += [0.0, 0.0, 0.05]
    edges = [[0, 1], [0, 2], [0, 3], [1, 2], [1, 3], [2, 3]]
    if near:
        # a fifth node (no cluster of its own) 0.2 a from node 0: node 0's
        # rest length, hence its step limit, is 0.25 * 0.2 a
        rest = np.vstack([rest, rest[0] + [0.2 * a, 0.0, 0.0]])
        edges.append([0, 4])
    n = rest.shape[0]
    sim = XPBD(oracle_path)
    sim.create(make_scene(rest, edges))
    quiet(sim, under_relax=0.7)
    compliance = 1e-9
    sim.set_constraints(distance=False, shape=False, angle=False, volume=True)
    sim.set_volume(compliance, 1, 64, 1)
    off = np.array([0] + [3] * n, np.int32)
    sim.ext_set_volume_clusters(off, np.array([1, 2, 3], np.int32), 3)
    sim.step()
    c = rest[:4].mean(0)
    cur = rest.copy()
    cur[:4] = c + (rest[:4] - c) * [1.0, 1.0, squash]
    cur = cur.astype(np.float32)
    sim.ext_set_state(cur, np.zeros_like(cur))
    sim.step()
    x = cur.astype(np.float64)
    V0, V = tet_volume(rest[:4].astype(np.float32).astype(np.float64)), tet_volume(x[:4])
    J = V / V0
    grad = tet_volume_grad(x[:4]) / V0
    alpha = compliance / DT ** 2
    dlam = -(J - 1.0) / ((grad * grad).sum() + alpha)
    dx = dlam * grad
    step = np.linalg.norm(dx, axis=1)
    limit = 0.25 * np.array([0.2 * a if near else a, a, a, a])
    assert (step > limit).any() == near
    dx = np.where((step > limit)[:, None], dx * (limit / step)[:, None], dx)
    moved = sim.positions().astype(np.float64) - x
    np.testing.assert_allclose(moved[:4], 0.7 * dx, rtol=2e-4, atol=1e-9)
    if near:
        assert not moved[4].any()
    sim.destroy()


def test_hinge_angle_step(oracle_path):
    a = 0.01
    pos = np.array([[0, 0, 0], [a, 0, 0], [0, a, 0]], np.float32)
    sim = XPBD(oracle_path)
    sim.create(make_scene(pos, [[0, 1], [0, 2]]))
    compliance, blend = 1e-8, 0.8
    sim.set_solver(iters=1, dt=DT, under_relax=1.0, vel_damping=0.0, angle_compliance=compliance,
                   angle_blend=blend)
    sim.set_constraints(distance=False, shape=False, angle=True, volume=False)
    sim.set_ground(False, (0, 0, 1), 0.0, gravity=0.0)
    sim.step()
    th = np.deg2rad(80.0)
    cur = np.array([[0, 0, 0], [a, 0, 0], [a * np.cos(th), a * np.sin(th), 0.0]], np.float32)
    sim.ext_set_state(cur, np.zeros_like(cur))
    sim.step()
    x = cur.astype(np.float64)

    def cosang(y):
        v1, v2 = y[1] - y[0], y[2] - y[0]
        return np.dot(v1, v2) / np.linalg.norm(v1) / np.linalg.norm(v2)

    grad = np.zeros((3, 3))
    eps = 1e-9
    for j in range(3):
        for d in range(3):
            yp, ym = x.copy(), x.copy()
            yp[j, d] += eps
            ym[j, d] -= eps
            grad[j, d] = (cosang(yp) - cosang(ym)) / (2 * eps)
    C = cosang(x) - 0.0                      # rest angle 90 degrees
    dlam = np.clip(-C / ((grad * grad).sum() + compliance / DT ** 2), -0.1, 0.1)
    dx = dlam * grad
    assert np.linalg.norm(dx, axis=1).max() < 0.25 * a
    np.testing.assert_allclose(sim.positions().astype(np.float64) - x, blend * dx, rtol=1e-4, atol=1e-10)
    sim.destroy()
