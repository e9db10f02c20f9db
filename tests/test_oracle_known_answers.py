"""Known answers that pin the oracle's per-node passes (CPU only).

Each case has a closed form chosen so that a plausible mistake fails it: the
order of the velocity and position updates in predict (semi-implicit Euler
gives k(k+1)/2 g dt^2, the other order k(k-1)/2), the sign and the 1/(w_i +
w_j) of the distance correction, the restitution and Coulomb friction of the
ground contact, and the sphere / box / capsule signed distances with a
non-symmetric rotation (a transposed R puts the particle on the wrong side).
The oracle restates forward.cu:5954-5979 (predict), 6484-6571 (distance
Jacobi), 7405-7421 (ground), 7425-7518 (colliders), 6623-6661 and 7528-7599
(velocity).
"""
import numpy as np
import pytest

from harness import XPBD

LIB = None


@pytest.fixture(autouse=True)
def _oracle(oracle_path):
    global LIB
    LIB = oracle_path


def particles(pos, edges=(), rest=()):
    pos = np.asarray(pos, np.float32).reshape(-1, 3)
    n = pos.shape[0]
    edges = np.asarray(edges, np.int32).reshape(-1, 2)
    return {
        "pos": np.ascontiguousarray(pos),
        "scales": np.full((n, 3), 0.01, np.float32),
        "rots": np.tile(np.array([1, 0, 0, 0], np.float32), (n, 1)),
        "opacity": np.full(n, 0.9, np.float32),
        "edges": np.ascontiguousarray(edges),
        "rest": np.ascontiguousarray(np.asarray(rest, np.float32).reshape(-1)),
        "stiff": np.ones(edges.shape[0], np.float32),
    }


def sim_with(scene, dt, iters=1, under_relax=1.0, damping=0.0, compliance=0.0):
    sim = XPBD(LIB)
    sim.create(scene)
    sim.set_solver(iters=iters, dt=dt, under_relax=under_relax, vel_damping=damping, dist_compliance=compliance)
    sim.set_constraints(distance=True, shape=False, angle=False, volume=False)
    return sim


def test_free_fall_is_semi_implicit_euler():
    # start at z = 0 so the step velocity dx / dt is not lost to the rounding
    # of a large coordinate
    dt, g, k = 0.01, 9.81, 12
    sim = sim_with(particles([[0.1, -0.2, 0.0]]), dt)
    sim.set_ground(False, (0, 0, 1), 0.0, gravity=g)
    for _ in range(k):
        sim.step()
    x, v = sim.ext_get_state()
    assert x[0, 0] == np.float32(0.1) and x[0, 1] == np.float32(-0.2)
    np.testing.assert_allclose(x[0, 2], -g * dt * dt * k * (k + 1) / 2, rtol=0, atol=5e-8)
    np.testing.assert_allclose(v[0], [0.0, 0.0, -g * dt * k], rtol=1e-5)
    sim.destroy()


def test_velocity_damping_scales_the_step_velocity():
    dt, g, d = 0.01, 9.81, 0.25
    sim = sim_with(particles([[0.0, 0.0, 0.0]]), dt, damping=d)
    sim.set_ground(False, (0, 0, 1), 0.0, gravity=g)
    sim.step()
    _, v = sim.ext_get_state()
    np.testing.assert_allclose(v[0, 2], -g * dt * (1.0 - d), rtol=1e-5)
    sim.destroy()


@pytest.mark.parametrize("compliance", [0.0, 1e-4])
def test_two_particle_distance_step(compliance):
    """One Jacobi iteration on a stretched pair: each end moves by
    (D - L) / (2 + alpha / dt^2) toward the other (w = 1 each)."""
    dt, L, D = 0.01, 0.1, 0.13
    sim = sim_with(particles([[0, 0, 1], [D, 0, 1]], [[0, 1]], [L]), dt, compliance=compliance)
    sim.set_ground(False, (0, 0, 1), 0.0, gravity=0.0)
    sim.step()
    x, v = sim.ext_get_state()
    at = compliance / (dt * dt)
    move = (D - L) / (2.0 + at)
    np.testing.assert_allclose(x[0], [move, 0, 1], atol=1e-7)
    np.testing.assert_allclose(x[1], [D - move, 0, 1], atol=1e-7)
    np.testing.assert_allclose(v[0, 0], move / dt, rtol=1e-5)
    np.testing.assert_allclose(v[1, 0], -move / dt, rtol=1e-5)
    sim.destroy()


def test_under_relaxation_scales_the_average_correction():
    dt, L, D, ur = 0.01, 0.1, 0.16, 0.6
    sim = sim_with(particles([[0, 0, 1], [D, 0, 1]], [[0, 1]], [L]), dt, under_relax=ur)
    sim.set_ground(False, (0, 0, 1), 0.0, gravity=0.0)
    sim.step()
    x, _ = sim.ext_get_state()
    np.testing.assert_allclose(x[1, 0] - x[0, 0], D - 2 * ur * (D - L) / 2, atol=1e-7)
    sim.destroy()


def test_ground_restitution_and_friction():
    """A particle hitting z = 0 at (1, 0, -2) m/s with no gravity (resting
    speed 0): projected onto the plane; normal speed 2e, tangential speed
    reduced by mu (1 + e) 2."""
    dt, e, mu = 0.01, 0.5, 0.1
    sim = sim_with(particles([[0.0, 0.0, 1.0]]), dt)
    sim.set_ground(True, (0, 0, 1), 0.0, friction=mu, restitution=e, gravity=0.0)
    sim.step()   # at rest far above the plane; the state can be set after a step
    sim.ext_set_state(np.array([[0.0, 0.0, 0.001]], np.float32), np.array([[1.0, 0.0, -2.0]], np.float32))
    sim.step()
    x, v = sim.ext_get_state()
    np.testing.assert_allclose(x[0], [dt * 1.0, 0.0, 0.0], atol=1e-7)
    np.testing.assert_allclose(v[0], [1.0 - mu * (1 + e) * 2.0, 0.0, 2.0 * e], rtol=1e-5, atol=1e-6)
    sim.destroy()


def test_ground_plane_with_tilted_normal_and_height():
    n = np.array([0.0, 0.6, 0.8], np.float32)
    h = 0.2
    p0 = np.array([0.3, 0.1, 0.05], np.float32)   # below the plane: n.p - h < 0
    sim = sim_with(particles([p0]), 0.01)
    sim.set_ground(True, n, h, friction=0.0, restitution=0.0, gravity=0.0)
    sim.step()
    x, _ = sim.ext_get_state()
    c = float(n @ p0) - h
    np.testing.assert_allclose(x[0], p0 - c * n, atol=1e-7)
    np.testing.assert_allclose(float(n @ x[0]), h, atol=1e-7)
    sim.destroy()


def pose(R, t):
    return np.concatenate([np.asarray(R, np.float32).reshape(9), np.asarray(t, np.float32)])


RX90 = np.array([[1, 0, 0], [0, 0, -1], [0, 1, 0]], np.float32)   # local z -> world +y


@pytest.mark.parametrize("case", ["sphere", "box", "capsule_side", "capsule_cap_rotated"])
def test_collider_pushes_to_the_surface(case):
    c = np.array([0.2, -0.1, 0.5], np.float32)
    margin = 0.003
    if case == "sphere":
        p0, typ, R, dims = c + [0.05, 0, 0], 0, np.eye(3), [0.1, 0, 0]
        want = c + [0.1 + margin, 0, 0]
    elif case == "box":
        # nearest face is +x (0.09 of 0.1); y and z are farther from their faces
        p0, typ, R, dims = c + [0.09, 0.05, -0.1], 1, np.eye(3), [0.1, 0.2, 0.3]
        want = c + [0.1 + margin, 0.05, -0.1]
    elif case == "capsule_side":
        p0, typ, R, dims = c + [0.02, 0, 0.05], 2, np.eye(3), [0.05, 0.1, 0]
        want = c + [0.05 + margin, 0, 0.05]
    else:
        # capsule axis along world +y after the rotation; a point past the +y
        # cap is pushed out along +y (a transposed R would use -y)
        p0, typ, R, dims = c + [0, 0.12, 0], 2, RX90, [0.05, 0.1, 0]
        want = c + [0, 0.15 + margin, 0]
    sim = sim_with(particles([p0]), 0.01)
    sim.set_ground(False, (0, 0, 1), 0.0, gravity=0.0)
    sim.set_colliders([typ], pose(R, c)[None, :], [dims], margin=margin, friction=0.0)
    sim.step()
    x, _ = sim.ext_get_state()
    np.testing.assert_allclose(x[0], want, atol=2e-7)
    hits, push, _ = sim.collider_stats(1)
    assert hits[0] == 1
    np.testing.assert_allclose(push[0], want - p0, atol=2e-7)
    sim.destroy()


def test_particle_outside_the_margin_is_untouched():
    c = np.array([0.0, 0.0, 0.5], np.float32)
    p0 = c + np.array([0.2, 0.0, 0.0], np.float32)
    sim = sim_with(particles([p0]), 0.01)
    sim.set_ground(False, (0, 0, 1), 0.0, gravity=0.0)
    sim.set_colliders([0], pose(np.eye(3), c)[None, :], [[0.1, 0, 0]], margin=0.05, friction=0.0)
    sim.step()
    x, _ = sim.ext_get_state()
    np.testing.assert_array_equal(x[0], p0)
    assert sim.collider_stats(1)[0][0] == 0
    sim.destroy()
