"""Rigid velocity, reset, launch and the press cycle
(native/xpbd_capi.cu:167-172, 263-303 -> forward.cu:6701-6712, 6790-6796,
4310-4427, 4551-4610, 9411-9424, 12095-12112).

CPU tests pin the oracle against closed forms: the rigid velocity field
v += dv + dw x (x - c) on unpinned nodes, the rest state after a reset, a
rigid tilt about the rest centroid (scipy's rotation-vector map) with the
velocity field of the launch, and the press schedule (plateaus at exactly
25/50/75/100 % of the largest displacement, each held for the 90-step dwell,
plates bounding every free node along the press axis).  GPU tests run the
same scripts on the HIP library and the oracle and compare every step.
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from harness import XPBD, max_abs, ulp_diff
import scenes

TOL_REL = 1e-5


def small_scene(n=2000, seed=5):
    return scenes.make_ellipsoid_scene(n, 8, seed=seed)


def setup(sim, scene, gravity=9.81):
    sim.create(scene)
    sim.set_solver(iters=4, dt=1.0 / 600.0, under_relax=0.6, vel_damping=0.01, dist_compliance=1e-6)
    sim.set_constraints(distance=True, shape=False, angle=False, volume=False)
    sim.set_ground(True, (0.0, 0.0, 1.0), 0.0, friction=0.5, restitution=0.0, gravity=gravity)
    sim.set_colliders([], [], [])


def _rigid_args():
    return (np.float32([0.3, -0.1, 0.05]), np.float32([0.5, 1.5, -2.0]), np.float32([0.01, 0.02, 0.08]))


def _launch_args():
    return (np.float32([0.2, 0.0, 1.0]), np.float32([0.0, 3.0, 0.5]), np.float32([0.1, -0.3, 0.2]))


# ---------------------------------------------------------------------------
# Oracle pins (CPU)
# ---------------------------------------------------------------------------
def test_oracle_rigid_velocity_closed_form(oracle_path):
    scene = small_scene(600)
    sim = XPBD(oracle_path)
    setup(sim, scene)
    for _ in range(3):
        sim.step()
    held = np.arange(0, 600, 7)
    pos0, _ = sim.ext_get_state()
    sim.set_attached(held, pos0[held])        # pinned nodes (inverse mass 0) keep their velocity
    x, v = sim.ext_get_state()
    dv, dw, c = _rigid_args()
    sim.add_rigid_velocity(dv, dw, c)
    _, v2 = sim.ext_get_state()
    r = x.astype(np.float64) - c
    want = v.astype(np.float64) + dv + np.cross(dw.astype(np.float64), r)
    free = np.ones(600, bool)
    free[held] = False
    np.testing.assert_allclose(v2[free], want[free], rtol=0, atol=1e-6)
    np.testing.assert_array_equal(v2[~free], v[~free])
    sim.destroy()


def test_oracle_reset_and_launch_closed_form(oracle_path):
    scene = small_scene(600)
    sim = XPBD(oracle_path)
    setup(sim, scene)
    for _ in range(5):
        sim.step()
    sim.set_attached([3, 4], scene["pos"][[3, 4]] + 0.01)
    sim.reset()
    x, v = sim.ext_get_state()
    np.testing.assert_array_equal(x, scene["pos"])
    assert not v.any()
    np.testing.assert_array_equal(sim.ext_inv_mass(), np.ones(600, np.float32))
    lin, ang, tilt = _launch_args()
    sim.launch(lin, ang, tilt)
    x, v = sim.ext_get_state()
    rest = scene["pos"].astype(np.float64)
    com = rest.mean(0)
    R = Rotation.from_rotvec(tilt.astype(np.float64)).as_matrix()
    r = (rest - com) @ R.T
    np.testing.assert_allclose(x, com + r, rtol=0, atol=2e-6)
    np.testing.assert_allclose(v, lin + np.cross(ang.astype(np.float64), r), rtol=0, atol=2e-5)
    sim.destroy()


def _press_run(sim, scene, steps, autolog=True, axis=2, ramp=0.15, pct=0.3, from_lo=False):
    setup(sim, scene)
    sim.step()
    sim.press_autolog(autolog)
    sim.press_start(axis, ramp, pct, from_lo)
    states, pos = [], []
    for _ in range(steps):
        sim.step()
        states.append(sim.press_state())
        pos.append(sim.positions())
    return states, pos


def test_oracle_press_schedule_and_plates(oracle_path):
    scene = small_scene(800)
    sim = XPBD(oracle_path)
    dt, ramp, pct = 1.0 / 600.0, 0.6, 0.3
    states, pos = _press_run(sim, scene, 520, ramp=ramp, pct=pct)
    rest = scene["pos"][:, 2]
    extent = float(rest.max() - rest.min())
    maxd = pct * extent
    cur = np.array([s[3] for s in states], np.float64)
    assert all(s[0] for s in states)
    np.testing.assert_allclose([s[4] for s in states], maxd, rtol=1e-6)
    assert np.all(np.diff(cur) >= 0.0)
    # plateaus: each checkpoint is hit exactly and held for the 90-step dwell
    for frac in (0.25, 0.5, 0.75, 1.0):
        target = np.float32(np.float32(frac) * np.float32(maxd))
        held = np.flatnonzero(cur == target)
        assert held.size >= 90, (frac, held.size)
        assert np.all(np.diff(held[:91]) == 1)
    # between plateaus the plate moves ramp * dt per step
    step = np.diff(cur)
    moving = step[(step > 0) & (step < 0.9 * ramp * dt)]   # the clipped steps at the checkpoints
    assert moving.size <= 4
    # the moving (hi) plate and the fixed (lo) plate bound every free node
    for (act, lo, hi, c, m), p in zip(states, pos):
        assert hi == np.float32(rest.max() - c) or abs(hi - (rest.max() - c)) <= 1e-6
        assert p[:, 2].max() <= hi + 1e-7 and p[:, 2].min() >= lo - 1e-7
    sim.press_stop()
    assert not sim.press_state()[0]
    np.testing.assert_array_equal(sim.ext_inv_mass(), np.ones(800, np.float32))
    sim.destroy()


def test_oracle_press_without_autolog_ramps_linearly(oracle_path):
    scene = small_scene(500)
    sim = XPBD(oracle_path)
    dt, ramp, pct = 1.0 / 600.0, 0.9, 0.2
    states, _ = _press_run(sim, scene, 100, autolog=False, ramp=ramp, pct=pct, axis=0, from_lo=True)
    rest = scene["pos"][:, 0]
    maxd = pct * float(rest.max() - rest.min())
    cur = np.array([s[3] for s in states], np.float64)
    k = np.arange(1, 101)
    np.testing.assert_allclose(cur, np.minimum(k * ramp * dt, maxd), rtol=1e-4, atol=1e-7)
    lo = np.array([s[1] for s in states])
    np.testing.assert_allclose(lo, rest.min() + cur, rtol=0, atol=1e-6)   # the lo plate moves
    sim.press_stop()
    sim.press_autolog(True)
    sim.destroy()


# ---------------------------------------------------------------------------
# GPU parity
# ---------------------------------------------------------------------------
def _compare(gpu, cpu, extent, what):
    a, b = gpu.positions(), cpu.positions()
    err = max_abs(a, b) / extent
    assert err <= TOL_REL, f"{what}: |dx|/extent {err:.3e} (ulp {ulp_diff(a, b)})"
    return err


@pytest.mark.gpu
def test_rigid_velocity_parity(product_path, oracle_path):
    scene = small_scene()
    extent = scenes.scene_extent(scene)
    gpu, cpu = XPBD(product_path), XPBD(oracle_path)
    worst = 0.0
    for sim in (gpu, cpu):
        setup(sim, scene)
        for _ in range(3):
            sim.step()
        sim.set_attached(np.arange(0, 2000, 11), scene["pos"][::11][:len(range(0, 2000, 11))])
        sim.add_rigid_velocity(*_rigid_args())
    _, vg = gpu.ext_get_state()
    _, vc = cpu.ext_get_state()
    assert ulp_diff(vg, vc) == 0
    for _ in range(20):
        gpu.step()
        cpu.step()
        worst = max(worst, _compare(gpu, cpu, extent, "rigid velocity"))
    print(f"rigid velocity: 20 substeps, max |dx|/extent {worst:.3e}")
    gpu.destroy()
    cpu.destroy()


@pytest.mark.gpu
def test_reset_and_launch_parity(product_path, oracle_path):
    scene = small_scene()
    extent = scenes.scene_extent(scene)
    gpu, cpu = XPBD(product_path), XPBD(oracle_path)
    for sim in (gpu, cpu):
        setup(sim, scene)
        for _ in range(10):
            sim.step()
        sim.set_attached([1, 2, 3], scene["pos"][[1, 2, 3]] + 0.02)
        sim.step()
        sim.reset()
    np.testing.assert_array_equal(gpu.positions(), scene["pos"])
    np.testing.assert_array_equal(gpu.ext_inv_mass(), np.ones(2000, np.float32))
    for _ in range(5):
        gpu.step()
        cpu.step()
        _compare(gpu, cpu, extent, "after reset")
    lin, ang, tilt = _launch_args()
    gpu.launch(lin, ang, tilt)
    cpu.launch(lin, ang, tilt)
    xg, vg = gpu.ext_get_state()
    xc, vc = cpu.ext_get_state()
    assert ulp_diff(xg, xc) == 0 and ulp_diff(vg, vc) == 0
    worst = 0.0
    for _ in range(60):
        gpu.step()
        cpu.step()
        worst = max(worst, _compare(gpu, cpu, extent, "after launch"))
    print(f"launch: 60 substeps, max |dx|/extent {worst:.3e}")
    gpu.destroy()
    cpu.destroy()


@pytest.mark.gpu
@pytest.mark.parametrize("autolog", [True, False])
def test_press_cycle_parity(product_path, oracle_path, autolog):
    scene = small_scene()
    extent = scenes.scene_extent(scene)
    gpu, cpu = XPBD(product_path), XPBD(oracle_path)
    for sim in (gpu, cpu):
        setup(sim, scene)
        sim.step()
        sim.press_autolog(autolog)
        sim.press_start(2, 0.6, 0.3, False)
    worst = 0.0
    steps = 480 if autolog else 150
    for k in range(steps):
        gpu.step()
        cpu.step()
        sg, sc = gpu.press_state(), cpu.press_state()
        assert sg == sc, (k, sg, sc)
        worst = max(worst, _compare(gpu, cpu, extent, f"press step {k}"))
    for sim in (gpu, cpu):
        sim.press_stop()
    np.testing.assert_array_equal(gpu.ext_inv_mass(), cpu.ext_inv_mass())
    assert gpu.press_state() == cpu.press_state()
    for _ in range(30):
        gpu.step()
        cpu.step()
        worst = max(worst, _compare(gpu, cpu, extent, "after press stop"))
    print(f"press (autolog {autolog}): {steps} + 30 substeps, max |dx|/extent {worst:.3e}")
    for sim in (gpu, cpu):
        sim.press_autolog(True)
        sim.destroy()
