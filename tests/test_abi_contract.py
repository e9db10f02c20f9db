"""The C ABI's argument checks, call-order rules and setter clamps
(native/xpbd_capi.cu:66-310; clamps of forward.cu:3849-3965, 4430-4446,
4464-4476, 4478-4500, 4532-4540, 4896-4921), on both libraries.

Nothing here touches a GPU: the HIP library's checks run before any device
call, and the settings are read back through xpbd_ext_get_settings.  The
expected values are written from the setters' stated rules (lower bounds,
[0, 1] blends, normalised ground normal with a +z fallback, ...), not copied
from either implementation.
"""
import ctypes
import math

import numpy as np
import pytest

from conftest import ORACLE_LIB, PRODUCT_LIB
from xpbd_bridge import bind

LIBS = [("oracle", ORACLE_LIB), ("product", PRODUCT_LIB)]


@pytest.fixture(params=LIBS, ids=[n for n, _ in LIBS])
def lib(request):
    name, path = request.param
    import os
    if not os.path.isfile(path):
        pytest.fail(f"{name} library missing; run __graft_entry__.build()")
    L = bind(ctypes.CDLL(path))
    L.xpbd_ext_get_settings.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.xpbd_ext_get_settings.restype = ctypes.c_int
    yield L
    # persistent settings back to what the other tests expect
    L.xpbd_destroy()
    L.xpbd_set_solver(20, 1 / 60, 0.6, 0.59, 1.0, 1.0, 1e-3, 0.0, 0.1, 0.0, 0.2)
    L.xpbd_set_constraints(1, 0, 0, 0)
    L.xpbd_set_volume(1e-6, 1, 64, 1)
    L.xpbd_set_ground(0, 0.0, 0.0, 1.0, 0.0, 0.5, 0.3, 0.0, 0.0, 0.0)
    L.xpbd_set_object_shape(0.0)
    L.xpbd_set_object_shape_gpu(1)
    L.xpbd_set_self_collision(0, 1.5, 2.0, 0)
    L.xpbd_set_shape_robust(1)
    L.xpbd_press_autolog(1)


def settings(L):
    out = np.zeros(42, np.float32)
    assert L.xpbd_ext_get_settings(out.ctypes.data_as(ctypes.c_void_p), 42) == 42
    return out


def err(L):
    return L.xpbd_last_error().decode()


def tiny_inputs(n=4):
    pos = np.zeros((n, 3), np.float32)
    pos[:, 0] = np.arange(n)
    sc = np.full((n, 3), 0.01, np.float32)
    rot = np.tile(np.array([1, 0, 0, 0], np.float32), (n, 1))
    op = np.full(n, 0.9, np.float32)
    edges = np.array([[0, 1], [1, 2]], np.int32)
    rest = np.ones(2, np.float32)
    stiff = np.ones(2, np.float32)
    return pos, sc, rot, op, edges, rest, stiff


def test_create_rejects_invalid_arguments(lib):
    L = lib
    f32p = ctypes.POINTER(ctypes.c_float)
    i32p = ctypes.POINTER(ctypes.c_int)
    raw = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_int, f32p, f32p, f32p, f32p, ctypes.c_int, i32p, f32p, f32p)(
        ("xpbd_create", L))
    pos, sc, rot, op, edges, rest, stiff = tiny_inputs()
    P = lambda a: a.ctypes.data_as(f32p)   # noqa: E731
    I = lambda a: a.ctypes.data_as(i32p)   # noqa: E731
    cases = {
        "n = 0": (0, P(pos), P(sc), P(rot), P(op), 2, I(edges), P(rest), P(stiff)),
        "n < 0": (-3, P(pos), P(sc), P(rot), P(op), 2, I(edges), P(rest), P(stiff)),
        "null pos": (4, None, P(sc), P(rot), P(op), 2, I(edges), P(rest), P(stiff)),
        "null scales": (4, P(pos), None, P(rot), P(op), 2, I(edges), P(rest), P(stiff)),
        "null rots": (4, P(pos), P(sc), None, P(op), 2, I(edges), P(rest), P(stiff)),
        "null opacity": (4, P(pos), P(sc), P(rot), None, 2, I(edges), P(rest), P(stiff)),
        "e < 0": (4, P(pos), P(sc), P(rot), P(op), -1, I(edges), P(rest), P(stiff)),
        "null edges": (4, P(pos), P(sc), P(rot), P(op), 2, None, P(rest), P(stiff)),
        "null rest": (4, P(pos), P(sc), P(rot), P(op), 2, I(edges), None, P(stiff)),
        "null stiff": (4, P(pos), P(sc), P(rot), P(op), 2, I(edges), P(rest), None),
    }
    for what, args in cases.items():
        assert raw(*args) == -1, what
        assert err(L) == "xpbd_create: invalid arguments", what
        assert L.xpbd_count() == 0, what


def test_call_order_before_create(lib):
    L = lib
    L.xpbd_destroy()
    assert L.xpbd_step() == -1 and err(L) == "xpbd_step: call xpbd_create first"
    assert L.xpbd_set_contact_shape(1, 0.2, 0.0) == -1
    assert err(L) == "xpbd_set_contact_shape: call xpbd_create first"
    assert L.xpbd_positions_device() == 0


def _expect_no_step(L):
    """Every call that needs a first step fails, with the reference's message."""
    pos = np.zeros(12, np.float32)
    v3 = np.zeros(3, np.float32)
    idx = np.zeros(1, np.int32)
    assert L.xpbd_set_attached(1, idx, v3, 1.0) == -1 and err(L) == "xpbd_set_attached: no step yet"
    assert L.xpbd_add_rigid_velocity(v3, v3, v3) == -1 and err(L) == "xpbd_add_rigid_velocity: no step yet"
    assert L.xpbd_reset() == -1 and err(L) == "xpbd_reset: no step yet"
    assert L.xpbd_launch(v3, v3, v3) == -1 and err(L) == "xpbd_launch: no step yet"
    assert L.xpbd_press_start(1, 0.1, 0.3, 0) == -1 and err(L) == "xpbd_press_start: step once first"
    assert L.xpbd_positions_host(pos, 4) == -1 and err(L) == "xpbd_positions_host: no step yet"
    sc = np.zeros(12, np.float32)
    q = np.zeros(16, np.float32)
    assert L.xpbd_compute_shapes(0.01, sc, q, 4) == -1 and err(L) == "xpbd_compute_shapes: no step yet"
    assert L.xpbd_positions_device() == 0


def test_call_order_before_first_step(lib):
    L = lib
    L.xpbd_destroy()
    _expect_no_step(L)   # no instance at all
    if L.xpbd_ext_backend().decode() == "oracle-cpu":
        # created but not stepped (the HIP library needs a device to create)
        pos, sc, rot, op, edges, rest, stiff = tiny_inputs()
        assert L.xpbd_create(4, pos.ravel(), sc.ravel(), rot.ravel(), op, 2, edges.ravel(), rest, stiff) == 0
        assert L.xpbd_count() == 4
        _expect_no_step(L)
        assert L.xpbd_set_contact_shape(0, 0.2, 0.0) == 0


def test_oracle_size_checks_after_step(oracle_path):
    L = bind(ctypes.CDLL(oracle_path))
    pos, sc, rot, op, edges, rest, stiff = tiny_inputs()
    assert L.xpbd_create(4, pos.ravel(), sc.ravel(), rot.ravel(), op, 2, edges.ravel(), rest, stiff) == 0
    assert L.xpbd_step() == 0
    out = np.zeros(15, np.float32)
    assert L.xpbd_positions_host(out, 5) == -1 and err(L) == "xpbd_positions_host: size mismatch"
    assert L.xpbd_positions_host(out, 4) == 0
    q = np.zeros(20, np.float32)
    assert L.xpbd_compute_shapes(0.01, out, q, 3) == -1 and err(L) == "xpbd_compute_shapes: size mismatch"
    L.xpbd_destroy()
    assert L.xpbd_count() == 0


def test_solver_clamps(lib):
    L = lib
    L.xpbd_set_solver(0, -1.0, -0.5, -2.0, -1.0, -3.0, -1e-3, -5.0, 2.5, -1.0, -0.5)
    s = settings(L)
    assert s[0] == 1                          # iterations at least 1
    assert s[1] == np.float32(1e-5)           # dt at least 1e-5
    assert list(s[2:8]) == [0.0] * 6          # non-negative scalars
    assert s[8] == 1.0                        # shape blend in [0, 1]
    assert s[9] == 0.0 and s[10] == 0.0       # angle compliance >= 0, blend in [0, 1]
    L.xpbd_set_solver(7, 0.01, 0.5, 0.1, 2.0, 0.5, 1e-6, 3e-7, 0.4, 1e-2, 0.3)
    s = settings(L)
    np.testing.assert_array_equal(s[:11], np.float32([7, 0.01, 0.5, 0.1, 2.0, 0.5, 1e-6, 3e-7, 0.4, 1e-2, 0.3]))


def test_constraint_switches(lib):
    L = lib
    L.xpbd_set_constraints(0, 5, -1, 1)
    assert list(settings(L)[11:15]) == [0.0, 1.0, 1.0, 1.0]


def test_volume_clamps(lib):
    L = lib
    L.xpbd_set_volume(-1e-3, 20, 1, 0)
    s = settings(L)
    assert s[15] == np.float32(-1e-3)         # compliance is stored as given
    assert (s[16], s[17], s[18]) == (8, 8, 1)
    L.xpbd_set_volume(2e-6, 0, 100000, 12)
    s = settings(L)
    assert (s[16], s[17], s[18]) == (1, 8192, 8)


def test_ground_clamps(lib):
    L = lib
    L.xpbd_set_ground(1, 3.0, 0.0, 4.0, math.inf, -1.0, 7.0, -2.0, -3.0, -9.81)
    s = settings(L)
    assert s[19] == 1
    np.testing.assert_allclose(s[20:23], [0.6, 0.0, 0.8], rtol=1e-6)
    assert s[23] == 0.0                       # non-finite height -> 0
    assert (s[24], s[25], s[26], s[27], s[28]) == (0.0, 1.0, 0.0, 0.0, 0.0)
    L.xpbd_set_ground(0, 0.0, 0.0, 0.0, -0.25, 0.7, -1.0, 0.01, 0.002, 9.81)
    s = settings(L)
    assert list(s[20:23]) == [0.0, 0.0, 1.0]  # zero normal falls back to +z
    assert s[23] == np.float32(-0.25) and s[25] == 0.0
    L.xpbd_set_ground(1, math.nan, 0.0, 1.0, 0.0, 0.5, 0.5, 0.0, 0.0, 0.0)
    assert list(settings(L)[20:23]) == [0.0, 0.0, 1.0]   # NaN normal too


def test_object_shape_clamps(lib):
    L = lib
    for v, want in ((5.0, 1.0), (-1.0, 0.0), (math.nan, 0.0), (0.25, 0.25)):
        L.xpbd_set_object_shape(v)
        s = settings(L)
        assert s[29] == np.float32(want) and s[30] == 50
    L.xpbd_set_object_shape_gpu(0)
    assert settings(L)[31] == 0
    L.xpbd_set_object_shape_gpu(3)
    assert settings(L)[31] == 1


def test_self_collision_clamps(lib):
    L = lib
    L.xpbd_set_self_collision(1, math.nan, math.inf, 1)
    s = settings(L)
    assert (s[32], s[33], s[34], s[35]) == (1, 1.5, 2.0, 1)   # non-finite -> the defaults
    L.xpbd_set_self_collision(1, 0.0, 0.0, 0)
    s = settings(L)
    assert s[33] == np.float32(0.1) and s[34] == np.float32(1.05) and s[35] == 0
    L.xpbd_set_self_collision(0, 100.0, 100.0, 0)
    s = settings(L)
    assert (s[32], s[33], s[34]) == (0, 10.0, 20.0)
    out = np.full(3, 7, np.int32)
    L.xpbd_self_collision_stats(out)
    assert out[0] == 0 and out[2] == 0        # disabling clears the candidate counts


def test_collider_list_clamps(lib):
    L = lib
    k = 40
    types = np.full(k, 2, np.int32)
    poses = np.tile(np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], np.float32), k)
    dims = np.tile(np.array([0.02, -0.04, 0.0], np.float32), k)
    L.xpbd_set_colliders(k, types, poses, dims, -0.5, math.nan)
    s = settings(L)
    assert (s[37], s[38], s[39]) == (32, 0.0, 0.0)   # at most 32; margin and friction >= 0, finite
    L.xpbd_set_colliders(-2, types, poses, dims, 0.01, 0.3)
    s = settings(L)
    assert s[37] == 0 and s[38] == np.float32(0.01) and s[39] == np.float32(0.3)
    L.xpbd_set_colliders(0, types, poses, dims, 0.0, 0.0)


def test_shape_robust_and_autolog_switches(lib):
    L = lib
    L.xpbd_set_shape_robust(0)
    L.xpbd_press_autolog(0)
    s = settings(L)
    assert s[36] == 0 and s[40] == 0
    L.xpbd_set_shape_robust(1)
    L.xpbd_press_autolog(1)
    s = settings(L)
    assert s[36] == 1 and s[40] == 1


def test_settings_persist_across_create_and_destroy(oracle_path):
    L = bind(ctypes.CDLL(oracle_path))
    L.xpbd_ext_get_settings.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.xpbd_ext_get_settings.restype = ctypes.c_int
    L.xpbd_set_solver(3, 0.005, 0.6, 0.01, 1.0, 1.0, 1e-6, 0.0, 0.1, 0.0, 0.2)
    pos, sc, rot, op, edges, rest, stiff = tiny_inputs()
    assert L.xpbd_create(4, pos.ravel(), sc.ravel(), rot.ravel(), op, 2, edges.ravel(), rest, stiff) == 0
    L.xpbd_destroy()
    assert settings(L)[0] == 3 and settings(L)[1] == np.float32(0.005)
    L.xpbd_set_solver(20, 1 / 60, 0.6, 0.59, 1.0, 1.0, 1e-3, 0.0, 0.1, 0.0, 0.2)
