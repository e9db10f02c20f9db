"""ctypes bridge to libxpbd_isaac.so — a drop-in for the reference's
src/xpbd_bridge.py.

Same module name, same ``load_inputs(d, name)`` and ``XPBD`` class with the
same methods, defaults, argument meaning and error behaviour
(``RuntimeError("<call>: <xpbd_last_error>")``), bound to the same 29 C
symbols (src/xpbd_bridge.py:45-271).  The library path is the argument, then
``$XPBD_DLL``, then the ``libxpbd_isaac.so`` built next to this file.

Extensions beyond the reference class are prefixed ``ext_``.
"""
import ctypes
import os

import numpy as np

from splat_io import read_ply

_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_LIB = os.path.join(_HERE, "libxpbd_isaac.so")


def load_inputs(d, name):
    """Inputs for xpbd_create from <d>/<name>_crop.ply and <d>/<name>_graph.npz
    (src/xpbd_bridge.py:17-41): activated scales, normalised quaternions,
    activated opacity, and the CSR graph reduced to undirected i<j edges."""
    _, data = read_ply(os.path.join(d, f"{name}_crop.ply"))
    g = np.load(os.path.join(d, f"{name}_graph.npz"), allow_pickle=False)
    pos = np.stack([data["x"], data["y"], data["z"]], axis=1).astype(np.float32)
    if not np.array_equal(pos, g["pos_rest"]):
        raise ValueError("graph pos_rest and the cropped PLY positions differ; re-run the graph import")
    scales = np.exp(np.stack([data["scale_0"], data["scale_1"], data["scale_2"]], axis=1)).astype(np.float32)
    rots = np.stack([data["rot_0"], data["rot_1"], data["rot_2"], data["rot_3"]], axis=1).astype(np.float32)
    rots /= np.linalg.norm(rots, axis=1, keepdims=True)
    opacity = (1.0 / (1.0 + np.exp(-data["opacity"].astype(np.float64)))).astype(np.float32)
    owner = np.repeat(np.arange(len(pos), dtype=np.int32), g["count"])
    idx = g["idx"]
    keep = idx > owner
    edges = np.stack([owner[keep], idx[keep]], axis=1).astype(np.int32)
    return {
        "pos": np.ascontiguousarray(pos),
        "scales": np.ascontiguousarray(scales),
        "rots": np.ascontiguousarray(rots),
        "opacity": np.ascontiguousarray(opacity),
        "edges": np.ascontiguousarray(edges),
        "rest": np.ascontiguousarray(g["dist"][keep].astype(np.float32)),
        "stiff": np.ascontiguousarray(g["stiff"][keep].astype(np.float32)),
    }


def bind(lib):
    """Declare argtypes / restypes of every exported symbol on a CDLL."""
    L = lib
    f32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    i32 = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    c_int, c_float = ctypes.c_int, ctypes.c_float
    L.xpbd_last_error.restype = ctypes.c_char_p
    L.xpbd_create.argtypes = [c_int, f32, f32, f32, f32, c_int, i32, f32, f32]
    L.xpbd_create.restype = c_int
    L.xpbd_destroy.restype = None
    L.xpbd_count.restype = c_int
    L.xpbd_set_solver.argtypes = [c_int] + [c_float] * 10
    L.xpbd_set_solver.restype = None
    L.xpbd_set_constraints.argtypes = [c_int] * 4
    L.xpbd_set_constraints.restype = None
    L.xpbd_set_volume.argtypes = [c_float, c_int, c_int, c_int]
    L.xpbd_set_volume.restype = None
    L.xpbd_set_ground.argtypes = [c_int] + [c_float] * 9
    L.xpbd_set_ground.restype = None
    L.xpbd_set_object_shape.argtypes = [c_float]
    L.xpbd_set_object_shape.restype = None
    L.xpbd_set_shape_robust.argtypes = [c_int]
    L.xpbd_set_shape_robust.restype = None
    L.xpbd_set_object_shape_gpu.argtypes = [c_int]
    L.xpbd_set_object_shape_gpu.restype = None
    L.xpbd_set_self_collision.argtypes = [c_int, c_float, c_float, c_int]
    L.xpbd_set_self_collision.restype = None
    L.xpbd_self_collision_stats.argtypes = [i32]
    L.xpbd_self_collision_stats.restype = None
    L.xpbd_set_colliders.argtypes = [c_int, i32, f32, f32, c_float, c_float]
    L.xpbd_set_colliders.restype = None
    L.xpbd_collider_stats.argtypes = [i32, f32, f32, c_int]
    L.xpbd_collider_stats.restype = c_int
    L.xpbd_add_rigid_velocity.argtypes = [f32, f32, f32]
    L.xpbd_add_rigid_velocity.restype = c_int
    L.xpbd_set_attached.argtypes = [c_int, i32, f32, c_float]
    L.xpbd_set_attached.restype = c_int
    L.xpbd_set_contact_shape.argtypes = [c_int, c_float, c_float]
    L.xpbd_set_contact_shape.restype = c_int
    L.xpbd_step.restype = c_int
    L.xpbd_positions_device.restype = ctypes.c_ulonglong
    L.xpbd_positions_host.argtypes = [f32, c_int]
    L.xpbd_positions_host.restype = c_int
    L.xpbd_reset.restype = c_int
    L.xpbd_launch.argtypes = [f32, f32, f32]
    L.xpbd_launch.restype = c_int
    L.xpbd_volume_stats.argtypes = [c_int, ctypes.POINTER(c_float), ctypes.POINTER(c_float)]
    L.xpbd_volume_stats.restype = None
    L.xpbd_compute_shapes.argtypes = [c_float, f32, f32, c_int]
    L.xpbd_compute_shapes.restype = c_int
    L.xpbd_press_start.argtypes = [c_int, c_float, c_float, c_int]
    L.xpbd_press_start.restype = c_int
    L.xpbd_press_stop.restype = None
    L.xpbd_press_autolog.argtypes = [c_int]
    L.xpbd_press_autolog.restype = None
    L.xpbd_press_state.argtypes = [ctypes.POINTER(c_float)] * 4
    L.xpbd_press_state.restype = c_int
    L.xpbd_ext_get_state.argtypes = [ctypes.c_void_p, ctypes.c_void_p, c_int]
    L.xpbd_ext_get_state.restype = c_int
    L.xpbd_ext_set_state.argtypes = [ctypes.c_void_p, ctypes.c_void_p, c_int]
    L.xpbd_ext_set_state.restype = c_int
    L.xpbd_ext_get_inv_mass.argtypes = [f32, c_int]
    L.xpbd_ext_get_inv_mass.restype = c_int
    L.xpbd_ext_backend.restype = ctypes.c_char_p
    L.xpbd_ext_profile.argtypes = [c_int, ctypes.c_char_p, c_int]
    L.xpbd_ext_profile.restype = c_int
    L.xpbd_ext_set_volume_clusters.argtypes = [ctypes.c_void_p, ctypes.c_void_p, c_int]
    L.xpbd_ext_set_volume_clusters.restype = c_int
    L.xpbd_ext_volume_topology.argtypes = [ctypes.c_void_p] * 5 + [c_int]
    L.xpbd_ext_volume_topology.restype = c_int
    L.xpbd_ext_set_node_order.argtypes = [c_int]
    L.xpbd_ext_set_node_order.restype = c_int
    L.xpbd_ext_self_collision_grid_info.argtypes = [ctypes.c_void_p]
    L.xpbd_ext_self_collision_grid_info.restype = c_int
    L.xpbd_ext_set_shape_clusters.argtypes = [ctypes.c_void_p, c_int]
    L.xpbd_ext_set_shape_clusters.restype = c_int
    L.xpbd_ext_set_gauss_seidel.argtypes = [c_int, c_int]
    L.xpbd_ext_set_gauss_seidel.restype = c_int
    L.xpbd_ext_gs_info.argtypes = [ctypes.c_void_p]
    L.xpbd_ext_gs_info.restype = c_int
    L.xpbd_ext_distance_gs_edges.argtypes = [ctypes.c_void_p] * 4
    L.xpbd_ext_distance_gs_edges.restype = c_int
    L.xpbd_ext_volume_gs_order.argtypes = [ctypes.c_void_p] * 2
    L.xpbd_ext_volume_gs_order.restype = c_int
    L.xpbd_ext_set_self_collision_grid.argtypes = [c_float, c_int]
    L.xpbd_ext_set_self_collision_grid.restype = c_int
    L.xpbd_ext_self_collision_lists.argtypes = [ctypes.c_void_p] * 4 + [c_int]
    L.xpbd_ext_self_collision_lists.restype = c_int
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class XPBD:
    SPHERE, BOX, CAPSULE = 0, 1, 2

    def __init__(self, dll_path=None):
        path = dll_path or os.environ.get("XPBD_DLL") or DEFAULT_LIB
        if not os.path.isfile(path):
            raise FileNotFoundError(
                f"XPBD library not found: {path}. Build it with 3dgs-isaac-sim-xpbd_amd/build_lib.py."
            )
        self.path = path
        self.lib = bind(ctypes.CDLL(path))
        self.has_attach = hasattr(self.lib, "xpbd_set_attached")
        self.has_contact_shape = hasattr(self.lib, "xpbd_set_contact_shape")
        self.n = 0

    @property
    def backend(self):
        return self.lib.xpbd_ext_backend().decode()

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what}: {self.lib.xpbd_last_error().decode(errors='replace')}")

    def create(self, inp):
        e = inp["edges"].shape[0]
        edges = inp["edges"].ravel() if e else np.zeros(2, np.int32)
        rest = inp["rest"] if e else np.zeros(1, np.float32)
        stiff = inp["stiff"] if e else np.zeros(1, np.float32)
        self._check(self.lib.xpbd_create(inp["pos"].shape[0], inp["pos"].ravel(), inp["scales"].ravel(),
                                         inp["rots"].ravel(), inp["opacity"], e, edges, rest, stiff), "xpbd_create")
        self.n = inp["pos"].shape[0]
        ed = np.asarray(inp["edges"]).reshape(-1, 2)
        self._edges_valid = int(np.count_nonzero((ed[:, 0] != ed[:, 1]) & (ed >= 0).all(1) & (ed < self.n).all(1)))

    def set_solver(self, iters=20, dt=1 / 60, under_relax=0.6, vel_damping=0.01, stiffness_scale=1.0,
                   inv_mass_scale=1.0, dist_compliance=0.0, shape_compliance=0.0, shape_blend=0.1,
                   angle_compliance=1e-2, angle_blend=0.2):
        self.lib.xpbd_set_solver(iters, dt, under_relax, vel_damping, stiffness_scale, inv_mass_scale,
                                 dist_compliance, shape_compliance, shape_blend, angle_compliance, angle_blend)

    def set_constraints(self, distance=True, shape=True, angle=False, volume=True):
        self.lib.xpbd_set_constraints(int(distance), int(shape), int(angle), int(volume))

    def set_volume(self, compliance=1e-6, ring_k=3, max_members=2048, leader_min_hop=2):
        self.lib.xpbd_set_volume(compliance, ring_k, max_members, leader_min_hop)

    def set_ground(self, enabled, up, height, friction=0.5, restitution=0.0, contact_radius=0.0,
                   contact_slop=0.0, gravity=0.0):
        self.lib.xpbd_set_ground(int(enabled), float(up[0]), float(up[1]), float(up[2]), height, friction,
                                 restitution, contact_radius, contact_slop, gravity)

    def set_object_shape(self, stiffness):
        self.lib.xpbd_set_object_shape(stiffness)

    def set_object_shape_gpu(self, enabled):
        self.lib.xpbd_set_object_shape_gpu(int(enabled))

    def set_colliders(self, types, poses, dims, margin=0.0, friction=0.0):
        """Kinematic colliders: types (K,) 0 sphere / 1 box / 2 capsule; poses (K,12) R row-major + t;
        dims (K,3). An empty list switches colliders off."""
        types = np.ascontiguousarray(np.asarray(types, dtype=np.int32).reshape(-1))
        k = len(types)
        poses = np.ascontiguousarray(np.asarray(poses, dtype=np.float32).reshape(k, 12) if k else np.zeros((1, 12), np.float32))
        dims = np.ascontiguousarray(np.asarray(dims, dtype=np.float32).reshape(k, 3) if k else np.zeros((1, 3), np.float32))
        if not k:
            types = np.zeros(1, dtype=np.int32)
        self.lib.xpbd_set_colliders(k, types, poses.ravel(), dims.ravel(), float(margin), float(friction))

    def collider_stats(self, count):
        hits = np.zeros(max(count, 1), dtype=np.int32)
        push = np.zeros(max(count, 1) * 3, dtype=np.float32)
        cen = np.zeros(max(count, 1) * 3, dtype=np.float32)
        n = self.lib.xpbd_collider_stats(hits, push, cen, count)
        return hits[:n], push[:3 * n].reshape(n, 3), cen[:3 * n].reshape(n, 3)

    def add_rigid_velocity(self, dv, dw=(0.0, 0.0, 0.0), center=(0.0, 0.0, 0.0)):
        a = [np.ascontiguousarray(v, dtype=np.float32) for v in (dv, dw, center)]
        self._check(self.lib.xpbd_add_rigid_velocity(*a), "xpbd_add_rigid_velocity")

    def set_attached(self, idx, pos, weight=1.0):
        idx = np.ascontiguousarray(idx, dtype=np.int32).ravel()
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        if len(idx) == 0:
            idx, pos = np.zeros(1, np.int32), np.zeros((1, 3), np.float32)
            self._check(self.lib.xpbd_set_attached(0, idx, pos.ravel(), float(weight)), "xpbd_set_attached")
            return
        self._check(self.lib.xpbd_set_attached(len(idx), idx, pos.ravel(), float(weight)), "xpbd_set_attached")

    def clear_attached(self):
        self._check(self.lib.xpbd_set_attached(0, np.zeros(1, np.int32), np.zeros(3, np.float32), 1.0),
                    "xpbd_set_attached")

    def set_contact_shape(self, enabled, tau=0.2, radius_cap=0.0):
        self._check(self.lib.xpbd_set_contact_shape(int(enabled), float(tau), float(radius_cap)),
                    "xpbd_set_contact_shape")
        return True

    def set_self_collision(self, enabled, radius_scale=1.5, exclude_scale=2.0, within_body=False):
        self.lib.xpbd_set_self_collision(int(enabled), radius_scale, exclude_scale, int(within_body))

    def self_collision_stats(self):
        out = np.zeros(3, dtype=np.int32)
        self.lib.xpbd_self_collision_stats(out)
        return int(out[0]), int(out[1]), int(out[2])

    def set_shape_robust(self, enabled):
        self.lib.xpbd_set_shape_robust(int(enabled))

    def step(self):
        self._check(self.lib.xpbd_step(), "xpbd_step")

    def positions(self):
        out = np.empty((self.n, 3), dtype=np.float32)
        self._check(self.lib.xpbd_positions_host(out.ravel(), self.n), "xpbd_positions_host")
        return out

    def shapes(self, deform_eps=1e-2):
        sc = np.empty((self.n, 3), dtype=np.float32)
        q = np.empty((self.n, 4), dtype=np.float32)
        k = self.lib.xpbd_compute_shapes(float(deform_eps), sc.ravel(), q.ravel(), self.n)
        if k < 0:
            self._check(-1, "xpbd_compute_shapes")
        return sc, q, k

    def shapes_device(self, deform_eps=1e-2):
        """compute_shapes with the outputs left in device memory (null host
        pointers, as the ABI allows); returns the deformed count."""
        fn = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int)(
            ("xpbd_compute_shapes", self.lib))
        k = fn(float(deform_eps), None, None, self.n)
        if k < 0:
            self._check(-1, "xpbd_compute_shapes")
        return k

    def positions_device_ptr(self):
        return int(self.lib.xpbd_positions_device())

    def reset(self):
        self._check(self.lib.xpbd_reset(), "xpbd_reset")

    def launch(self, lin=(0, 0, 0), ang=(0, 0, 0), tilt=(0, 0, 0)):
        a = [np.ascontiguousarray(v, dtype=np.float32) for v in (lin, ang, tilt)]
        self._check(self.lib.xpbd_launch(*a), "xpbd_launch")

    def volume_stats(self, enable=True):
        m, s = ctypes.c_float(0.0), ctypes.c_float(0.0)
        self.lib.xpbd_volume_stats(int(enable), ctypes.byref(m), ctypes.byref(s))
        return m.value, s.value

    def press_start(self, axis, ramp_per_sec, max_disp_pct, from_lo):
        self._check(self.lib.xpbd_press_start(int(axis), float(ramp_per_sec), float(max_disp_pct), int(from_lo)),
                    "xpbd_press_start")

    def press_autolog(self, enabled):
        self.lib.xpbd_press_autolog(int(enabled))

    def press_stop(self):
        self.lib.xpbd_press_stop()

    def press_state(self):
        v = [ctypes.c_float(0.0) for _ in range(4)]
        active = self.lib.xpbd_press_state(*[ctypes.byref(x) for x in v])
        return bool(active), v[0].value, v[1].value, v[2].value, v[3].value

    def destroy(self):
        self.lib.xpbd_destroy()
        self.n = 0

    # ---- extensions -----------------------------------------------------
    def ext_get_state(self):
        pos = np.empty((self.n, 3), np.float32)
        vel = np.empty((self.n, 3), np.float32)
        self._check(self.lib.xpbd_ext_get_state(_ptr(pos), _ptr(vel), self.n), "xpbd_ext_get_state")
        return pos, vel

    def ext_set_state(self, pos=None, vel=None):
        pos = None if pos is None else np.ascontiguousarray(pos, np.float32)
        vel = None if vel is None else np.ascontiguousarray(vel, np.float32)
        self._check(self.lib.xpbd_ext_set_state(_ptr(pos), _ptr(vel), self.n), "xpbd_ext_set_state")

    def ext_profile(self, substeps):
        """Advance `substeps` substeps eagerly with per-kernel HIP event timing.
        Returns {kernel: (total_ms, launches)}."""
        import json
        buf = ctypes.create_string_buffer(1 << 16)
        self._check(self.lib.xpbd_ext_profile(int(substeps), buf, len(buf)), "xpbd_ext_profile")
        return {k: (float(v[0]), int(v[1])) for k, v in json.loads(buf.value.decode()).items()}

    def ext_angle_pairs(self):
        """Umbrella angle pairing per directed entry (reference CSR order):
        (pairNext int32[m], restCos float32[m])."""
        fn = self.lib.xpbd_ext_angle_pairs
        fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        fn.restype = ctypes.c_int
        m = 2 * int(self._edges_valid)
        pair = np.empty(max(m, 1), np.int32)
        cos = np.empty(max(m, 1), np.float32)
        self._check(fn(_ptr(pair), _ptr(cos), m), "xpbd_ext_angle_pairs")
        return pair[:m], cos[:m]

    def ext_set_self_collision_fine(self, enabled):
        """Candidate search of one-object bodies through the predicted-position
        grid (True, default) or the reference's coarse walk (False); same lists."""
        fn = self.lib.xpbd_ext_set_self_collision_fine
        fn.argtypes = [ctypes.c_int]
        fn.restype = ctypes.c_int
        self._check(fn(int(bool(enabled))), "xpbd_ext_set_self_collision_fine")

    def ext_set_group_lanes(self, lanes):
        """Lanes per node of the grouped fine candidate scan (8, 16, 32), 0 for
        the one-wave-per-node scan, -1 (default) by mean body size; same lists."""
        fn = self.lib.xpbd_ext_set_group_lanes
        fn.argtypes = [ctypes.c_int]
        fn.restype = ctypes.c_int
        self._check(fn(int(lanes)), "xpbd_ext_set_group_lanes")

    def ext_autotune(self, substeps=2, hash_table=False):
        """Pick the fastest launch shapes on the current state (restored
        afterwards; results unchanged).  Returns the tuner's report."""
        import json
        fn = self.lib.xpbd_ext_autotune
        fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        fn.restype = ctypes.c_int
        buf = ctypes.create_string_buffer(4096)
        self._check(fn(int(substeps), 1 if hash_table else 0, buf, len(buf)), "xpbd_ext_autotune")
        return json.loads(buf.value.decode())

    def ext_set_volume_clusters(self, off=None, members=None, min_members=4):
        """Explicit volume clusters: node i leads {i} + members[off[i]:off[i+1]].
        None returns to the k-ring rule."""
        if off is None:
            self._check(self.lib.xpbd_ext_set_volume_clusters(None, None, 4), "xpbd_ext_set_volume_clusters")
            return
        off = np.ascontiguousarray(off, np.int32)
        members = np.ascontiguousarray(members, np.int32)
        if members.size == 0:
            members = np.zeros(1, np.int32)
        self._check(self.lib.xpbd_ext_set_volume_clusters(_ptr(off), _ptr(members), int(min_members)),
                    "xpbd_ext_set_volume_clusters")

    def ext_volume_topology(self):
        """(leader flag, member count, clusters containing node, material type, rest det) per node."""
        n = self.n
        lead, mc, rc, mt = (np.zeros(n, np.int32) for _ in range(4))
        dr = np.zeros(n, np.float32)
        self._check(self.lib.xpbd_ext_volume_topology(_ptr(lead), _ptr(mc), _ptr(rc), _ptr(mt), _ptr(dr), n),
                    "xpbd_ext_volume_topology")
        return lead, mc, rc, mt, dr

    def ext_volume_members(self):
        """Members of every volume cluster, one int32 array per node (empty for
        non-leaders), in the builder's order."""
        fn = self.lib.xpbd_ext_volume_members
        fn.argtypes = [ctypes.c_void_p, ctypes.c_int]
        fn.restype = ctypes.c_int
        counts = self.ext_volume_topology()[1]
        flat = np.zeros(max(int(counts.sum()), 1), np.int32)
        total = fn(_ptr(flat), flat.size)
        if total < 0:
            self._check(-1, "xpbd_ext_volume_members")
        return np.split(flat[:total], np.cumsum(counts)[:-1])

    def ext_inv_mass(self):
        out = np.empty(self.n, np.float32)
        self._check(self.lib.xpbd_ext_get_inv_mass(out, self.n), "xpbd_ext_get_inv_mass")
        return out

    def ext_set_node_order(self, mode):
        """Device node layout of the next instance: 1 Morton (default), 0 caller order."""
        self._check(self.lib.xpbd_ext_set_node_order(int(mode)), "xpbd_ext_set_node_order")

    def ext_self_collision_grid_info(self):
        """(cell edge, d_c, search radius) of the last self-collision build."""
        info = np.zeros(3, np.float32)
        self._check(self.lib.xpbd_ext_self_collision_grid_info(_ptr(info)), "xpbd_ext_self_collision_grid_info")
        return tuple(float(v) for v in info)

    def ext_set_shape_clusters(self, cluster_of=None):
        """Shape matching over node clusters (cluster id per node, -1 = none); None = 1-ring."""
        if cluster_of is None:
            self._check(self.lib.xpbd_ext_set_shape_clusters(None, self.n), "xpbd_ext_set_shape_clusters")
            return
        c = np.ascontiguousarray(cluster_of, np.int32)
        self._check(self.lib.xpbd_ext_set_shape_clusters(_ptr(c), int(c.size)), "xpbd_ext_set_shape_clusters")

    def ext_set_gauss_seidel(self, distance, volume):
        """Graph-coloured Gauss-Seidel instead of the Jacobi distance / gather volume passes."""
        self._check(self.lib.xpbd_ext_set_gauss_seidel(int(distance), int(volume)), "xpbd_ext_set_gauss_seidel")

    def ext_gs_info(self):
        """dict of the colouring sizes (after the first step)."""
        info = np.zeros(6, np.int32)
        self._check(self.lib.xpbd_ext_gs_info(_ptr(info)), "xpbd_ext_gs_info")
        keys = ("edges", "edge_colors", "one_way", "clusters", "cluster_colors", "cluster_skipped")
        return dict(zip(keys, (int(v) for v in info)))

    def ext_distance_gs_edges(self):
        """(ab[E, 2], rest[E], stiff[E], colorOff[C+1]) of the distance colouring."""
        info = self.ext_gs_info()
        m, c = info["edges"], info["edge_colors"]
        ab = np.zeros((max(m, 1), 2), np.int32)
        rest = np.zeros(max(m, 1), np.float32)
        stiff = np.zeros(max(m, 1), np.float32)
        off = np.zeros(c + 1, np.int32)
        self._check(self.lib.xpbd_ext_distance_gs_edges(_ptr(ab), _ptr(rest), _ptr(stiff), _ptr(off)),
                    "xpbd_ext_distance_gs_edges")
        return ab[:m], rest[:m], stiff[:m], off

    def ext_volume_gs_order(self):
        """(leaders in colour-major order, colorOff[C+1]) of the volume colouring."""
        info = self.ext_gs_info()
        order = np.zeros(max(info["clusters"], 1), np.int32)
        off = np.zeros(info["cluster_colors"] + 1, np.int32)
        self._check(self.lib.xpbd_ext_volume_gs_order(_ptr(order), _ptr(off)), "xpbd_ext_volume_gs_order")
        return order[:info["clusters"]], off

    def ext_set_self_collision_grid(self, cell_size=0.0, table_size=0):
        """Fixed self-collision hash cell (0 = displacement rule) and table size (0 = default)."""
        self._check(self.lib.xpbd_ext_set_self_collision_grid(float(cell_size), int(table_size)),
                    "xpbd_ext_set_self_collision_grid")

    def ext_self_collision_lists(self):
        """(keys[n], stable key order[n], counts[n], contacts[n, 32]) of the last build."""
        n = self.n
        keys = np.zeros(n, np.uint32)
        order = np.zeros(n, np.int32)
        counts = np.zeros(n, np.int32)
        contacts = np.zeros((n, 32), np.int32)
        self._check(self.lib.xpbd_ext_self_collision_lists(_ptr(keys), _ptr(order), _ptr(counts), _ptr(contacts), n),
                    "xpbd_ext_self_collision_lists")
        return keys, order, counts, contacts
