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
from pathlib import Path

import numpy as np

from splat_io import read_ply

_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_LIB = os.path.join(_HERE, "libxpbd_isaac.so")


def _columns(table, *names):
    """float32 (N, len(names)) array of the named fields of a structured array."""
    return np.column_stack([np.asarray(table[f], np.float32) for f in names])


def load_inputs(d, name):
    """xpbd_create inputs from <d>/<name>_crop.ply and <d>/<name>_graph.npz,
    both in USD order (the loader of src/xpbd_bridge.py:17-41).

    PLY fields: x/y/z, scale_0..2 (log scales), rot_0..3 (w, x, y, z, not
    normalised) and opacity (logit).  NPZ: the bidirectional CSR (count, idx,
    dist, stiff) and the positions it was built on (pos_rest), which must
    equal the PLY's.  Scales are exponentiated, quaternions normalised,
    opacity passed through the logistic function (in double), and each
    undirected edge kept once, as its (i, j) entry with i < j, with that
    entry's rest length and stiffness."""
    folder = Path(d)
    _, verts = read_ply(folder / f"{name}_crop.ply")
    with np.load(folder / f"{name}_graph.npz", allow_pickle=False) as graph:
        count, nbr = graph["count"], graph["idx"]
        dist, stiff, anchor = graph["dist"], graph["stiff"], graph["pos_rest"]
    pos = _columns(verts, "x", "y", "z")
    if not np.array_equal(pos, anchor):
        raise ValueError("graph pos_rest and the cropped PLY positions differ; re-run the graph import")
    quat = _columns(verts, "rot_0", "rot_1", "rot_2", "rot_3")
    quat /= np.linalg.norm(quat, axis=1, keepdims=True)
    logit = np.asarray(verts["opacity"], np.float64)
    row_of = np.repeat(np.arange(pos.shape[0], dtype=np.int32), count)
    upper = np.flatnonzero(nbr > row_of)
    inp = {
        "pos": pos,
        "scales": np.exp(_columns(verts, "scale_0", "scale_1", "scale_2")).astype(np.float32),
        "rots": quat,
        "opacity": (1.0 / (1.0 + np.exp(-logit))).astype(np.float32),
        "edges": np.column_stack([row_of[upper], nbr[upper]]).astype(np.int32),
        "rest": dist[upper].astype(np.float32),
        "stiff": stiff[upper].astype(np.float32),
    }
    return {k: np.ascontiguousarray(v) for k, v in inp.items()}


_F32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_I32 = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_INT, _FLT, _PTR = ctypes.c_int, ctypes.c_float, ctypes.c_void_p
_FOUT = ctypes.POINTER(ctypes.c_float)

# symbol: (result type, argument types), as declared in include/xpbd_capi.h.
# The first 29 are the reference's ABI (xpbd_capi.cu:66-310); the others are
# this library's extensions, present in both the HIP library and the oracle.
ABI = {
    "xpbd_last_error": (ctypes.c_char_p, []),
    "xpbd_create": (_INT, [_INT, _F32, _F32, _F32, _F32, _INT, _I32, _F32, _F32]),
    "xpbd_destroy": (None, []),
    "xpbd_count": (_INT, []),
    "xpbd_set_solver": (None, [_INT] + [_FLT] * 10),
    "xpbd_set_constraints": (None, [_INT] * 4),
    "xpbd_set_volume": (None, [_FLT, _INT, _INT, _INT]),
    "xpbd_set_ground": (None, [_INT] + [_FLT] * 9),
    "xpbd_set_object_shape": (None, [_FLT]),
    "xpbd_set_shape_robust": (None, [_INT]),
    "xpbd_set_object_shape_gpu": (None, [_INT]),
    "xpbd_set_self_collision": (None, [_INT, _FLT, _FLT, _INT]),
    "xpbd_self_collision_stats": (None, [_I32]),
    "xpbd_set_colliders": (None, [_INT, _I32, _F32, _F32, _FLT, _FLT]),
    "xpbd_collider_stats": (_INT, [_I32, _F32, _F32, _INT]),
    "xpbd_add_rigid_velocity": (_INT, [_F32, _F32, _F32]),
    "xpbd_set_attached": (_INT, [_INT, _I32, _F32, _FLT]),
    "xpbd_set_contact_shape": (_INT, [_INT, _FLT, _FLT]),
    "xpbd_step": (_INT, []),
    "xpbd_positions_device": (ctypes.c_ulonglong, []),
    "xpbd_positions_host": (_INT, [_F32, _INT]),
    "xpbd_reset": (_INT, []),
    "xpbd_launch": (_INT, [_F32, _F32, _F32]),
    "xpbd_volume_stats": (None, [_INT, _FOUT, _FOUT]),
    "xpbd_compute_shapes": (_INT, [_FLT, _F32, _F32, _INT]),
    "xpbd_press_start": (_INT, [_INT, _FLT, _FLT, _INT]),
    "xpbd_press_stop": (None, []),
    "xpbd_press_autolog": (None, [_INT]),
    "xpbd_press_state": (_INT, [_FOUT] * 4),
    "xpbd_ext_get_state": (_INT, [_PTR, _PTR, _INT]),
    "xpbd_ext_set_state": (_INT, [_PTR, _PTR, _INT]),
    "xpbd_ext_get_inv_mass": (_INT, [_F32, _INT]),
    "xpbd_ext_backend": (ctypes.c_char_p, []),
    "xpbd_ext_profile": (_INT, [_INT, ctypes.c_char_p, _INT]),
    "xpbd_ext_set_volume_clusters": (_INT, [_PTR, _PTR, _INT]),
    "xpbd_ext_volume_topology": (_INT, [_PTR] * 5 + [_INT]),
    "xpbd_ext_set_node_order": (_INT, [_INT]),
    "xpbd_ext_self_collision_grid_info": (_INT, [_PTR]),
    "xpbd_ext_set_shape_clusters": (_INT, [_PTR, _INT]),
    "xpbd_ext_set_gauss_seidel": (_INT, [_INT, _INT]),
    "xpbd_ext_gs_info": (_INT, [_PTR]),
    "xpbd_ext_distance_gs_edges": (_INT, [_PTR] * 4),
    "xpbd_ext_volume_gs_order": (_INT, [_PTR] * 2),
    "xpbd_ext_set_self_collision_grid": (_INT, [_FLT, _INT]),
    "xpbd_ext_self_collision_lists": (_INT, [_PTR] * 4 + [_INT]),
}

# extensions of the HIP library only (looked up when used)
ABI_OPTIONAL = {
    "xpbd_ext_angle_pairs": (_INT, [_PTR, _PTR, _INT]),
    "xpbd_ext_set_self_collision_fine": (_INT, [_INT]),
    "xpbd_ext_set_group_lanes": (_INT, [_INT]),
    "xpbd_ext_self_collision_paths": (_INT, [_PTR]),
    "xpbd_ext_set_device_builders": (_INT, [_INT]),
    "xpbd_ext_autotune": (_INT, [_INT, _INT, ctypes.c_char_p, _INT]),
    "xpbd_ext_volume_members": (_INT, [_PTR, _INT]),
    "xpbd_ext_frame_outputs": (_INT, [_PTR, _PTR, _PTR, _INT, _FLT, _PTR, _FLT, ctypes.POINTER(ctypes.c_double), _INT]),
}


def bind(lib):
    """Declare result and argument types of every symbol in ABI (and of
    those in ABI_OPTIONAL the library exports) on a CDLL; returns it."""
    for table, required in ((ABI, True), (ABI_OPTIONAL, False)):
        for name, (res, args) in table.items():
            if not required and not hasattr(lib, name):
                continue
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f32(v, shape=None):
    a = np.ascontiguousarray(v, dtype=np.float32)
    return a if shape is None else a.reshape(shape)


def _or_dummy(a, rows, width):
    """`a` itself, or one zero row of `width` when there are no rows: the
    library takes a valid pointer even for an empty list."""
    return a if rows else np.zeros(width, a.dtype)


class XPBD:
    """The solver singleton behind the C ABI, with the reference bridge's
    methods (src/xpbd_bridge.py:45-271): same names, defaults, argument
    meaning; failing calls raise RuntimeError("<symbol>: <xpbd_last_error>")."""

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

    def _call(self, symbol, *args):
        self._check(getattr(self.lib, symbol)(*args), symbol)

    # ---- instance ---------------------------------------------------------
    def create(self, inp):
        pos = inp["pos"]
        edges = np.asarray(inp["edges"]).reshape(-1, 2)
        m = edges.shape[0]
        self._call("xpbd_create", pos.shape[0], pos.ravel(), inp["scales"].ravel(), inp["rots"].ravel(),
                   inp["opacity"], m, _or_dummy(inp["edges"].ravel(), m, 2), _or_dummy(inp["rest"], m, 1),
                   _or_dummy(inp["stiff"], m, 1))
        self.n = pos.shape[0]
        in_range = (edges >= 0).all(1) & (edges < self.n).all(1)
        self._edges_valid = int(np.count_nonzero(in_range & (edges[:, 0] != edges[:, 1])))

    def destroy(self):
        self.lib.xpbd_destroy()
        self.n = 0

    # ---- settings ---------------------------------------------------------
    def set_solver(self, iters=20, dt=1 / 60, under_relax=0.6, vel_damping=0.01, stiffness_scale=1.0,
                   inv_mass_scale=1.0, dist_compliance=0.0, shape_compliance=0.0, shape_blend=0.1,
                   angle_compliance=1e-2, angle_blend=0.2):
        self.lib.xpbd_set_solver(iters, dt, under_relax, vel_damping, stiffness_scale, inv_mass_scale,
                                 dist_compliance, shape_compliance, shape_blend, angle_compliance, angle_blend)

    def set_constraints(self, distance=True, shape=True, angle=False, volume=True):
        self.lib.xpbd_set_constraints(*(int(bool(f)) for f in (distance, shape, angle, volume)))

    def set_volume(self, compliance=1e-6, ring_k=3, max_members=2048, leader_min_hop=2):
        self.lib.xpbd_set_volume(compliance, ring_k, max_members, leader_min_hop)

    def set_ground(self, enabled, up, height, friction=0.5, restitution=0.0, contact_radius=0.0,
                   contact_slop=0.0, gravity=0.0):
        ux, uy, uz = (float(c) for c in up[:3])
        self.lib.xpbd_set_ground(int(enabled), ux, uy, uz, height, friction, restitution, contact_radius,
                                 contact_slop, gravity)

    def set_object_shape(self, stiffness):
        self.lib.xpbd_set_object_shape(stiffness)

    def set_object_shape_gpu(self, enabled):
        self.lib.xpbd_set_object_shape_gpu(int(enabled))

    def set_shape_robust(self, enabled):
        self.lib.xpbd_set_shape_robust(int(enabled))

    def set_self_collision(self, enabled, radius_scale=1.5, exclude_scale=2.0, within_body=False):
        self.lib.xpbd_set_self_collision(int(enabled), radius_scale, exclude_scale, int(within_body))

    def set_contact_shape(self, enabled, tau=0.2, radius_cap=0.0):
        self._call("xpbd_set_contact_shape", int(enabled), float(tau), float(radius_cap))
        return True

    # ---- colliders, grippers, impulses ----------------------------------
    def set_colliders(self, types, poses, dims, margin=0.0, friction=0.0):
        """Kinematic colliders, one-way (collider -> Gaussians), in input
        coordinates; may be called again before any step.  types (K,): 0 sphere,
        1 box, 2 capsule; poses (K, 12): local-to-input rotation (row-major 9)
        then centre; dims (K, 3): box half extents, sphere (radius, 0, 0),
        capsule (radius, half length along local z, 0).  An empty list switches
        colliders off."""
        kinds = np.ascontiguousarray(np.asarray(types, dtype=np.int32).ravel())
        k = kinds.size
        frames = _f32(poses, (k, 12)) if k else np.zeros((1, 12), np.float32)
        sizes = _f32(dims, (k, 3)) if k else np.zeros((1, 3), np.float32)
        self.lib.xpbd_set_colliders(k, _or_dummy(kinds, k, 1), frames.ravel(), sizes.ravel(), float(margin),
                                    float(friction))

    def collider_stats(self, count):
        """Per collider of the last step: (Gaussians touched (K,), summed push
        (K, 3), contact centre (K, 3)), input units."""
        cap = max(int(count), 1)
        hits = np.zeros(cap, np.int32)
        push = np.zeros((cap, 3), np.float32)
        centre = np.zeros((cap, 3), np.float32)
        k = self.lib.xpbd_collider_stats(hits, push.ravel(), centre.ravel(), count)
        return hits[:k], push[:k], centre[:k]

    def add_rigid_velocity(self, dv, dw=(0.0, 0.0, 0.0), center=(0.0, 0.0, 0.0)):
        """Adds dv + dw x (x - center) to every Gaussian's velocity (between steps)."""
        self._call("xpbd_add_rigid_velocity", _f32(dv), _f32(dw), _f32(center))

    def set_attached(self, idx, pos, weight=1.0):
        """Held Gaussians (gripper fingers): idx pinned at pos (N, 3, input
        coordinates), called every step with the new positions; weight of one
        held point in the object-shape rigid fit."""
        held = np.ascontiguousarray(idx, dtype=np.int32).ravel()
        where = _f32(pos, (-1, 3))
        m = held.size
        self._call("xpbd_set_attached", m, _or_dummy(held, m, 1), _or_dummy(where.ravel(), m, 3), float(weight))

    def clear_attached(self):
        self.set_attached(np.zeros(0, np.int32), np.zeros((0, 3), np.float32))

    # ---- stepping and results -------------------------------------------
    def step(self):
        self._call("xpbd_step")

    def positions(self):
        out = np.empty((self.n, 3), dtype=np.float32)
        self._check(self.lib.xpbd_positions_host(out.ravel(), self.n), "xpbd_positions_host")
        return out

    def positions_device_ptr(self):
        return int(self.lib.xpbd_positions_device())

    def shapes(self, deform_eps=1e-2):
        """Deformed Gaussian shapes: (activated scales (N, 3), quaternions
        (N, 4) w x y z, number deformed); Gaussians that moved at most
        deform_eps keep their own."""
        scales = np.empty((self.n, 3), dtype=np.float32)
        quats = np.empty((self.n, 4), dtype=np.float32)
        k = self.lib.xpbd_compute_shapes(float(deform_eps), scales.ravel(), quats.ravel(), self.n)
        if k < 0:
            self._check(k, "xpbd_compute_shapes")
        return scales, quats, k

    def shapes_device(self, deform_eps=1e-2):
        """compute_shapes with the outputs left in device memory (null host
        pointers, as the ABI allows); returns the deformed count."""
        fn = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int)(
            ("xpbd_compute_shapes", self.lib))
        k = fn(float(deform_eps), None, None, self.n)
        if k < 0:
            self._check(k, "xpbd_compute_shapes")
        return k

    def self_collision_stats(self):
        """(candidates, active Gaussians, over-capacity replacements) of the last step."""
        out = np.zeros(3, dtype=np.int32)
        self.lib.xpbd_self_collision_stats(out)
        return tuple(int(v) for v in out)

    def volume_stats(self, enable=True):
        mean, spread = ctypes.c_float(), ctypes.c_float()
        self.lib.xpbd_volume_stats(int(enable), ctypes.byref(mean), ctypes.byref(spread))
        return mean.value, spread.value

    # ---- scripted motion ------------------------------------------------
    def reset(self):
        self._call("xpbd_reset")

    def launch(self, lin=(0, 0, 0), ang=(0, 0, 0), tilt=(0, 0, 0)):
        self._call("xpbd_launch", _f32(lin), _f32(ang), _f32(tilt))

    def press_start(self, axis, ramp_per_sec, max_disp_pct, from_lo):
        """Plate press along axis 0/1/2; the low plate moves when from_lo;
        ramp in input units per second."""
        self._call("xpbd_press_start", int(axis), float(ramp_per_sec), float(max_disp_pct), int(from_lo))

    def press_autolog(self, enabled):
        self.lib.xpbd_press_autolog(int(enabled))

    def press_stop(self):
        self.lib.xpbd_press_stop()

    def press_state(self):
        """(active, lo plate, hi plate, current displacement, maximum displacement)."""
        vals = (ctypes.c_float * 4)()
        active = self.lib.xpbd_press_state(*(ctypes.cast(ctypes.byref(vals, 4 * i), _FOUT) for i in range(4)))
        return (bool(active),) + tuple(float(v) for v in vals)

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
        """Fine candidate scan: 1 one lane per node in the grid's order, 0 one
        wave per node, -1 (default) the lane scan; same lists."""
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

    def ext_self_collision_paths(self):
        """Node counts per candidate-search path of the last build: lane scan
        (lists of at most 32), lane scan with the slot rule, coarse walk,
        outside the grid."""
        out = np.zeros(4, np.int32)
        self._check(self.lib.xpbd_ext_self_collision_paths(_ptr(out)), "xpbd_ext_self_collision_paths")
        return tuple(int(v) for v in out)

    def ext_set_device_builders(self, mode):
        """Where k-ring volume member lists are built: 0 host, 1 GPU, -1 by size."""
        self._check(self.lib.xpbd_ext_set_device_builders(int(mode)), "xpbd_ext_set_device_builders")

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
