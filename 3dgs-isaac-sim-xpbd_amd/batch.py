"""ctypes wrapper of the batched many-object API (include/xpbd_capi.h, part 3).

``XPBDBatch(scenes)`` creates one batch from a list of scene dicts shaped like
the reference loader's output (src/xpbd_bridge.py:17-41).  Its setters mirror
the ``XPBD`` methods of ``xpbd_bridge`` with an object index first (``obj=-1``
means every object); errors raise ``RuntimeError("<call>: <message>")`` as
the reference bridge does (src/xpbd_bridge.py:115-117).
"""
import ctypes
import json
import os

import numpy as np

from xpbd_bridge import DEFAULT_LIB

_VP = ctypes.c_void_p


def _p(a):
    return None if a is None else a.ctypes.data_as(_VP)


def bind_batch(lib):
    c_int, c_float = ctypes.c_int, ctypes.c_float
    sig = {
        "xpbd_batch_create": ([c_int] + [_VP] * 9, _VP),
        "xpbd_batch_destroy": ([_VP], None),
        "xpbd_batch_last_error": ([_VP], ctypes.c_char_p),
        "xpbd_batch_objects": ([_VP], c_int),
        "xpbd_batch_count": ([_VP, c_int], c_int),
        "xpbd_batch_set_solver": ([_VP, c_int, c_int] + [c_float] * 10, c_int),
        "xpbd_batch_set_constraints": ([_VP, c_int] + [c_int] * 4, c_int),
        "xpbd_batch_set_volume": ([_VP, c_int, c_float, c_int, c_int, c_int], c_int),
        "xpbd_batch_set_ground": ([_VP, c_int, c_int] + [c_float] * 9, c_int),
        "xpbd_batch_set_object_shape": ([_VP, c_int, c_float], c_int),
        "xpbd_batch_set_self_collision": ([_VP, c_int, c_int, c_float, c_float, c_int], c_int),
        "xpbd_batch_set_self_collision_grid": ([_VP, c_int, c_float, c_int], c_int),
        "xpbd_batch_set_gauss_seidel": ([_VP, c_int, c_int, c_int], c_int),
        "xpbd_batch_set_shape_robust": ([_VP, c_int, c_int], c_int),
        "xpbd_batch_set_volume_clusters": ([_VP, c_int, _VP, _VP, c_int], c_int),
        "xpbd_batch_set_shape_clusters": ([_VP, c_int, _VP, c_int], c_int),
        "xpbd_batch_set_colliders": ([_VP, c_int, c_int, _VP, _VP, _VP, c_float, c_float], c_int),
        "xpbd_batch_set_attached": ([_VP, c_int, _VP, _VP, _VP, _VP, c_float], c_int),
        "xpbd_batch_add_rigid_velocity": ([_VP, c_int, _VP, _VP, _VP], c_int),
        "xpbd_batch_reset": ([_VP, c_int], c_int),
        "xpbd_batch_prepare": ([_VP], c_int),
        "xpbd_batch_step": ([_VP, c_int], c_int),
        "xpbd_batch_step_async": ([_VP, c_int], c_int),
        "xpbd_batch_sync": ([_VP], c_int),
        "xpbd_batch_positions_device": ([_VP], ctypes.c_ulonglong),
        "xpbd_batch_positions_host": ([_VP, c_int, _VP, c_int], c_int),
        "xpbd_batch_get_state": ([_VP, c_int, _VP, _VP, c_int], c_int),
        "xpbd_batch_set_state": ([_VP, c_int, _VP, _VP, c_int], c_int),
        "xpbd_batch_compute_shapes": ([_VP, c_int, c_float, _VP, _VP, c_int], c_int),
        "xpbd_batch_shapes_device": ([_VP, _VP], c_int),
        "xpbd_batch_self_collision_stats": ([_VP, c_int, _VP], c_int),
        "xpbd_batch_collider_stats": ([_VP, c_int, _VP, _VP, _VP, c_int], c_int),
        "xpbd_batch_copy_object": ([_VP, c_int, _VP, c_int], c_int),
        "xpbd_batch_autotune": ([_VP, c_int, c_int, ctypes.c_char_p, c_int], c_int),
        "xpbd_batch_get_tuning": ([_VP, _VP], c_int),
        "xpbd_batch_set_tuning": ([_VP, _VP], c_int),
        "xpbd_batch_profile": ([_VP, c_int, ctypes.c_char_p, c_int], c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    return lib


_LIBS = {}


def load_batch_lib(path=None):
    path = path or os.environ.get("XPBD_DLL") or DEFAULT_LIB
    if path not in _LIBS:
        if not os.path.isfile(path):
            raise FileNotFoundError(f"XPBD library not found: {path}. Build it with 3dgs-isaac-sim-xpbd_amd/build_lib.py.")
        _LIBS[path] = bind_batch(ctypes.CDLL(path))
    return _LIBS[path]


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a if shape is None else a.reshape(shape)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class XPBDBatch:
    """Many independent objects stepped together (one launch set per substep)."""
    SPHERE, BOX, CAPSULE = 0, 1, 2

    def __init__(self, scenes, dll_path=None):
        self.lib = load_batch_lib(dll_path)
        counts = [int(s["pos"].shape[0]) for s in scenes]
        ecounts = [int(s["edges"].shape[0]) for s in scenes]
        self.node_off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
        self.edge_off = np.concatenate([[0], np.cumsum(ecounts)]).astype(np.int32)
        cat = lambda k, dt, w: np.ascontiguousarray(np.concatenate([np.asarray(s[k], dt).reshape(-1, w) for s in scenes]))
        pos, sc, rt = cat("pos", np.float32, 3), cat("scales", np.float32, 3), cat("rots", np.float32, 4)
        op = np.ascontiguousarray(np.concatenate([np.asarray(s["opacity"], np.float32) for s in scenes]))
        ed = cat("edges", np.int32, 2) if sum(ecounts) else np.zeros((1, 2), np.int32)
        rest = np.ascontiguousarray(np.concatenate([np.asarray(s["rest"], np.float32) for s in scenes])) \
            if sum(ecounts) else np.zeros(1, np.float32)
        stiff = np.ascontiguousarray(np.concatenate([np.asarray(s["stiff"], np.float32) for s in scenes])) \
            if sum(ecounts) else np.zeros(1, np.float32)
        h = self.lib.xpbd_batch_create(len(scenes), _p(self.node_off), _p(pos), _p(sc), _p(rt), _p(op),
                                       _p(self.edge_off), _p(ed), _p(rest), _p(stiff))
        if not h:
            raise RuntimeError("xpbd_batch_create: " + self.lib.xpbd_batch_last_error(None).decode(errors="replace"))
        self.h = h
        self.counts = counts

    # -- lifetime -------------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.xpbd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what}: {self.lib.xpbd_batch_last_error(self.h).decode(errors='replace')}")

    @property
    def objects(self):
        return len(self.counts)

    def count(self, obj=-1):
        return self.lib.xpbd_batch_count(self.h, obj)

    # -- settings (XPBD defaults) ----------------------------------------------
    def set_solver(self, obj=-1, iters=20, dt=1 / 60, under_relax=0.6, vel_damping=0.01, stiffness_scale=1.0,
                   inv_mass_scale=1.0, dist_compliance=0.0, shape_compliance=0.0, shape_blend=0.1,
                   angle_compliance=1e-2, angle_blend=0.2):
        self._check(self.lib.xpbd_batch_set_solver(self.h, obj, iters, dt, under_relax, vel_damping, stiffness_scale,
                                                   inv_mass_scale, dist_compliance, shape_compliance, shape_blend,
                                                   angle_compliance, angle_blend), "xpbd_batch_set_solver")

    def set_constraints(self, obj=-1, distance=True, shape=True, angle=False, volume=True):
        self._check(self.lib.xpbd_batch_set_constraints(self.h, obj, int(distance), int(shape), int(angle), int(volume)),
                    "xpbd_batch_set_constraints")

    def set_volume(self, obj=-1, compliance=1e-6, ring_k=3, max_members=2048, leader_min_hop=2):
        self._check(self.lib.xpbd_batch_set_volume(self.h, obj, compliance, ring_k, max_members, leader_min_hop),
                    "xpbd_batch_set_volume")

    def set_ground(self, obj, enabled, up, height, friction=0.5, restitution=0.0, contact_radius=0.0,
                   contact_slop=0.0, gravity=0.0):
        self._check(self.lib.xpbd_batch_set_ground(self.h, obj, int(enabled), float(up[0]), float(up[1]), float(up[2]),
                                                   height, friction, restitution, contact_radius, contact_slop,
                                                   gravity), "xpbd_batch_set_ground")

    def set_object_shape(self, obj, stiffness):
        self._check(self.lib.xpbd_batch_set_object_shape(self.h, obj, stiffness), "xpbd_batch_set_object_shape")

    def set_self_collision(self, obj, enabled, radius_scale=1.5, exclude_scale=2.0, within_body=False):
        self._check(self.lib.xpbd_batch_set_self_collision(self.h, obj, int(enabled), radius_scale, exclude_scale,
                                                           int(within_body)), "xpbd_batch_set_self_collision")

    def set_self_collision_grid(self, obj, cell_size=0.0, table_size=0):
        self._check(self.lib.xpbd_batch_set_self_collision_grid(self.h, obj, float(cell_size), int(table_size)),
                    "xpbd_batch_set_self_collision_grid")

    def set_gauss_seidel(self, obj, distance, volume):
        self._check(self.lib.xpbd_batch_set_gauss_seidel(self.h, obj, int(distance), int(volume)),
                    "xpbd_batch_set_gauss_seidel")

    def set_shape_robust(self, obj, enabled):
        self._check(self.lib.xpbd_batch_set_shape_robust(self.h, obj, int(enabled)), "xpbd_batch_set_shape_robust")

    def set_volume_clusters(self, obj, off=None, members=None, min_members=4):
        if off is None:
            self._check(self.lib.xpbd_batch_set_volume_clusters(self.h, obj, None, None, int(min_members)),
                        "xpbd_batch_set_volume_clusters")
            return
        off, members = _i32(off), _i32(members)
        if members.size == 0:
            members = np.zeros(1, np.int32)
        self._check(self.lib.xpbd_batch_set_volume_clusters(self.h, obj, _p(off), _p(members), int(min_members)),
                    "xpbd_batch_set_volume_clusters")

    def set_shape_clusters(self, obj, cluster_of=None):
        c = None if cluster_of is None else _i32(cluster_of)
        n = self.counts[obj] if c is None else c.size
        self._check(self.lib.xpbd_batch_set_shape_clusters(self.h, obj, _p(c), n), "xpbd_batch_set_shape_clusters")

    def set_colliders(self, obj, types, poses, dims, margin=0.0, friction=0.0):
        types = _i32(np.asarray(types).reshape(-1))
        k = len(types)
        poses = _f32(poses, (k, 12)) if k else np.zeros((1, 12), np.float32)
        dims = _f32(dims, (k, 3)) if k else np.zeros((1, 3), np.float32)
        if not k:
            types = np.zeros(1, np.int32)
        self._check(self.lib.xpbd_batch_set_colliders(self.h, obj, k, _p(types), _p(poses), _p(dims), float(margin),
                                                      float(friction)), "xpbd_batch_set_colliders")

    def set_attached(self, requests, weight=1.0):
        """requests: list of (obj, idx, pos) — idx local ids, pos (k,3);
        an empty idx releases that object."""
        if not requests:
            return
        objs = _i32([r[0] for r in requests])
        sizes = [len(np.asarray(r[1]).reshape(-1)) for r in requests]
        off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
        idx = _i32(np.concatenate([np.asarray(r[1], np.int64).reshape(-1) for r in requests])) if off[-1] else \
            np.zeros(1, np.int32)
        pos = _f32(np.concatenate([np.asarray(r[2], np.float32).reshape(-1, 3) for r in requests])) if off[-1] else \
            np.zeros((1, 3), np.float32)
        self._check(self.lib.xpbd_batch_set_attached(self.h, len(requests), _p(objs), _p(off), _p(idx), _p(pos),
                                                     float(weight)), "xpbd_batch_set_attached")

    def add_rigid_velocity(self, obj, dv, dw=(0.0, 0.0, 0.0), center=(0.0, 0.0, 0.0)):
        a = [_f32(v) for v in (dv, dw, center)]
        self._check(self.lib.xpbd_batch_add_rigid_velocity(self.h, obj, *[_p(x) for x in a]),
                    "xpbd_batch_add_rigid_velocity")

    def reset(self, obj=-1):
        self._check(self.lib.xpbd_batch_reset(self.h, obj), "xpbd_batch_reset")

    # -- stepping ---------------------------------------------------------------
    def prepare(self):
        self._check(self.lib.xpbd_batch_prepare(self.h), "xpbd_batch_prepare")

    def step(self, substeps=1):
        self._check(self.lib.xpbd_batch_step(self.h, substeps), "xpbd_batch_step")

    def step_async(self, substeps=1):
        self._check(self.lib.xpbd_batch_step_async(self.h, substeps), "xpbd_batch_step_async")

    def sync(self):
        self._check(self.lib.xpbd_batch_sync(self.h), "xpbd_batch_sync")

    # -- outputs ----------------------------------------------------------------
    def positions_device_ptr(self):
        return int(self.lib.xpbd_batch_positions_device(self.h))

    def positions(self, obj):
        out = np.empty((self.counts[obj], 3), np.float32)
        self._check(self.lib.xpbd_batch_positions_host(self.h, obj, _p(out), self.counts[obj]),
                    "xpbd_batch_positions_host")
        return out

    def get_state(self, obj):
        p = np.empty((self.counts[obj], 3), np.float32)
        v = np.empty_like(p)
        self._check(self.lib.xpbd_batch_get_state(self.h, obj, _p(p), _p(v), self.counts[obj]), "xpbd_batch_get_state")
        return p, v

    def set_state(self, obj, pos=None, vel=None):
        p = None if pos is None else _f32(pos)
        v = None if vel is None else _f32(vel)
        self._check(self.lib.xpbd_batch_set_state(self.h, obj, _p(p), _p(v), self.counts[obj]), "xpbd_batch_set_state")

    def shapes(self, obj, deform_eps=1e-2):
        n = self.counts[obj]
        sc = np.empty((n, 3), np.float32)
        q = np.empty((n, 4), np.float32)
        k = self.lib.xpbd_batch_compute_shapes(self.h, obj, float(deform_eps), _p(sc), _p(q), n)
        if k < 0:
            self._check(-1, "xpbd_batch_compute_shapes")
        return sc, q, k

    def shapes_device(self, deform_eps=1e-2):
        """Every object's deformed shapes into device memory (one launch)."""
        self._check(self.lib.xpbd_batch_compute_shapes(self.h, -1, float(deform_eps), None, None, 0),
                    "xpbd_batch_compute_shapes")

    def shapes_device_ptrs(self):
        ptrs = np.zeros(3, np.uint64)
        p = self.lib.xpbd_batch_shapes_device(self.h, _p(ptrs))
        if p < 0:
            self._check(-1, "xpbd_batch_shapes_device")
        return int(ptrs[0]), int(ptrs[1]), int(ptrs[2]), p

    def self_collision_stats(self, obj):
        out = np.zeros(3, np.int32)
        self._check(self.lib.xpbd_batch_self_collision_stats(self.h, obj, _p(out)), "xpbd_batch_self_collision_stats")
        return int(out[0]), int(out[1]), int(out[2])

    def copy_object(self, dst_obj, src, src_obj):
        self._check(self.lib.xpbd_batch_copy_object(self.h, dst_obj, src.h, src_obj), "xpbd_batch_copy_object")

    # -- tuning and profiling -----------------------------------------------------
    def autotune(self, substeps=2, hash_tables=False):
        buf = ctypes.create_string_buffer(4096)
        self._check(self.lib.xpbd_batch_autotune(self.h, substeps, 1 if hash_tables else 0, buf, len(buf)),
                    "xpbd_batch_autotune")
        return json.loads(buf.value.decode())

    def tuning(self):
        t = np.zeros(7, np.int32)
        self._check(self.lib.xpbd_batch_get_tuning(self.h, _p(t)), "xpbd_batch_get_tuning")
        return [int(x) for x in t]

    def set_tuning(self, t):
        t = _i32(t)
        self._check(self.lib.xpbd_batch_set_tuning(self.h, _p(t)), "xpbd_batch_set_tuning")

    def profile(self, substeps):
        buf = ctypes.create_string_buffer(1 << 16)
        self._check(self.lib.xpbd_batch_profile(self.h, substeps, buf, len(buf)), "xpbd_batch_profile")
        return json.loads(buf.value.decode())
