"""Seeded synthetic Gaussian-graph scenes for the XPBD solver.

Every scene is a dict shaped like the reference loader's output
(src/xpbd_bridge.py:17-41): C-contiguous float32 ``pos`` (N,3), ``scales``
(N,3, activated), ``rots`` (N,4, w,x,y,z normalised), ``opacity`` (N,
activated), int32 ``edges`` (E,2) with i<j once each, float32 ``rest`` (E,) and
``stiff`` (E,).  Edges are the symmetric union of the k-nearest-neighbour
graph; rest lengths are the float64 distances of the float32 positions rounded
to float32.

The recipes follow the workload descriptions: an ellipsoid of Gaussians resting
on the z=0 plane, uniformly sampled by rejection with a PCG64 generator.
"""
import numpy as np
from scipy.spatial import cKDTree


def ellipsoid_points(n, semi=(0.25, 0.15, 0.10), centre=(0.0, 0.0, 0.10), seed=1):
    """n points uniform in an axis-aligned ellipsoid (rejection sampling)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    semi = np.asarray(semi, np.float64)
    out = np.empty((0, 3), np.float64)
    while out.shape[0] < n:
        u = rng.uniform(-1.0, 1.0, size=(max(2 * n, 1024), 3))
        u = u[(u * u).sum(1) <= 1.0]
        out = np.concatenate([out, u], 0)
    pts = out[:n] * semi + np.asarray(centre, np.float64)
    return pts.astype(np.float32), rng


def knn_union_edges(pos, k):
    """Undirected symmetric-union kNN edges (i<j), sorted, with float32 rest."""
    n = pos.shape[0]
    tree = cKDTree(pos.astype(np.float64))
    _, nb = tree.query(pos.astype(np.float64), k=k + 1, workers=-1)
    nb = nb[:, 1:].astype(np.int64)
    src = np.repeat(np.arange(n, dtype=np.int64), k)
    dst = nb.reshape(-1)
    a = np.minimum(src, dst)
    b = np.maximum(src, dst)
    keep = a != b
    key = np.unique(a[keep] * n + b[keep])
    e = np.stack([key // n, key % n], 1).astype(np.int32)
    d = pos[e[:, 0]].astype(np.float64) - pos[e[:, 1]].astype(np.float64)
    rest = np.sqrt((d * d).sum(1)).astype(np.float32)
    return np.ascontiguousarray(e), rest


def gaussian_attributes(n, rng):
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    scales = np.exp(rng.normal(-4.0, 0.3, size=(n, 3)))
    opacity = np.full(n, 0.9)
    return (q.astype(np.float32), scales.astype(np.float32), opacity.astype(np.float32))


def make_ellipsoid_scene(n, k, semi=(0.25, 0.15, 0.10), seed=1):
    semi = np.asarray(semi, np.float64)
    pos, rng = ellipsoid_points(n, semi=semi, centre=(0.0, 0.0, semi[2]), seed=seed)
    rots, scales, opacity = gaussian_attributes(n, rng)
    edges, rest = knn_union_edges(pos, k)
    return {
        "pos": np.ascontiguousarray(pos),
        "scales": np.ascontiguousarray(scales),
        "rots": np.ascontiguousarray(rots),
        "opacity": np.ascontiguousarray(opacity),
        "edges": edges,
        "rest": np.ascontiguousarray(rest),
        "stiff": np.ones(edges.shape[0], np.float32),
        "seed": seed,
        "semi": semi.astype(np.float32),
    }


def kmeans_clusters(pos, size=64, iters=20, seed=0):
    """Seeded Lloyd k-means with K = N / size clusters; returns the cluster of
    every node (int32).  Assignment by a k-d tree over the centroids, so one
    iteration is O(N log K)."""
    from scipy.spatial import cKDTree
    p = np.asarray(pos, np.float64)
    n = p.shape[0]
    k = max(1, n // size)
    rng = np.random.Generator(np.random.PCG64(seed))
    cent = p[rng.choice(n, size=k, replace=False)]
    lab = np.zeros(n, np.int64)
    for _ in range(iters):
        lab = cKDTree(cent).query(p)[1]
        cnt = np.bincount(lab, minlength=k).astype(np.float64)
        for a in range(3):
            s = np.bincount(lab, weights=p[:, a], minlength=k)
            cent[:, a] = np.where(cnt > 0, s / np.maximum(cnt, 1), cent[:, a])
    # empty clusters are dropped and the ids compacted
    used = np.unique(lab)
    remap = np.full(k, -1, np.int64)
    remap[used] = np.arange(len(used))
    return remap[lab].astype(np.int32)


def full_recipe(scene, cluster_size=64, seed=0):
    """The full single-object recipe: one volume cluster per node (the node and
    its 3 nearest neighbours, a tetrahedron), k-means shape clusters of about
    `cluster_size` nodes, self-collision within the body."""
    from scipy.spatial import cKDTree
    pos = scene["pos"]
    n = pos.shape[0]
    _, nn = cKDTree(pos).query(pos, 4)
    tet_off = np.arange(0, 3 * n + 1, 3, dtype=np.int32)
    tet_mem = np.ascontiguousarray(nn[:, 1:].astype(np.int32).ravel())
    table = 1 << 21 if n <= 300_000 else 1 << 23
    return {"tet_off": tet_off, "tet_mem": tet_mem,
            "clusters": kmeans_clusters(pos, size=cluster_size, seed=seed), "table": table}


def configure_full(sim, recipe, iters=4, substeps=10):
    """Solver settings of the full recipe (after create, before the first step).
    Self-collision keeps the reference's cell rule (search radius + twice the
    largest predicted move); the table size follows the recipe."""
    sim.set_solver(iters=iters, dt=1.0 / (60.0 * substeps), under_relax=0.6, vel_damping=0.01,
                   dist_compliance=1e-6, shape_compliance=0.0, shape_blend=0.1)
    sim.set_constraints(distance=True, shape=True, angle=False, volume=True)
    sim.ext_set_volume_clusters(recipe["tet_off"], recipe["tet_mem"], 3)
    sim.ext_set_shape_clusters(recipe["clusters"])
    sim.set_ground(True, (0.0, 0.0, 1.0), 0.0, friction=0.5, restitution=0.0, gravity=9.81)
    sim.set_object_shape(0.0)
    sim.ext_set_self_collision_grid(0.0, recipe["table"])
    sim.set_self_collision(True, 1.5, 2.0, True)


def merge_scenes(parts, offsets):
    """Several scenes as one graph (disjoint components), part p moved by
    offsets[p].  The components become the objects of object shape matching
    and cross-object self-collision."""
    keys = ("pos", "scales", "rots", "opacity", "rest", "stiff")
    out = {k: [] for k in keys}
    edges, base = [], 0
    for sc, off in zip(parts, offsets):
        out["pos"].append(sc["pos"] + np.asarray(off, np.float32))
        for k in keys[1:]:
            out[k].append(sc[k])
        edges.append(sc["edges"] + base)
        base += sc["pos"].shape[0]
    merged = {k: np.ascontiguousarray(np.concatenate(v).astype(np.float32)) for k, v in out.items()}
    merged["edges"] = np.ascontiguousarray(np.concatenate(edges).astype(np.int32))
    merged["parts"] = [sc["pos"].shape[0] for sc in parts]
    return merged


def stacked_bodies(n_each, k, semi=(0.1, 0.1, 0.1), gap=0.0, count=2, seed=5):
    """`count` ellipsoids stacked along z with `gap` between their surfaces
    (negative = overlap), the lowest resting on z = 0."""
    parts, offs = [], []
    for c in range(count):
        parts.append(make_ellipsoid_scene(n_each, k, semi=semi, seed=seed + c))
        offs.append((0.0, 0.0, c * (2.0 * semi[2] + gap)))
    return merge_scenes(parts, offs)


def scene_extent(scene):
    p = scene["pos"]
    return float(np.linalg.norm(p.max(0) - p.min(0)))


# ----------------------------------------------------------------------------
# Many-object workloads
# ----------------------------------------------------------------------------

def log_uniform_sizes(count, nmin, nmax, rng):
    return np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=count)).astype(np.int64)


def object_request(n, seed, k=12, cluster_size=64, semi_range=(0.05, 0.30), table=None):
    """One object of the many-object workloads: an ellipsoid of n Gaussians with
    semi-axes uniform in semi_range (seeded), the full recipe (table: hash
    cells, default nextpow2(2n) as the reference sizes it) and its own gripper
    pair."""
    rng = np.random.Generator(np.random.PCG64(seed))
    semi = rng.uniform(semi_range[0], semi_range[1], size=3)
    scene = make_ellipsoid_scene(int(n), k, semi=tuple(semi), seed=seed)
    recipe = full_recipe(scene, cluster_size=cluster_size, seed=seed)
    if table is None:
        t = 1024
        while t < 2 * int(n):
            t <<= 1
        table = t
    recipe["table"] = table
    script = GripperScript(scene, seed=seed + 1, attach=min(500, max(4, int(n) // 10)))
    return scene, recipe, script


def object_mix(count=64, nmin=5_000, nmax=200_000, seed=11, **kw):
    """`count` objects with sizes log-uniform in [nmin, nmax] (the batched
    workload: 64 objects, 5k-200k Gaussians, ~3.4M in total)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = log_uniform_sizes(count, nmin, nmax, rng)
    return [object_request(int(n), seed * 1000 + j, **kw) for j, n in enumerate(sizes)]


def configure_full_batch(batch, obj, recipe, iters=4, substeps=10):
    """configure_full for object `obj` of an XPBDBatch."""
    batch.set_solver(obj, iters=iters, dt=1.0 / (60.0 * substeps), under_relax=0.6, vel_damping=0.01,
                     dist_compliance=1e-6, shape_compliance=0.0, shape_blend=0.1)
    batch.set_constraints(obj, distance=True, shape=True, angle=False, volume=True)
    batch.set_volume_clusters(obj, recipe["tet_off"], recipe["tet_mem"], 3)
    batch.set_shape_clusters(obj, recipe["clusters"])
    batch.set_ground(obj, True, (0.0, 0.0, 1.0), 0.0, friction=0.5, restitution=0.0, gravity=9.81)
    batch.set_object_shape(obj, 0.0)
    batch.set_self_collision_grid(obj, 0.0, recipe["table"])
    batch.set_self_collision(obj, True, 1.5, 2.0, True)


# ----------------------------------------------------------------------------
# Scripted scene drivers
# ----------------------------------------------------------------------------

def config1_solver(sim):
    """20k-Gaussian plain XPBD recipe: distance only, ground, gravity."""
    sim.set_solver(iters=4, dt=1.0 / 600.0, under_relax=0.6, vel_damping=0.01, dist_compliance=1e-6)
    sim.set_constraints(distance=True, shape=False, angle=False, volume=False)
    sim.set_ground(True, (0.0, 0.0, 1.0), 0.0, friction=0.5, restitution=0.0, gravity=9.81)


def sphere_pose(frame, frames=100, start=(-0.30, 0.0, 0.10), travel=0.2):
    """Kinematic sphere centre moving +travel in x over ``frames`` frames."""
    t = min(max(frame / float(frames), 0.0), 1.0)
    c = np.array(start, np.float32) + np.array([travel * t, 0.0, 0.0], np.float32)
    pose = np.zeros(12, np.float32)
    pose[[0, 4, 8]] = 1.0
    pose[9:12] = c
    return pose


def config1_colliders(sim, frame, radius=0.05, frames=100):
    sim.set_colliders([0], sphere_pose(frame, frames)[None, :], [[radius, 0.0, 0.0]], margin=0.0, friction=0.0)


def _rot_axis_angle(axis, theta):
    k = np.asarray(axis, np.float64)
    k = k / np.linalg.norm(k)
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + np.sin(theta) * K + (1.0 - np.cos(theta)) * (K @ K)


def _frame_from_axis(n):
    """Rotation whose local z axis is n (capsule axis along local z)."""
    n = np.asarray(n, np.float64) / np.linalg.norm(n)
    a = np.array([1.0, 0.0, 0.0]) if abs(n[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    u = np.cross(n, a)
    u /= np.linalg.norm(u)
    w = np.cross(n, u)
    return np.stack([u, w, n], axis=1)   # columns = local x, y, z in world


class GripperScript:
    """Two capsule grippers (radius 0.02, half-length 0.04) that each attach the
    `attach` nodes nearest a seeded random surface point of the ellipsoid, pull
    them `pull` m along the outward normal and twist them `twist_deg` about the
    normal, linearly over `ramp_frames` frames, then release."""

    def __init__(self, scene, seed=7, grippers=2, attach=500, pull=0.1, twist_deg=30.0, ramp_frames=60,
                 radius=0.02, half_length=0.04, tree=None):
        rng = np.random.Generator(np.random.PCG64(seed))
        pos = scene["pos"].astype(np.float64)
        semi = np.asarray(scene["semi"], np.float64)
        centre = pos.mean(0)
        centre[2] = semi[2]
        self.radius, self.half_length = radius, half_length
        self.pull, self.twist = pull, np.deg2rad(twist_deg)
        self.ramp = ramp_frames
        tree = cKDTree(pos) if tree is None else tree   # a tree of scene["pos"] may be shared
        self.g = []
        for _ in range(grippers):
            d = rng.standard_normal(3)
            d /= np.linalg.norm(d)
            d[2] = abs(d[2])   # grip the upper half, pull away from the ground
            s = centre + semi * d
            nrm = (s - centre) / (semi * semi)
            nrm /= np.linalg.norm(nrm)
            k = min(attach, pos.shape[0])
            _, idx = tree.query(s, k=k)
            idx = np.sort(np.atleast_1d(idx)).astype(np.int32)
            self.g.append({"s": s, "n": nrm, "idx": idx, "rest": pos[idx], "R0": _frame_from_axis(nrm)})

    def transform(self, g, frame):
        t = min(max(frame / float(self.ramp), 0.0), 1.0)
        R = _rot_axis_angle(g["n"], self.twist * t)
        shift = g["n"] * (self.pull * t)
        return R, shift

    def colliders(self, frame):
        types, poses, dims = [], [], []
        for g in self.g:
            R, shift = self.transform(g, frame)
            c0 = g["s"] + g["n"] * self.radius
            c = g["s"] + R @ (c0 - g["s"]) + shift
            Rw = R @ g["R0"]
            pose = np.concatenate([Rw.reshape(-1), c]).astype(np.float32)
            types.append(2)
            poses.append(pose)
            dims.append([self.radius, self.half_length, 0.0])
        return np.array(types, np.int32), np.array(poses, np.float32), np.array(dims, np.float32)

    def attached(self, frame):
        """(idx, target positions) while gripping, or None after release."""
        if frame >= self.ramp:
            return None
        idx, tgt = [], []
        for g in self.g:
            R, shift = self.transform(g, frame)
            tgt.append(g["s"] + (g["rest"] - g["s"]) @ R.T + shift)
            idx.append(g["idx"])
        return np.concatenate(idx).astype(np.int32), np.concatenate(tgt).astype(np.float32)

    def apply(self, sim, frame, margin=0.0, friction=0.0):
        t, p, d = self.colliders(frame)
        sim.set_colliders(t, p, d, margin=margin, friction=friction)
        att = self.attached(frame)
        if att is None:
            if frame == self.ramp:
                sim.clear_attached()
        else:
            sim.set_attached(att[0], att[1], 1.0)
