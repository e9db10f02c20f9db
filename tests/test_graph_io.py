"""On-disk graphs (CPU only): the APGGRPH1 export format of the reference
viewer (forward.cu:4612-4663) and its conversion to the PLY-ordered NPZ
(tools/import_graph.py), checked by round trips through ``load_inputs``
(src/xpbd_bridge.py:17-41) back to the edges a scene started from."""
import numpy as np
import pytest

from harness import PKG  # noqa: F401  (puts the package on sys.path)
import graph_io
import scenes
import splat_io
from xpbd_bridge import load_inputs


def loop_csr(n, edges, rest, stiff):
    """Per-edge pushes into per-node lists (the loadGraph loop)."""
    nb = [[] for _ in range(n)]
    for (a, b), r, s in zip(edges, rest, stiff):
        if a < 0 or b < 0 or a >= n or b >= n or a == b:
            continue
        nb[a].append((b, r, s))
        nb[b].append((a, r, s))
    return nb


def test_csr_from_edges_follows_the_edge_order():
    rng = np.random.default_rng(2)
    n = 40
    edges = rng.integers(-2, n + 2, size=(150, 2))
    rest = rng.uniform(0.1, 1.0, 150).astype(np.float32)
    stiff = rng.uniform(0.5, 2.0, 150).astype(np.float32)
    off, cnt, idx, dist, st = graph_io.csr_from_edges(n, edges, rest, stiff)
    ref = loop_csr(n, edges, rest, stiff)
    for v in range(n):
        got = list(zip(idx[off[v]:off[v] + cnt[v]], dist[off[v]:off[v] + cnt[v]], st[off[v]:off[v] + cnt[v]]))
        assert [(int(a), float(b), float(c)) for a, b, c in got] == [(a, float(b), float(c)) for a, b, c in ref[v]]


def test_apggrph1_round_trip_and_rejections(tmp_path):
    rng = np.random.default_rng(5)
    pos = rng.normal(size=(9, 3)).astype(np.float32)
    off, cnt, idx, dist, st = graph_io.csr_from_edges(9, [[0, 1], [1, 2], [3, 8], [2, 0]], [1, 2, 3, 4], [1, 1, 2, 2])
    p = tmp_path / "g.bin"
    graph_io.write_apggrph1(p, pos, off, cnt, idx, dist, st)
    assert p.stat().st_size == 16 + 12 * 9 + 8 * 9 + 12 * len(idx)
    g = graph_io.read_apggrph1(p)
    np.testing.assert_array_equal(g["pos"], pos)
    for k, v in (("offset", off), ("count", cnt), ("idx", idx), ("dist", dist), ("stiff", st)):
        np.testing.assert_array_equal(g[k], v)
    raw = p.read_bytes()
    (tmp_path / "magic.bin").write_bytes(b"APGGRPH2" + raw[8:])
    (tmp_path / "short.bin").write_bytes(raw[:-4])
    for bad in ("magic.bin", "short.bin"):
        with pytest.raises(ValueError):
            graph_io.read_apggrph1(tmp_path / bad)


def viewer_export(tmp_path, n=800, seed=4):
    """A scene in PLY (USD) order, its graph exported in a shuffled viewer
    order, and the cropped PLY."""
    scene = scenes.make_ellipsoid_scene(n, 6, semi=(0.1, 0.08, 0.05), seed=seed)
    pos_u = scene["pos"]
    rng = np.random.default_rng(seed)
    perm = rng.permutation(n)                  # viewer node s is USD node perm[s]
    usd_to_view = np.empty(n, np.int64)
    usd_to_view[perm] = np.arange(n)
    ev = usd_to_view[scene["edges"]]
    graph_io.write_apggrph1(tmp_path / "toy_graph_sibr.bin", pos_u[perm],
                            *graph_io.csr_from_edges(n, ev, scene["rest"], scene["stiff"]))
    names = ["x", "y", "z", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3", "opacity"]
    data = np.zeros(n, np.dtype([(k, "<f4") for k in names]))
    data["x"], data["y"], data["z"] = pos_u[:, 0], pos_u[:, 1], pos_u[:, 2]
    ls = np.log(scene["scales"])
    data["scale_0"], data["scale_1"], data["scale_2"] = ls[:, 0], ls[:, 1], ls[:, 2]
    for k in range(4):
        data[f"rot_{k}"] = scene["rots"][:, k]
    data["opacity"] = 2.0
    splat_io.write_ply(tmp_path / "toy_crop.ply", splat_io.vertex_header(names, n), data)
    return scene, perm


def test_import_then_load_inputs_gives_back_the_edges(tmp_path):
    scene, perm = viewer_export(tmp_path)
    g = graph_io.read_apggrph1(tmp_path / "toy_graph_sibr.bin")
    _, data = splat_io.read_ply(tmp_path / "toy_crop.ply")
    usd = np.stack([data["x"], data["y"], data["z"]], axis=1)
    out = graph_io.import_graph(g, usd)
    np.testing.assert_array_equal(out["sibr_to_usd"], perm)
    # every USD node has its viewer node's list, in order, renumbered
    for u in range(0, len(perm), 37):
        s = int(np.nonzero(perm == u)[0][0])
        a = g["idx"][g["offset"][s]:g["offset"][s] + g["count"][s]]
        b = out["idx"][out["offset"][u]:out["offset"][u] + out["count"][u]]
        np.testing.assert_array_equal(perm[a], b)
    graph_io.save_graph_npz(tmp_path / "toy_graph.npz", out)
    inp = load_inputs(str(tmp_path), "toy")
    want = {(min(a, b), max(a, b)): (r, s) for (a, b), r, s in zip(scene["edges"], scene["rest"], scene["stiff"])}
    got = {(int(a), int(b)): (r, s) for (a, b), r, s in zip(inp["edges"], inp["rest"], inp["stiff"])}
    assert got == want
    np.testing.assert_array_equal(inp["pos"], scene["pos"])


def test_import_refuses_unmatched_or_ambiguous_positions(tmp_path):
    viewer_export(tmp_path, n=200)
    g = graph_io.read_apggrph1(tmp_path / "toy_graph_sibr.bin")
    _, data = splat_io.read_ply(tmp_path / "toy_crop.ply")
    usd = np.stack([data["x"], data["y"], data["z"]], axis=1)
    moved = usd.copy()
    moved[7, 0] = np.nextafter(moved[7, 0], np.float32(1e9))   # one bit off
    with pytest.raises(ValueError):
        graph_io.import_graph(g, moved)
    doubled = usd.copy()
    doubled[3] = doubled[4]
    with pytest.raises(ValueError):
        graph_io.import_graph(g, doubled)
    with pytest.raises(ValueError):
        graph_io.import_graph(g, usd[:-1])
