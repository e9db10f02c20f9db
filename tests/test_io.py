"""Asset loading (CPU only): the PLY reader/writer with the interface of the
reference's src/splat_io.py:9-59, and ``load_inputs`` (src/xpbd_bridge.py:17-41),
which turns ``<name>_crop.ply`` + ``<name>_graph.npz`` into xpbd_create
inputs.  The reference ships no assets (its .gitignore excludes them), so the
files are written here with known values: activations are checked against
their definitions (exp of the log-scales, unit quaternions, logistic opacity)
and the edges against the CSR they came from."""
import numpy as np
import pytest

from harness import PKG  # noqa: F401  (puts the package on sys.path)
import splat_io
from xpbd_bridge import load_inputs

NAMES = ["x", "y", "z", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3", "opacity", "f_dc_0"]


def vertex_array(n, seed=0):
    rng = np.random.default_rng(seed)
    data = np.zeros(n, np.dtype([(k, "<f4") for k in NAMES]))
    for k in NAMES:
        data[k] = rng.normal(0.0, 1.0, n).astype(np.float32)
    for k in ("scale_0", "scale_1", "scale_2"):
        data[k] = rng.normal(-4.0, 0.3, n).astype(np.float32)
    return data


def test_ply_round_trip_keeps_header_and_values(tmp_path):
    data = vertex_array(37)
    header = splat_io.vertex_header(NAMES, 99)   # count is rewritten on write
    header.insert(2, "comment made by a test")
    splat_io.write_ply(tmp_path / "a.ply", header, data)
    h2, d2 = splat_io.read_ply(tmp_path / "a.ply")
    assert h2[2] == "comment made by a test"
    assert "element vertex 37" in h2
    assert d2.dtype.names == tuple(NAMES)
    for k in NAMES:
        np.testing.assert_array_equal(d2[k], data[k])
    # a subset written under the read header
    splat_io.write_ply(tmp_path / "b.ply", h2, d2[5:9])
    _, d3 = splat_io.read_ply(tmp_path / "b.ply")
    np.testing.assert_array_equal(d3["opacity"], data["opacity"][5:9])


@pytest.mark.parametrize("bad", ["ascii", "double", "truncated", "two_elements"])
def test_ply_rejects_what_the_reference_rejects(tmp_path, bad):
    data = vertex_array(4)
    header = splat_io.vertex_header(NAMES, 4)
    if bad == "ascii":
        header[1] = "format ascii 1.0"
    elif bad == "double":
        header[3] = "property double x"
    elif bad == "two_elements":
        header.insert(-1, "element face 2")
    splat_io.write_ply(tmp_path / "c.ply", header, data)
    if bad == "truncated":
        raw = (tmp_path / "c.ply").read_bytes()
        (tmp_path / "c.ply").write_bytes(raw[:-10])
    with pytest.raises(ValueError):
        splat_io.read_ply(tmp_path / "c.ply")


def write_asset(tmp_path, n=6, seed=1):
    data = vertex_array(n, seed)
    splat_io.write_ply(tmp_path / "toy_crop.ply", splat_io.vertex_header(NAMES, n), data)
    pos = np.stack([data["x"], data["y"], data["z"]], axis=1)
    # a symmetric CSR: 0-1, 0-2, 1-2, 3-4, 2-5 (both directions, rows in any order)
    rows = {0: [2, 1], 1: [0, 2], 2: [5, 0, 1], 3: [4], 4: [3], 5: [2]}
    count = np.array([len(rows[i]) for i in range(n)], np.int32)
    offset = np.concatenate([[0], np.cumsum(count)[:-1]]).astype(np.int32)
    idx = np.concatenate([rows[i] for i in range(n)]).astype(np.int32)
    owner = np.repeat(np.arange(n), count)
    dist = np.linalg.norm(pos[owner] - pos[idx], axis=1).astype(np.float32)
    stiff = (1.0 + 0.1 * np.minimum(owner, idx) + 0.01 * np.maximum(owner, idx)).astype(np.float32)
    np.savez(tmp_path / "toy_graph.npz", offset=offset, count=count, idx=idx, dist=dist, stiff=stiff,
             pos_rest=pos.astype(np.float32), sibr_to_usd=np.arange(n, dtype=np.int32))
    return data, pos, (owner, idx, dist, stiff)


def test_load_inputs_activations_and_edges(tmp_path):
    data, pos, (owner, idx, dist, stiff) = write_asset(tmp_path)
    inp = load_inputs(str(tmp_path), "toy")
    np.testing.assert_array_equal(inp["pos"], pos.astype(np.float32))
    s = np.stack([data["scale_0"], data["scale_1"], data["scale_2"]], axis=1).astype(np.float64)
    np.testing.assert_allclose(inp["scales"], np.exp(s), rtol=1e-6)
    q = np.stack([data["rot_0"], data["rot_1"], data["rot_2"], data["rot_3"]], axis=1).astype(np.float64)
    np.testing.assert_allclose(inp["rots"], q / np.linalg.norm(q, axis=1, keepdims=True), rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(inp["opacity"], 1.0 / (1.0 + np.exp(-data["opacity"].astype(np.float64))), rtol=1e-6)
    # each undirected edge once, i < j, in CSR row order
    want = [(0, 2), (0, 1), (1, 2), (2, 5), (3, 4)]
    assert [tuple(e) for e in inp["edges"]] == want
    keep = idx > owner
    np.testing.assert_array_equal(inp["rest"], dist[keep])
    np.testing.assert_array_equal(inp["stiff"], stiff[keep])
    for k in ("pos", "scales", "rots", "opacity", "edges", "rest", "stiff"):
        assert inp[k].flags["C_CONTIGUOUS"]
    assert inp["edges"].dtype == np.int32 and inp["rest"].dtype == np.float32


def test_load_inputs_refuses_a_stale_graph(tmp_path):
    write_asset(tmp_path)
    g = dict(np.load(tmp_path / "toy_graph.npz"))
    g["pos_rest"] = g["pos_rest"] + np.float32(1e-3)
    np.savez(tmp_path / "toy_graph.npz", **g)
    with pytest.raises(ValueError):
        load_inputs(str(tmp_path), "toy")


def test_loaded_asset_steps_on_the_oracle(tmp_path, oracle_path):
    from xpbd_bridge import XPBD
    write_asset(tmp_path)
    inp = load_inputs(str(tmp_path), "toy")
    sim = XPBD(oracle_path)
    sim.create(inp)
    sim.set_solver(iters=2, dt=1 / 600)
    sim.step()
    assert np.isfinite(sim.positions()).all()
    sim.destroy()
