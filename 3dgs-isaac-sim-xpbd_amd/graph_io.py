"""On-disk physics graphs.

* ``APGGRPH1`` binary (the reference viewer's export, forward.cu:4612-4663):
  little endian ``"APGGRPH1" | int32 N | int32 M | float32 pos_rest[N*3] |
  int32 offset[N] | int32 count[N] | int32 idx[M] | float32 dist[M] |
  float32 stiff[M]``, nodes in the viewer's order.
* The NPZ graph read by ``xpbd_bridge.load_inputs`` (offset, count, idx,
  dist, stiff, pos_rest, sibr_to_usd), nodes in the order of the cropped PLY
  (tools/import_graph.py:124-126).

``import_graph`` converts the first into the second with the result of the
reference's tools/import_graph.py: every viewer node is matched to the PLY
splat with the same float32 position bit pattern, each USD node takes its
viewer node's neighbour list in the same order, and neighbour ids are
renumbered.
``csr_from_edges`` builds the symmetric CSR the solver derives from an edge
list (ChainMail::loadGraph, forward.cu:2702-2779), for writing graphs.
"""
import numpy as np

MAGIC = b"APGGRPH1"


def read_apggrph1(path):
    """Read an APGGRPH1 file: dict with pos (N,3) float32, offset, count,
    idx (int32) and dist, stiff (float32).  Raises ValueError on a wrong
    magic or a size that does not match N and M."""
    raw = open(path, "rb").read()
    if raw[:8] != MAGIC:
        raise ValueError(f"{path}: not an APGGRPH1 graph (magic {raw[:8]!r})")
    if len(raw) < 16:
        raise ValueError(f"{path}: truncated header")
    n, m = (int(v) for v in np.frombuffer(raw, "<i4", 2, 8))
    if n < 0 or m < 0:
        raise ValueError(f"{path}: negative sizes N={n} M={m}")
    want = 16 + 12 * n + 8 * n + 12 * m
    if len(raw) != want:
        raise ValueError(f"{path}: {len(raw)} bytes, the header's N={n} M={m} needs {want}")
    out, at = {}, 16
    for name, dt, count, shape in (("pos", "<f4", 3 * n, (n, 3)), ("offset", "<i4", n, (n,)),
                                   ("count", "<i4", n, (n,)), ("idx", "<i4", m, (m,)),
                                   ("dist", "<f4", m, (m,)), ("stiff", "<f4", m, (m,))):
        a = np.frombuffer(raw, dt, count, at)
        at += a.nbytes
        out[name] = a.reshape(shape).astype(a.dtype.newbyteorder("="))
    return out


def write_apggrph1(path, pos, offset, count, idx, dist, stiff):
    pos = np.ascontiguousarray(pos, "<f4").reshape(-1, 3)
    n, m = pos.shape[0], int(np.asarray(idx).size)
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(np.array([n, m], "<i4").tobytes())
        f.write(pos.tobytes())
        for a, dt in ((offset, "<i4"), (count, "<i4"), (idx, "<i4"), (dist, "<f4"), (stiff, "<f4")):
            f.write(np.ascontiguousarray(a, dt).tobytes())


def csr_from_edges(n, edges, rest, stiff):
    """Symmetric CSR of an undirected edge list as the solver builds it:
    both directions of every valid edge, each node's entries in edge order,
    edges with an end out of range or equal ends dropped.  Returns offset,
    count, idx (int32), dist, stiff (float32)."""
    e = np.asarray(edges, np.int64).reshape(-1, 2)
    rest = np.asarray(rest, np.float32)
    stiff = np.asarray(stiff, np.float32)
    ok = (e[:, 0] >= 0) & (e[:, 1] >= 0) & (e[:, 0] < n) & (e[:, 1] < n) & (e[:, 0] != e[:, 1])
    e, rest, stiff = e[ok], rest[ok], stiff[ok]
    # entry k of the doubled list: (a_k -> b_k) for k < E, (b -> a) after;
    # a node's entries follow the edge order in which it appears as either end
    src = np.stack([e[:, 0], e[:, 1]], axis=1).ravel()
    dst = np.stack([e[:, 1], e[:, 0]], axis=1).ravel()
    d2 = np.repeat(rest, 2)
    s2 = np.repeat(stiff, 2)
    order = np.argsort(src, kind="stable")
    count = np.bincount(src, minlength=n).astype(np.int32)
    offset = np.zeros(n, np.int32)
    offset[1:] = np.cumsum(count)[:-1]
    return offset, count, dst[order].astype(np.int32), d2[order], s2[order]


def _bit_rows(p):
    """(n, 3) uint32 bit patterns of float32 positions."""
    return np.ascontiguousarray(p, np.float32).reshape(-1, 3).view(np.uint32)


def match_positions(viewer_pos, splat_pos):
    """For every viewer node, the splat with bit-identical float32 position.
    Both sets are sorted lexicographically by bit pattern and must then be
    equal row for row, which makes the match one-to-one.  Raises ValueError
    when the sizes differ, a splat position repeats, or a viewer node has no
    splat."""
    a, b = _bit_rows(viewer_pos), _bit_rows(splat_pos)
    if a.shape[0] != b.shape[0]:
        raise ValueError(f"graph has {a.shape[0]} nodes, the PLY {b.shape[0]} splats")
    oa, ob = np.lexsort(a.T[::-1]), np.lexsort(b.T[::-1])
    sa, sb = a[oa], b[ob]
    repeated = int((sb[1:] == sb[:-1]).all(1).sum())
    if repeated:
        raise ValueError(f"{repeated} splat positions occur more than once; the match is ambiguous")
    differ = (sa != sb).any(1)
    if differ.any():
        # viewer rows absent from the splats (the sorted runs no longer line up)
        keys = lambda r: r.view(np.dtype((np.void, 12))).ravel()
        missing = int((~np.isin(keys(sa), keys(sb))).sum())
        raise ValueError(f"{missing} graph nodes have no splat with the same position bits")
    viewer_to_splat = np.empty(a.shape[0], np.int64)
    viewer_to_splat[oa] = ob
    return viewer_to_splat


def import_graph(graph, usd_pos):
    """Viewer-order graph (read_apggrph1's dict) -> NPZ fields in the order of
    the PLY positions ``usd_pos`` (N, 3) (what tools/import_graph.py:73-126
    writes).  Splat u takes the neighbour list of its viewer node, entries in
    the same order, neighbours renumbered to splats (-1 where the viewer's id
    is out of range); dist and stiff follow their entries."""
    usd = np.asarray(usd_pos, np.float32).reshape(-1, 3)
    viewer_to_splat = match_positions(graph["pos"], usd)
    n = usd.shape[0]
    splat_to_viewer = np.empty(n, np.int64)
    splat_to_viewer[viewer_to_splat] = np.arange(n)
    count = np.asarray(graph["count"], np.int64)[splat_to_viewer]
    offset = np.concatenate([[0], np.cumsum(count)[:-1]]).astype(np.int64) if n else np.zeros(0, np.int64)
    # entry e of the new CSR (row u, slot e - offset[u]) reads viewer entry
    # viewer_offset[row] + slot
    shift = np.asarray(graph["offset"], np.int64)[splat_to_viewer] - offset
    src = np.repeat(shift, count) + np.arange(int(count.sum()), dtype=np.int64)
    old = np.asarray(graph["idx"], np.int64)[src]
    inside = (old >= 0) & (old < n)
    idx = np.full(old.shape, -1, np.int32)
    idx[inside] = viewer_to_splat[old[inside]]
    return {
        "offset": offset.astype(np.int32), "count": count.astype(np.int32), "idx": idx,
        "dist": np.asarray(graph["dist"], np.float32)[src], "stiff": np.asarray(graph["stiff"], np.float32)[src],
        "pos_rest": usd, "sibr_to_usd": viewer_to_splat.astype(np.int32),
    }


def save_graph_npz(path, g):
    np.savez_compressed(path, **{k: g[k] for k in ("offset", "count", "idx", "dist", "stiff", "pos_rest",
                                                   "sibr_to_usd")})
