"""On-disk physics graphs.

* ``APGGRPH1`` binary (the reference viewer's export, forward.cu:4612-4663):
  little endian ``"APGGRPH1" | int32 N | int32 M | float32 pos_rest[N*3] |
  int32 offset[N] | int32 count[N] | int32 idx[M] | float32 dist[M] |
  float32 stiff[M]``, nodes in the viewer's order.
* The NPZ graph read by ``xpbd_bridge.load_inputs`` (offset, count, idx,
  dist, stiff, pos_rest, sibr_to_usd), nodes in the order of the cropped PLY
  (tools/import_graph.py:124-126).

``import_graph`` converts the first into the second the way the reference's
tools/import_graph.py does: every viewer node is matched to the PLY splat with
the same float32 position bit pattern, each USD node takes its viewer node's
neighbour list in the same order, and neighbour ids are renumbered.
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


def _position_keys(p):
    return np.ascontiguousarray(p, np.float32).view(np.dtype((np.void, 12))).ravel()


def import_graph(graph, usd_pos):
    """Viewer-order graph (read_apggrph1's dict) -> NPZ fields in the order
    of the PLY positions ``usd_pos`` (N,3).  Every viewer node must match
    exactly one splat by its float32 bit pattern."""
    pos_s = np.asarray(graph["pos"], np.float32)
    usd = np.asarray(usd_pos, np.float32).reshape(-1, 3)
    n = pos_s.shape[0]
    if usd.shape[0] != n:
        raise ValueError(f"graph has {n} nodes, the PLY {usd.shape[0]} splats")
    ku, ks = _position_keys(usd), _position_keys(pos_s)
    uniq, first, cnt = np.unique(ku, return_index=True, return_counts=True)
    at = np.clip(np.searchsorted(uniq, ks), 0, max(len(uniq) - 1, 0))
    if n and not (uniq[at] == ks).all():
        raise ValueError(f"{int((uniq[at] != ks).sum())} graph nodes have no splat with the same position bits")
    if (cnt > 1).any():
        raise ValueError(f"{int((cnt > 1).sum())} splat positions occur more than once; the match is ambiguous")
    sibr_to_usd = first[at].astype(np.int64)
    usd_to_sibr = np.empty(n, np.int64)
    usd_to_sibr[sibr_to_usd] = np.arange(n)
    off_s = np.asarray(graph["offset"], np.int64)
    cnt_s = np.asarray(graph["count"], np.int64)
    count = cnt_s[usd_to_sibr]
    offset = np.zeros(n, np.int64)
    if n:
        offset[1:] = np.cumsum(count)[:-1]
    total = int(count.sum())
    # entry t of USD node u comes from entry t of viewer node usd_to_sibr[u]
    src_start = np.repeat(off_s[usd_to_sibr], count)
    within = np.arange(total) - np.repeat(offset, count)
    gather = src_start + within
    idx_s = np.asarray(graph["idx"], np.int64)[gather]
    valid = (idx_s >= 0) & (idx_s < n)
    idx = np.where(valid, sibr_to_usd[np.clip(idx_s, 0, max(n - 1, 0))], -1).astype(np.int32)
    return {
        "offset": offset.astype(np.int32), "count": count.astype(np.int32), "idx": idx,
        "dist": np.asarray(graph["dist"], np.float32)[gather], "stiff": np.asarray(graph["stiff"], np.float32)[gather],
        "pos_rest": usd, "sibr_to_usd": sibr_to_usd.astype(np.int32),
    }


def save_graph_npz(path, g):
    np.savez_compressed(path, **{k: g[k] for k in ("offset", "count", "idx", "dist", "stiff", "pos_rest",
                                                   "sibr_to_usd")})
