"""Gaussian-splat PLY files: binary little-endian, one ``vertex`` element whose
properties are all ``float``.  Same interface as the reference's
src/splat_io.py:9-59 — ``read_ply(path)`` gives ``(header lines, structured
vertex array)`` and ``write_ply(path, header, data)`` writes ``data`` under
``header`` with the vertex count set to ``len(data)``.

The header is ASCII lines up to ``end_header``; the body is ``count`` packed
records of little-endian float32, one per property in header order (PLY 1.0,
Turk 1994).  The body is memory-mapped and copied out once."""
import mmap
from pathlib import Path

import numpy as np

_END = b"end_header"
_FORMAT = "format binary_little_endian 1.0"


def _split_header(buf, path):
    """Header lines (without line ends) and the body offset of a PLY image."""
    pos = 0
    while True:
        nl = buf.find(b"\n", pos)
        if nl < 0:
            raise ValueError(f"{path}: PLY header has no end_header")
        if buf[pos:nl].rstrip(b"\r") == _END:
            return [ln.rstrip("\r") for ln in buf[:nl].decode("ascii").split("\n")], nl + 1
        pos = nl + 1


def _vertex_layout(header, path):
    """(vertex count, property names) of a float-only, vertex-only header."""
    if _FORMAT not in header:
        raise ValueError(f"{path}: only binary_little_endian 1.0 PLY files are supported")
    count, names = None, []
    for line in header:
        words = line.split()
        kind = words[0] if words else ""
        if kind == "element":
            if count is not None or words[1:2] != ["vertex"] or len(words) != 3:
                raise ValueError(f"{path}: expected one vertex element, got {line!r}")
            count = int(words[2])
        elif kind == "property":
            if words[1:2] != ["float"] or len(words) != 3:
                raise ValueError(f"{path}: only float properties are supported, got {line!r}")
            names.append(words[2])
    if count is None:
        raise ValueError(f"{path}: PLY header has no vertex element")
    return count, names


def read_ply(path):
    """Read a splat PLY.  Returns (header lines without line ends, structured
    array with one little-endian float32 field per property)."""
    path = Path(path)
    with path.open("rb") as f, mmap.mmap(f.fileno(), 0, access=mmap.ACCESS_READ) as image:
        header, body = _split_header(image, path)
        count, names = _vertex_layout(header, path)
        record = np.dtype([(nm, "<f4") for nm in names])
        have = (len(image) - body) // record.itemsize if record.itemsize else 0
        if have < count:
            raise ValueError(f"{path}: expected {count} vertices, read {have}")
        data = np.frombuffer(image, dtype=record, count=count, offset=body).copy()
    return header, data


def write_ply(path, header, data):
    """Write ``data`` (structured array) under ``header`` (lines as read_ply
    returns them); the ``element vertex`` line takes the array's length."""
    lines = [f"element vertex {len(data)}" if ln.startswith("element vertex") else ln for ln in header]
    with Path(path).open("wb") as f:
        f.write("".join(ln + "\n" for ln in lines).encode("ascii"))
        f.write(np.ascontiguousarray(data).tobytes())


def vertex_header(names, count):
    """Header lines of a float-property vertex PLY (for new files)."""
    return (["ply", _FORMAT, f"element vertex {count}"]
            + [f"property float {nm}" for nm in names] + ["end_header"])
