"""Binary little-endian PLY reader/writer for float-only vertex properties, with
the interface of the reference's src/splat_io.py:9-59: ``read_ply(path)``
returns ``(header lines, structured vertex array)`` and ``write_ply(path,
header, data)`` writes ``data`` under ``header`` with its vertex count
updated."""
from pathlib import Path

import numpy as np


def read_ply(path):
    """Read a binary_little_endian 1.0 PLY whose single element is ``vertex``
    with ``float`` properties.  Returns (header lines without line ends,
    structured array with one little-endian float32 field per property)."""
    path = Path(path)
    header = []
    with path.open("rb") as f:
        while True:
            line = f.readline()
            if not line:
                raise ValueError(f"{path}: PLY header has no end_header")
            text = line.decode("ascii").rstrip("\r\n")
            header.append(text)
            if text == "end_header":
                break
        body = f.tell()
    if "format binary_little_endian 1.0" not in header:
        raise ValueError(f"{path}: only binary_little_endian 1.0 PLY files are supported")
    count, names = None, []
    for text in header:
        tok = text.split()
        if not tok:
            continue
        if tok[0] == "element":
            if len(tok) != 3 or tok[1] != "vertex" or count is not None:
                raise ValueError(f"{path}: expected one vertex element, got {text!r}")
            count = int(tok[2])
        elif tok[0] == "property":
            if len(tok) != 3 or tok[1] != "float":
                raise ValueError(f"{path}: only float properties are supported, got {text!r}")
            names.append(tok[2])
    if count is None:
        raise ValueError(f"{path}: PLY header has no vertex element")
    dtype = np.dtype([(n, "<f4") for n in names])
    data = np.fromfile(path, dtype=dtype, count=count, offset=body)
    if len(data) != count:
        raise ValueError(f"{path}: expected {count} vertices, read {len(data)}")
    return header, data


def write_ply(path, header, data):
    """Write ``data`` (structured array) under ``header`` (lines as returned
    by read_ply); the ``element vertex`` line takes the array's length."""
    lines = [f"element vertex {len(data)}" if h.startswith("element vertex") else h for h in header]
    with Path(path).open("wb") as f:
        f.write(("\n".join(lines) + "\n").encode("ascii"))
        f.write(np.ascontiguousarray(data).tobytes())


def vertex_header(names, count):
    """Header lines of a float-property vertex PLY (for new files)."""
    return (["ply", "format binary_little_endian 1.0", f"element vertex {count}"]
            + [f"property float {n}" for n in names] + ["end_header"])
