"""Binary little-endian PLY reader/writer for float-only vertex properties
(the format of the reference's src/splat_io.py:9-59)."""
import numpy as np


def read_ply(path):
    """Return (property names, dict name -> float32 array) of the vertex element."""
    with open(path, "rb") as f:
        if f.readline().strip() != b"ply":
            raise ValueError(f"{path}: not a PLY file")
        fmt = f.readline().strip()
        if fmt != b"format binary_little_endian 1.0":
            raise ValueError(f"{path}: only binary_little_endian 1.0 is supported, got {fmt!r}")
        count, names, in_vertex = None, [], False
        while True:
            line = f.readline()
            if not line:
                raise ValueError(f"{path}: header has no end_header")
            tok = line.split()
            if not tok:
                continue
            if tok[0] == b"end_header":
                break
            if tok[0] == b"element":
                in_vertex = tok[1] == b"vertex"
                if in_vertex:
                    count = int(tok[2])
            elif tok[0] == b"property" and in_vertex:
                if tok[1] != b"float":
                    raise ValueError(f"{path}: only float vertex properties are supported")
                names.append(tok[2].decode())
        if count is None:
            raise ValueError(f"{path}: no vertex element")
        raw = np.frombuffer(f.read(4 * len(names) * count), dtype="<f4").reshape(count, len(names))
    return names, {n: np.ascontiguousarray(raw[:, k]) for k, n in enumerate(names)}


def write_ply(path, data, names=None):
    names = list(names or data.keys())
    n = len(data[names[0]])
    arr = np.stack([np.asarray(data[k], np.float32) for k in names], axis=1).astype("<f4")
    header = ["ply", "format binary_little_endian 1.0", f"element vertex {n}"]
    header += [f"property float {k}" for k in names] + ["end_header"]
    with open(path, "wb") as f:
        f.write(("\n".join(header) + "\n").encode())
        f.write(arr.tobytes())
