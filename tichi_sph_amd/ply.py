"""PLY point files: `read_ply` / `write_ply` of sph_util.py:83-99.

The reference reads and writes PLY through the `plyfile` package, which is
not available here; this is a self-contained reader/writer of the subset the
reference uses: the `vertex` element's x, y, z, in ASCII or binary
(little/big endian) PLY, with any other elements and properties skipped.
`write_ply` writes ASCII PLY (the reference's text=True) with float32 x, y, z
per vertex, formatted with 9 significant digits so a read returns the
float32 values exactly.
"""
import numpy as np

_TYPES = {
    "char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1", "short": "i2", "int16": "i2",
    "ushort": "u2", "uint16": "u2", "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4",
    "float": "f4", "float32": "f4", "double": "f8", "float64": "f8",
}


def _parse_header(f):
    if f.readline().strip() != b"ply":
        raise ValueError("not a PLY file")
    fmt = None
    elements = []  # [name, count, [(prop, type) or (prop, ('list', count_t, item_t))]]
    while True:
        line = f.readline()
        if not line:
            raise ValueError("PLY header has no end_header")
        tok = line.decode("ascii", "replace").split()
        if not tok or tok[0] in ("comment", "obj_info"):
            continue
        if tok[0] == "format":
            fmt = tok[1]
        elif tok[0] == "element":
            elements.append([tok[1], int(tok[2]), []])
        elif tok[0] == "property":
            if tok[1] == "list":
                elements[-1][2].append((tok[4], ("list", _TYPES[tok[2]], _TYPES[tok[3]])))
            else:
                elements[-1][2].append((tok[2], _TYPES[tok[1]]))
        elif tok[0] == "end_header":
            break
    if fmt not in ("ascii", "binary_little_endian", "binary_big_endian"):
        raise ValueError(f"unsupported PLY format {fmt}")
    return fmt, elements


def read_ply(path):
    """Vertex positions (x, y, z) as an [N, 3] array of the file's type."""
    with open(path, "rb") as f:
        fmt, elements = _parse_header(f)
        body = f.read()
    if fmt == "ascii":
        lines = body.decode("ascii").split("\n")
        pos = 0
        for name, count, props in elements:
            rows = []
            for _ in range(count):
                while not lines[pos].strip():
                    pos += 1
                rows.append(lines[pos].split())
                pos += 1
            if name == "vertex":
                names = [p for p, _ in props]
                types = dict(props)
                cols = [names.index(c) for c in ("x", "y", "z")]
                dt = np.result_type(*[np.dtype(types[c]) for c in ("x", "y", "z")])
                return np.array([[float(r[c]) for c in cols] for r in rows], dtype=dt)
        raise ValueError("PLY file has no vertex element")
    end = "<" if fmt == "binary_little_endian" else ">"
    off = 0
    for name, count, props in elements:
        if any(isinstance(t, tuple) for _, t in props):
            if name == "vertex":
                raise ValueError("list properties in the vertex element are not supported")
            # variable-size rows: walk them
            for _ in range(count):
                for _, t in props:
                    if isinstance(t, tuple):
                        n = int(np.frombuffer(body, end + t[1], 1, off)[0])
                        off += np.dtype(t[1]).itemsize + n * np.dtype(t[2]).itemsize
                    else:
                        off += np.dtype(t).itemsize
            continue
        dt = np.dtype([(p, end + t) for p, t in props])
        if name == "vertex":
            v = np.frombuffer(body, dt, count, off)
            return np.stack([v["x"], v["y"], v["z"]], axis=1).astype(np.result_type(v["x"], v["y"], v["z"]))
        off += dt.itemsize * count
    raise ValueError("PLY file has no vertex element")


def write_ply(path, frame_num, dim, num, pos):
    """sph_util.py:90-99: the first `num` positions as vertices (z = 0 in
    2D) to `<path>_<frame_num>.ply`, ASCII."""
    pos = np.asarray(pos)
    if dim not in (2, 3):
        print("write_ply(): dim exceeds default values")
        return
    xyz = np.zeros((num, 3), dtype=np.float32)
    xyz[:, :dim] = pos[:num, :dim]
    fname = str(path) + "_" + str(frame_num) + ".ply"
    with open(fname, "w") as f:
        f.write(f"ply\nformat ascii 1.0\nelement vertex {num}\nproperty float x\nproperty float y\n"
                f"property float z\nend_header\n")
        for x, y, z in xyz:
            f.write(f"{x:.9g} {y:.9g} {z:.9g}\n")
    return fname
