"""PLY ingest (mh_ply_read) against files written here from known arrays.

The expected mesh follows from how the files are made: vertices as written (float32),
polygons as fans (first, previous, current), collapsed triangles removed by moving the
last triangle into their place (Mesh::removeDegenerateTriangles), the optional 3x4
transform applied in the reference's accumulation order. Host only.
"""

import struct

import numpy as np
import pytest

from rfmanta_amd import _lib as L
from rfmanta_amd import api

TYPES = {"char": "b", "uchar": "B", "short": "h", "ushort": "H", "int": "i", "uint": "I",
         "float": "f", "double": "d"}


def write_ply(path, verts, faces, fmt="ascii", vtype="float", ctype="uchar", itype="int",
              extra_vertex_prop=True, extra_element=True):
    lines = ["ply", f"format {fmt} 1.0", "comment made by the test"]
    lines += [f"element vertex {len(verts)}", f"property {vtype} x", f"property {vtype} y",
              f"property {vtype} z"]
    if extra_vertex_prop:
        lines.append("property uchar red")
    lines += [f"element face {len(faces)}", f"property list {ctype} {itype} vertex_indices"]
    if extra_element:
        lines += ["element edge 2", "property int vertex1", "property int vertex2"]
    lines.append("end_header")
    head = ("\n".join(lines) + "\n").encode()
    if fmt == "ascii":
        body = []
        for v in verts:
            conv = float if vtype in ("float", "double") else int
            body.append(" ".join(repr(conv(x)) for x in v) + (" 7" if extra_vertex_prop else ""))
        for f in faces:
            body.append(" ".join(str(x) for x in [len(f)] + list(f)))
        if extra_element:
            body += ["0 1", "1 2"]
        data = ("\n".join(body) + "\n").encode()
    else:
        e = "<" if fmt == "binary_little_endian" else ">"
        data = b""
        for v in verts:
            conv = float if vtype in ("float", "double") else int
            data += struct.pack(e + TYPES[vtype] * 3, *[conv(x) for x in v])
            if extra_vertex_prop:
                data += struct.pack(e + "B", 7)
        for f in faces:
            data += struct.pack(e + TYPES[ctype], len(f))
            data += struct.pack(e + TYPES[itype] * len(f), *f)
        if extra_element:
            data += struct.pack(e + "ii", 0, 1) + struct.pack(e + "ii", 1, 2)
    with open(path, "wb") as fh:
        fh.write(head + data)


def fan(faces):
    out = []
    for f in faces:
        for k in range(2, len(f)):
            out.append((f[0], f[k - 1], f[k]))
    return out


def remove_degenerate(v, tris):
    tris = list(tris)
    t = 0
    while t < len(tris):
        a, b, c = (v[j].astype(np.float32) for j in tris[t])
        cr = np.cross(a - b, c - b).astype(np.float32)
        if np.all(cr == 0):
            tris[t] = tris[-1]
            tris.pop()
            continue
        t += 1
    return tris


@pytest.mark.parametrize("fmt", ["ascii", "binary_little_endian", "binary_big_endian"])
@pytest.mark.parametrize("vtype,ctype,itype", [("float", "uchar", "int"),
                                               ("double", "uint", "uint"),
                                               ("short", "uchar", "ushort")])
def test_ply_round_trip(tmp_path, fmt, vtype, ctype, itype):
    rng = np.random.default_rng(1)
    if vtype == "short":
        v = rng.integers(-200, 200, size=(40, 3)).astype(np.float64)
    else:
        v = rng.normal(size=(40, 3))
    faces = [list(rng.choice(40, size=int(k), replace=False)) for k in rng.integers(3, 7, 30)]
    faces.append([0, 1])  # fewer than three vertices: no triangle
    v[5] = v[6]  # two equal vertices: faces through both collapse
    faces += [[5, 6, 7], [6, 9, 5, 11]]
    p = tmp_path / "m.ply"
    write_ply(p, v, faces, fmt, vtype, ctype, itype)
    gv, gi = api.read_ply(p)
    v32 = v.astype(np.float32)
    assert np.array_equal(gv, v32)
    want = remove_degenerate(v32, fan(faces))
    assert [tuple(t) for t in gi.tolist()] == want
    assert len(want) < len(fan(faces))  # some were removed


def test_ply_transform(tmp_path):
    rng = np.random.default_rng(2)
    v = rng.normal(size=(10, 3)).astype(np.float32)
    p = tmp_path / "t.ply"
    write_ply(p, v, [[0, 1, 2], [3, 4, 5, 6]])
    m = rng.normal(size=(3, 4)).astype(np.float32)
    gv, _ = api.read_ply(p, m)
    want = np.empty_like(v)
    for a in range(3):
        t = m[a, 0] * v[:, 0]
        t = (t + m[a, 1] * v[:, 1]).astype(np.float32)
        t = (t + m[a, 2] * v[:, 2]).astype(np.float32)
        want[:, a] = (t + m[a, 3]).astype(np.float32)
    assert np.array_equal(gv, want)


def test_ply_without_faces_drops_last_vertex(tmp_path):
    v = np.arange(12, dtype=np.float32).reshape(4, 3)
    p = tmp_path / "nf.ply"
    write_ply(p, v, [])
    gv, gi = api.read_ply(p)
    assert len(gv) == 3 and len(gi) == 0


def test_ply_errors(tmp_path):
    with pytest.raises(L.MhError):
        api.read_ply(tmp_path / "missing.ply")
    p = tmp_path / "bad.ply"
    p.write_bytes(b"not a ply\n")
    with pytest.raises(L.MhError):
        api.read_ply(p)
    q = tmp_path / "range.ply"
    write_ply(q, np.zeros((3, 3)), [[0, 1, 7]])
    with pytest.raises(L.MhError):
        api.read_ply(q)
