"""OBJ ingest (mh_obj_read) against files written here.

The small files are worked by hand through the rules of glmReadOBJ and ObjGroup
(Model/Readers/glm/glm.cc:575-960, Model/Groups/ObjGroup.cc:184-316): groups named
"<group>_MAT_<material>", emitted in reverse order of creation, faces before the first
group statement going to the group created last, material slots 1 + glm's index (0 when
there is no material library), face normals appended for corners without a normal,
MH_NO_TEXTURE_INDEX for corners without texture coordinates, degenerate triangles replaced
by the last one (Mesh::removeDegenerateTriangles). The expected arrays are written out as
literals. Host only.
"""

import numpy as np
import pytest

from rfmanta_amd import _lib as L
from rfmanta_amd import api

NT = 0xFFFFFFFF  # Mesh::kNoTextureIndex


def test_groups_and_material_slots(tmp_path):
    (tmp_path / "m.mtl").write_text("# two materials\nnewmtl green\nKd 0 1 0\n\n"
                                    "newmtl red\nKd 1 0 0\nNs 10\n")
    verts = "\n".join(f"v {k} {k * k % 7} {k % 3}" for k in range(1, 10))
    (tmp_path / "s.obj").write_text(f"""mtllib m.mtl
{verts}
f 1 2 3
g A
f 2 3 4
usemtl red
f 3 4 5
g B
f 4 5 6
usemtl nothere
f 5 6 7
g A
f 6 7 8
usemtl red
f 7 8 9
""")
    m = api.read_obj(tmp_path / "s.obj")
    # Faces T0..T6 are (k, k+1, k+2) zero-based. Groups by creation: No Group, A_MAT_No
    # Material {T1}, A_MAT_red {T2, T6}, B_MAT_red {T3}, B_MAT_nothere {T4},
    # A_MAT_nothere {T0 (early), T5}; emitted last-created first.
    order = [0, 5, 4, 3, 2, 6, 1]
    assert m["idx"].tolist() == [[k, k + 1, k + 2] for k in order]
    # glm materials: 0 NO_ASSIGNED_MATERIAL, 1 green, 2 red; unknown names resolve to 0.
    assert m["face_material"].tolist() == [1, 1, 1, 3, 3, 3, 1]
    assert m["material_names"] == ["", "NO_ASSIGNED_MATERIAL", "green", "red"]
    assert m["normals"] is None and m["texcoords"] is None
    assert m["verts"].shape == (9, 3) and m["verts"][4].tolist() == [5.0, 4.0, 2.0]


def test_without_material_library_every_slot_is_default(tmp_path):
    (tmp_path / "s.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\n"
                                    "usemtl red\nf 1 2 3\ng X\nf 2 4 3\n")
    m = api.read_obj(tmp_path / "s.obj")
    assert m["idx"].tolist() == [[1, 3, 2], [0, 1, 2]]  # X_MAT_red first
    assert m["face_material"].tolist() == [0, 0]
    assert m["material_names"] == [""]


def test_corner_forms_normals_texcoords_and_degenerate_removal(tmp_path):
    (tmp_path / "n.obj").write_text("""# corner forms
v 0 0 0
v 1 0 0
v 0 1 0
v 0 0 1
v 1 1 1
vn 0 0 1
vn 1 0 0
vt 0.25 0.5
vt 0.75 0.125 0.0
s off
f 1//2 1//1 2//2
f 1//1 2//2 3//1
f 1 2 4
f 2/1/2 3/2/1 5/1/2 4/2/1
f 1/2 3/1 5/-1
""")
    m = api.read_obj(tmp_path / "n.obj")
    # T0 (collapsed: corners 1 and 2 are the same vertex) is replaced by T5, the last.
    assert m["idx"].tolist() == [[0, 2, 4], [0, 1, 2], [0, 1, 3], [1, 2, 4], [1, 4, 3]]
    assert m["normal_idx"].tolist() == [[5, 6, 7], [0, 1, 0], [2, 3, 4], [1, 0, 1], [1, 1, 0]]
    assert m["tex_idx"].tolist() == [[1, 0, NT], [NT] * 3, [NT] * 3, [0, 1, 0], [0, 0, 1]]
    # Appended face normals Cross(v1 - v0, v2 - v0): (0,-1,0) for (0,1,3), (1,0,-1) for
    # (0,2,4), once per corner without a normal.
    assert m["normals"].tolist() == [[0, 0, 1], [1, 0, 0]] + [[0, -1, 0]] * 3 + [[1, 0, -1]] * 3
    assert m["texcoords"].tolist() == [[0.25, 0.5, 1.0], [0.75, 0.125, 1.0]]
    assert m["face_material"].tolist() == [0] * 5


def test_polygons_negative_indices_and_line_endings(tmp_path):
    rng = np.random.default_rng(4)
    v = rng.normal(size=(30, 3)).astype(np.float32)
    lines = [f"v {a!r} {b!r} {c!r}" for a, b, c in v.tolist()]
    want = []
    for f in range(25):
        k = int(rng.integers(3, 8))
        corners = rng.choice(30, size=k, replace=False)
        toks = [str(c + 1) if rng.random() < 0.5 else str(c - 30) for c in corners]
        lines.append("f " + " ".join(toks) + ("  # trailing comment" if f % 4 == 0 else ""))
        want += [[int(corners[0]), int(corners[j - 1]), int(corners[j])] for j in range(2, k)]
    (tmp_path / "p.obj").write_bytes(("\r\n".join(lines) + "\r\n").encode())
    m = api.read_obj(tmp_path / "p.obj")
    assert np.array_equal(m["verts"], v)
    assert m["idx"].tolist() == want  # random corners of a normal cloud: none collapse


def test_obj_errors(tmp_path):
    with pytest.raises(L.MhError):
        api.read_obj(tmp_path / "missing.obj")
    cases = {
        "nomtl.obj": "mtllib absent.mtl\nv 0 0 0\n",
        "range.obj": "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n",
        "zero.obj": "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n",
        "mixed.obj": "v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1//1 2//1 3\n",
        "short.obj": "v 0 0 0\nv 1 0 0\nf 1 2\n",
        "normal.obj": "v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1//1 2//1 3//2\n",
    }
    for name, text in cases.items():
        (tmp_path / name).write_text(text)
        with pytest.raises(L.MhError):
            api.read_obj(tmp_path / name)


def test_obj_and_ply_of_one_mesh_agree(tmp_path):
    """The same polygons written as OBJ (no groups) and as PLY read back to the same Mesh:
    both readers fan polygons the same way and remove the same collapsed triangles."""
    import struct
    rng = np.random.default_rng(6)
    v = rng.normal(size=(50, 3)).astype(np.float32)
    v[7] = v[8]
    faces = [list(rng.choice(50, size=int(k), replace=False)) for k in rng.integers(3, 7, 40)]
    faces += [[7, 8, 9], [9, 8, 7, 10]]
    obj = [f"v {a!r} {b!r} {c!r}" for a, b, c in v.tolist()]
    obj += ["f " + " ".join(str(x + 1) for x in f) for f in faces]
    (tmp_path / "m.obj").write_text("\n".join(obj) + "\n")
    head = (f"ply\nformat binary_little_endian 1.0\nelement vertex {len(v)}\nproperty float x\n"
            f"property float y\nproperty float z\nelement face {len(faces)}\n"
            "property list uchar int vertex_indices\nend_header\n").encode()
    body = b"".join(struct.pack("<3f", *p) for p in v.tolist())
    body += b"".join(struct.pack("<B", len(f)) + struct.pack(f"<{len(f)}i", *f) for f in faces)
    (tmp_path / "m.ply").write_bytes(head + body)
    m = api.read_obj(tmp_path / "m.obj")
    pv, pi = api.read_ply(tmp_path / "m.ply")
    assert np.array_equal(m["verts"], pv)
    assert np.array_equal(m["idx"], pi)
    assert len(pi) < sum(len(f) - 2 for f in faces)  # collapsed triangles were removed
