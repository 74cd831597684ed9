"""mh_shade (hit positions, normals, texture coordinates on the device) against the
oracle's restatement, bit for bit, on the hits of a device trace."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from rfmanta_amd import _lib as L  # noqa: E402
from rfmanta_amd import api, scenes  # noqa: E402

KEYS = ("px", "py", "pz", "nx", "ny", "nz", "gx", "gy", "gz", "tu", "tv", "tw")


@pytest.mark.parametrize("kind", [L.MH_PRIM_KENSLER_SHIRLEY, L.MH_PRIM_WALD])
@pytest.mark.parametrize("attrs", ["none", "normals", "texcoords", "both"])
def test_shade_matches_oracle(oracle, kind, attrs):
    print(f"shade kind={kind} attrs={attrs}", flush=True)
    v, i = scenes.config0_scene()
    rays = scenes.config0_camera(128, 128)
    rng = np.random.default_rng(3)
    normals = nidx = tex = tidx = None
    if attrs in ("normals", "both"):
        normals = rng.normal(size=(len(v) + 7, 3)).astype(np.float32)
        nidx = rng.integers(0, len(normals), size=i.shape).astype(np.uint32)
    if attrs in ("texcoords", "both"):
        tex = rng.random((len(v) // 2, 3)).astype(np.float32)
        tidx = rng.integers(0, len(tex), size=i.shape).astype(np.uint32)
        tidx[::5, 0] = 0xFFFFFFFF  # Mesh::kNoTextureIndex: these triangles have none
    with api.Scene(v, i) as s:
        s.build(prim_kind=kind)
        s.set_attributes(normals, nidx, tex, tidx)
        dr = api.DeviceRays.upload(rays)
        dh = api.DeviceHits.alloc(dr.n)
        s.trace(dr, dh)
        api.sync()
        h = dh.download()
        g = s.shade(dr, dh)
    assert (h["prim"] >= 0).sum() > 1000
    o = oracle.shade(v, i, rays, h, prim_kind=kind, normals=normals, normal_idx=nidx, tex=tex,
                     tex_idx=tidx)
    for k, key in enumerate(KEYS):
        assert np.array_equal(g[key].view(np.uint32), o[:, k].view(np.uint32)), key


def test_obj_mesh_attributes_shade_like_oracle(oracle, tmp_path):
    """A mesh read by mh_obj_read (per-corner normals and texture coordinates, some
    corners without) feeds mh_scene_set_attributes directly."""
    print("shade obj mesh", flush=True)
    v, i = scenes.config0_scene()
    rng = np.random.default_rng(5)
    vn = rng.normal(size=(500, 3)).astype(np.float32)
    vt = rng.random((300, 2)).astype(np.float32)
    lines = [f"v {a!r} {b!r} {c!r}" for a, b, c in v.tolist()]
    lines += [f"vn {a!r} {b!r} {c!r}" for a, b, c in vn.tolist()]
    lines += [f"vt {a!r} {b!r}" for a, b in vt.tolist()]
    for t, tri in enumerate(i.tolist()):
        if t % 7 == 3:  # corners without normals or texture coordinates
            lines.append("f " + " ".join(str(x + 1) for x in tri))
        else:
            lines.append("f " + " ".join(f"{x + 1}/{rng.integers(1, 301)}/{rng.integers(1, 501)}"
                                         for x in tri))
    (tmp_path / "c0.obj").write_text("\n".join(lines) + "\n")
    m = api.read_obj(tmp_path / "c0.obj")
    assert m["verts"].shape == v.shape and len(m["normals"]) > len(vn)
    rays = scenes.config0_camera(128, 128)
    with api.Scene(m["verts"], m["idx"]) as s:
        s.build()
        s.set_attributes(m["normals"], m["normal_idx"], m["texcoords"], m["tex_idx"])
        dr = api.DeviceRays.upload(rays)
        dh = api.DeviceHits.alloc(dr.n)
        s.trace(dr, dh)
        api.sync()
        h = dh.download()
        g = s.shade(dr, dh)
    assert (h["prim"] >= 0).sum() > 1000
    o = oracle.shade(m["verts"], m["idx"], rays, h, prim_kind=L.MH_PRIM_KENSLER_SHIRLEY,
                     normals=m["normals"], normal_idx=m["normal_idx"], tex=m["texcoords"],
                     tex_idx=m["tex_idx"])
    for k, key in enumerate(KEYS):
        assert np.array_equal(g[key].view(np.uint32), o[:, k].view(np.uint32)), key
