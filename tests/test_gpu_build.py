"""Device build (mh_scene_build with approximate = 1): DynBVH::rebuild with the
approximate SAH at every node, against the oracle's restatement record for record.

The exported tree (ids and every record in use, isLeftCheaper and isLargeSubtree bits
included) must equal the oracle's approximate build bit for bit, unused records must be
blank, and traces over it must equal the oracle's traces over its own tree. Meshes cover
the phase boundaries (one node, nodes just above and below the one-lane size), a soup, a
heightfield, coincident triangles (a large leaf), both signs of zero, and a split that
falls back to the middle. At 1M triangles the checks are the tree's own invariants.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from rfmanta_amd import api, scenes  # noqa: E402


def _meshes():
    out = []
    for n in (1, 2, 3, 64, 65, 128, 129, 300):
        out.append((f"soup{n}",) + scenes.soup(n, seed=n))
    out.append(("soup20000",) + scenes.soup(20000, seed=3))
    out.append(("heightfield",) + scenes.heightfield(150, 90, seed=1))
    t = np.array([[0, 0, 0], [0.1, 0, 0], [0, 0.1, 0]], np.float32)
    vd = np.tile(t, (700, 1))
    idd = np.arange(len(vd), dtype=np.uint32).reshape(-1, 3)
    vs, is_ = scenes.soup(1500, seed=9)
    out.append(("duplicates", np.concatenate([vd, vs]), np.concatenate([idd, is_ + len(vd)])))
    v, i = scenes.soup(3000, seed=5)
    v = v.copy()
    rng = np.random.default_rng(1)
    m = rng.random(v.shape) < 0.2
    v[m] = np.where(rng.random(m.sum()) < 0.5, np.float32(0.0), np.float32(-0.0))
    out.append(("signed_zeros", v, i))
    from test_oracle_approx_build import _ulp_cluster
    v0, _ = _ulp_cluster()
    v = np.tile(v0, (30, 1))
    out.append(("middle_fallback", v, np.arange(len(v), dtype=np.uint32).reshape(-1, 3)))
    return out


@pytest.mark.parametrize("name,v,i", _meshes(), ids=lambda x: x if isinstance(x, str) else "")
def test_device_build_matches_oracle(oracle, name, v, i):
    print(f"device build {name} T={len(i)}", flush=True)
    on, oi, onn = oracle.build_bvh_approx(v, i)
    with api.Scene(v, i) as s:
        s.build(approximate=True)
        gi, gn = s.export_dynbvh()
        assert s.stats()["nodes"] == onn
        assert np.array_equal(gi, oi)
        a = gn[:onn].view(np.uint8).reshape(onn, 32)
        b = on[:onn].view(np.uint8).reshape(onn, 32)
        bad = np.nonzero((a != b).any(1))[0]
        assert len(bad) == 0, (name, bad[:5], gn[bad[:3]], on[bad[:3]])
        assert gn[onn:].tobytes() == on[onn:].tobytes()
        rays = scenes.random_rays(4096, 11, (-1.3, -1.3, -1.3), (1.3, 1.3, 1.3))
        g = s.trace_host(rays)
        o = oracle.trace(v, i, on, oi, rays, packet_size=1)
        assert np.array_equal(g["prim"], o["prim"])
        assert np.array_equal(g["t"].view(np.uint32), o["t"].view(np.uint32))
        sh = s.trace_host(rays, any_hit=True)
        osh = oracle.trace(v, i, on, oi, rays, any_hit=True, packet_size=1)
        assert np.array_equal(sh["occluded"], osh["occluded"])
        assert np.array_equal(sh["prim"], osh["prim"])


def test_device_build_then_update(oracle):
    """The rotating update starts from the device build's costs and sizes."""
    v, i = scenes.soup(4000, seed=12)
    on, oi, onn = oracle.build_bvh_approx(v, i)
    st = oracle.RotState(on, oi, onn)
    with api.Scene(v, i) as s:
        s.build(approximate=True)
        for f in range(3):
            vd = v.copy()
            vd[:, 1] += (0.3 * np.sin(4 * v[:, 0] + 0.9 * f)).astype(np.float32)
            st.update(vd, i)
            s.set_vertices(vd)
            s.update()
            api.sync()
            gi, gn = s.export_dynbvh()
            from test_gpu_update import canonical_equal
            ok, msg, count = canonical_equal(gn, gi, st.nodes, st.ids)
            assert ok, (f, msg)


def test_rebuild_between_updates(oracle):
    """Updates after a host build, then a device rebuild of the displaced mesh, then
    updates again: every tree equals the oracle's (the update's per-slot leaf map and level
    lists are set up again for the new tree)."""
    from test_gpu_update import canonical_equal
    v, i = scenes.soup(3000, seed=21)

    def frame(f):
        vd = v.copy()
        vd[:, 1] += (0.3 * np.sin(4 * v[:, 0] + 0.7 * f)).astype(np.float32)
        return vd

    with api.Scene(v, i) as s:
        s.build()
        ids, nodes = s.export_dynbvh()
        st = oracle.RotState(nodes, ids, s.stats()["nodes"])
        for f in range(2):
            vd = frame(f)
            st.update(vd, i)
            s.set_vertices(vd)
            s.update()
            api.sync()
            gi, gn = s.export_dynbvh()
            ok, msg, _ = canonical_equal(gn, gi, st.nodes, st.ids)
            assert ok, ("host build", f, msg)
        vd = frame(2)
        s.set_vertices(vd)
        s.build(approximate=True)
        on, oi, onn = oracle.build_bvh_approx(vd, i)
        st = oracle.RotState(on, oi, onn)
        for f in range(3, 5):
            vd = frame(f)
            st.update(vd, i)
            s.set_vertices(vd)
            s.update()
            api.sync()
            gi, gn = s.export_dynbvh()
            ok, msg, _ = canonical_equal(gn, gi, st.nodes, st.ids)
            assert ok, ("device rebuild", f, msg)


def test_device_build_1m_invariants():
    """997,002-triangle heightfield: boxes are exact unions, every object in one sorted
    leaf, every record reachable, and the closest hits of the full-SAH tree and the
    approximate tree agree away from ties."""
    v, i = scenes.heightfield(1000, 500, seed=0)
    T = len(i)
    print(f"device build 1M T={T}", flush=True)
    with api.Scene(v, i) as s:
        s.build(approximate=True)
        ids, nodes = s.export_dynbvh()
        nn = s.stats()["nodes"]
        rays = scenes.random_rays(1 << 16, 21, (-1.0, 0.3, -1.0), (1.0, 0.6, 1.0))
        ga = s.trace_host(rays)
    assert np.array_equal(np.sort(ids), np.arange(T))
    tri = v[i]
    tlo, thi = tri.min(1), tri.max(1)
    # walk: children boxes inside parents, leaf boxes equal to the union of their triangles
    child = nodes["child"][:nn]
    cnt = nodes["children"][:nn]
    inner = np.nonzero(cnt == 0)[0]
    leaf = np.nonzero(cnt > 0)[0]
    assert len(inner) + len(leaf) == nn and len(leaf) == len(inner) + 1
    c = child[inner]
    assert np.all(c % 2 == 1) and np.all(c + 1 < nn)
    lo, hi = nodes["lo"], nodes["hi"]
    u_lo = np.minimum(lo[c], lo[c + 1])
    u_hi = np.maximum(hi[c], hi[c + 1])
    assert np.array_equal(u_lo[inner != 0], lo[inner][inner != 0])
    assert np.array_equal(u_hi[inner != 0], hi[inner][inner != 0])
    order = np.argsort(child[leaf])
    starts = child[leaf][order].astype(np.int64)
    counts = cnt[leaf][order].astype(np.int64)
    assert starts[0] == 0 and np.array_equal(starts[1:], np.cumsum(counts)[:-1])
    assert starts[-1] + counts[-1] == T  # leaves partition the slots
    inside = np.ones(T - 1, bool)
    inside[starts[1:] - 1] = False
    assert np.all(np.diff(ids)[inside] > 0)  # leaf ids sorted
    slo, shi = tlo[ids], thi[ids]
    assert np.array_equal(np.minimum.reduceat(slo, starts, axis=0), lo[leaf[order]])
    assert np.array_equal(np.maximum.reduceat(shi, starts, axis=0), hi[leaf[order]])
    with api.Scene(v, i) as s2:
        s2.build()
        gf = s2.trace_host(rays)
    same = ga["prim"] == gf["prim"]
    close = np.abs(ga["t"] - gf["t"]) <= 1e-4
    assert np.all(same | close)
