"""The reference's update with several render threads, restated in the oracle.

DynBVH::parallelUpdateBounds with numProcs > 1 (DynBVH.cc:2034-2065) updates the subtrees
seedSubtrees collects (the maximal ones not marked isLargeSubtree) first and then, from the
root, the large nodes above them (updateBounds<true>); largeSubtreeSize is
records / (20 * numProcs) (DynBVH.cc:945, 1776). Each node is still refitted and rotated
exactly once, after its children, and rotateNode reads and writes only its own subtree, so
the tree is the one-thread tree: these tests show it frame after frame, with rotations and
with leaf splits, and pin the isLargeSubtree bits to the subtree sizes.
"""

import numpy as np
import pytest

from rfmanta_amd import scenes


def canonical(ga, gi, oa, oi, bits_mask):
    g = np.zeros(1, np.int64)
    o = np.zeros(1, np.int64)
    n = 0
    while len(g):
        n += len(g)
        for f in ("lo", "hi", "axis", "children"):
            if ga[f][g].tobytes() != oa[f][o].tobytes():
                return False, f, n
        if (ga["bits"][g] & bits_mask).tobytes() != (oa["bits"][o] & bits_mask).tobytes():
            return False, "bits", n
        c = ga["children"][g].astype(np.int64)
        leaf = c > 0
        if leaf.any():
            cnt = c[leaf]
            rel = np.arange(int(cnt.sum())) - np.repeat(np.cumsum(cnt) - cnt, cnt)
            sg = np.repeat(ga["child"][g[leaf]].astype(np.int64), cnt) + rel
            so = np.repeat(oa["child"][o[leaf]].astype(np.int64), cnt) + rel
            if not np.array_equal(gi[sg], oi[so]):
                return False, "ids", n
        gc = ga["child"][g[~leaf]].astype(np.int64)
        oc = oa["child"][o[~leaf]].astype(np.int64)
        g = np.concatenate([gc, gc + 1])
        o = np.concatenate([oc, oc + 1])
    return True, "", n


def sizes_and_bits(nodes, limit):
    """Subtree record counts from the root; True when every reachable record's bit 1 says
    size > limit (leaves never)."""
    ok = True
    order = []
    stack = [0]
    while stack:
        r = stack.pop()
        order.append(r)
        if nodes["children"][r] == 0:
            stack += [int(nodes["child"][r]), int(nodes["child"][r]) + 1]
    size = {}
    for r in reversed(order):
        if nodes["children"][r]:
            size[r] = 1
            ok &= (nodes["bits"][r] & 2) == 0
        else:
            c = int(nodes["child"][r])
            size[r] = size[c] + size[c + 1] + 1
            ok &= bool(nodes["bits"][r] & 2) == (size[r] > limit)
    return ok


FRAMES = {
    # heightfield, smooth displacement: rotations every frame, no leaf splits
    "heightfield": lambda: (scenes.heightfield(160, 80, seed=5),
                            lambda v, f: np.column_stack(
                                [v[:, 0], v[:, 1] + np.float32(0.05) *
                                 np.sin(np.float32(4.0) * v[:, 0] + np.float32(0.7 * f)),
                                 v[:, 2]]).astype(np.float32)),
    # soup, growing random motion: leaves split
    "soup_splits": lambda: (scenes.soup(6000, seed=7, sigma=0.05),
                            lambda v, f: (v + np.random.default_rng(f).normal(
                                0, 0.3, v.shape).astype(np.float32) * (f + 1) / 5)
                            .astype(np.float32)),
}


@pytest.mark.parametrize("mesh", sorted(FRAMES))
@pytest.mark.parametrize("workers", [2, 5, 64])
def test_worker_schedule_builds_the_one_thread_tree(oracle, mesh, workers):
    (v, i), move = FRAMES[mesh]()
    nodes, ids, used = oracle.build_bvh(v, i)
    one = oracle.RotState(nodes, ids, used)
    many_nodes = nodes.copy()
    oracle.mark_large(many_nodes, workers)
    limit = len(nodes) // (20 * workers)
    assert sizes_and_bits(many_nodes, limit)
    assert (many_nodes["bits"] & 2).sum() > (nodes["bits"] & 2).sum()  # more seeds
    many = oracle.RotState(many_nodes, ids, used)
    changed = 0
    for f in range(5):
        vd = move(v, f)
        before = one.nodes.copy()
        one.update(vd, i)
        many.update(vd, i, workers=workers)
        ok, what, n = canonical(one.nodes, one.ids, many.nodes, many.ids, 0xFD)
        assert ok, (f, what)
        assert one.next_free.value == many.next_free.value
        changed += not np.array_equal(before["child"], one.nodes["child"])
    assert changed >= 3  # the frames do rotate the tree
    if mesh == "soup_splits":
        assert one.next_free.value > used  # and split leaves


def test_mark_large_plants_the_threshold(oracle):
    """A threshold off by a factor (20 instead of 20 * workers) moves bits: caught."""
    v, i = scenes.heightfield(120, 60, seed=2)
    nodes, _, _ = oracle.build_bvh(v, i)
    for workers in (1, 3, 8):
        x = nodes.copy()
        oracle.mark_large(x, workers)
        assert sizes_and_bits(x, len(x) // (20 * workers))
        if workers > 1:
            assert not sizes_and_bits(x, len(x) // 20)
