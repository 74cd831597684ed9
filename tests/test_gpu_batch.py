"""Persistent multi-batch frames: parity with the plain single-ray path and the oracle,
HIP-graph capture and replay, batch-table rewrites, and the autotuner's cache."""

import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from rfmanta_amd import api, scenes  # noqa: E402


def _bits(a):
    return np.asarray(a).view(np.uint32)


@pytest.fixture(scope="module")
def small():
    """A heightfield with a mixed set of uneven batches (all three ray kinds)."""
    v, i = scenes.heightfield(160, 80, seed=9)
    sizes = [1, 63, 64, 65, 1000, 4097, 20000, 7]
    kinds = ["primary", "shadow", "bounce", "primary", "bounce", "shadow", "primary", "bounce"]
    batches = []
    start = 0
    for k, n in zip(kinds, sizes):
        rays = scenes.batch_rays(k, n, v, i, 3, start, eye=(0, 0.9, -1.1))
        if k == "shadow":
            rays["masked"] = (np.arange(n) % 11 == 0).astype(np.uint8)
        batches.append((k, rays))
        start += n
    return v, i, batches


def _device(batches):
    out = []
    for k, rays in batches:
        any_hit = k == "shadow"
        out.append((api.DeviceRays.upload(rays), api.DeviceHits.alloc(len(rays["ox"]), any_hit),
                    any_hit))
    return out


@pytest.mark.parametrize("cfg", [
    # per-lane refill, pools of 64 and 256 rays
    {"layout": 0, "stack_lds": 8, "block": 256, "blocks_per_cu": 8, "refill": 16, "chunk": 64},
    {"layout": 1, "stack_lds": 8, "block": 64, "blocks_per_cu": 4, "refill": 1, "chunk": 256},
    # whole-wave chunks from the counter, with and without the XCD remap
    {"layout": 0, "stack_lds": 32, "block": 128, "blocks_per_cu": 2, "refill": 64, "chunk": 64},
    {"layout": 0, "stack_lds": 8, "block": 256, "blocks_per_cu": 3, "refill": 64, "chunk": 1024,
     "xcd_remap": 1},
    {"layout": 1, "stack_lds": 16, "block": 256, "blocks_per_cu": 1, "refill": 32, "chunk": 128,
     "xcd_remap": 1},
    # static stride, no counter
    {"layout": 0, "stack_lds": 16, "block": 64, "blocks_per_cu": 8, "refill": 64, "chunk": 0,
     "xcd_remap": 1},
    {"layout": 1, "stack_lds": 8, "block": 128, "blocks_per_cu": 2, "refill": 64, "chunk": 0},
    # the PairNodes as four field arrays: static stride, counter chunks, per-lane refill
    {"layout": 2, "stack_lds": 8, "block": 256, "blocks_per_cu": 4, "refill": 64, "chunk": 0},
    {"layout": 2, "stack_lds": 16, "block": 128, "blocks_per_cu": 2, "refill": 64, "chunk": 256},
    {"layout": 2, "stack_lds": 8, "block": 256, "blocks_per_cu": 5, "refill": 32, "chunk": 128,
     "xcd_remap": 1},
    # handed out from the end of the table: per-lane refill (also per XCD), counter chunks,
    # static stride (the last chunk in hand-out order is the short one at ray 0)
    {"layout": 0, "stack_lds": 8, "block": 128, "blocks_per_cu": 6, "refill": 32, "chunk": 256,
     "reverse": 1},
    {"layout": 0, "stack_lds": 8, "block": 256, "blocks_per_cu": 3, "refill": 16, "chunk": 192,
     "xcd_remap": 1, "reverse": 1},
    {"layout": 0, "stack_lds": 16, "block": 128, "blocks_per_cu": 2, "refill": 64, "chunk": 320,
     "reverse": 1},
    {"layout": 1, "stack_lds": 8, "block": 256, "blocks_per_cu": 2, "refill": 64, "chunk": 0,
     "reverse": 1},
])
def test_frame_matches_oracle(oracle, small, cfg):
    v, i, batches = small
    nodes, ids, _ = oracle.build_bvh(v, i)
    with api.Scene(v, i) as s:
        s.load_dynbvh(ids, nodes)
        dev = _device(batches)
        f = api.Frame(s, dev)
        f.set_config(cfg)
        full = {"chunk": 64, "xcd_remap": 0, "reverse": 0, **cfg}
        assert f.config() == full
        f.launch()
        f.launch()  # a second launch checks that the counters reset themselves
        api.sync()
        for (k, rays), (dr, dh, any_hit) in zip(batches, dev):
            g = dh.download()
            o = oracle.trace(v, i, nodes, ids, rays, any_hit=any_hit, packet_size=1)
            assert np.array_equal(g["prim"], o["prim"]), (k, len(rays["ox"]))
            assert np.array_equal(_bits(g["t"]), _bits(o["t"])), k
            if any_hit:
                assert np.array_equal(g["occluded"], o["occluded"])
            else:
                assert np.array_equal(_bits(g["u"]), _bits(o["u"]))
                assert np.array_equal(_bits(g["v"]), _bits(o["v"]))
        f.close()


LANE_CFG = dict(layout=0, stack_lds=8, block=64, blocks_per_cu=24, refill=32, chunk=64,
                xcd_remap=0)


@pytest.mark.parametrize("cfg", [None, LANE_CFG], ids=["default", "lane_refill"])
def test_graph_replays_and_table_rewrite(small, cfg):
    """Replays of a captured frame (whole-wave and per-lane refill schedules, whose work
    counters reset themselves) equal the eager results, with eager work in between."""
    print(f"graph replays cfg={cfg}", flush=True)
    v, i, batches = small
    with api.Scene(v, i) as s:
        s.build()
        dev = _device(batches)
        # eager results of the plain path, per batch
        ref = []
        for dr, dh, any_hit in dev:
            s.trace(dr, dh, any_hit=any_hit)
            api.sync()
            ref.append(dh.download())
        stream = api.stream_create()
        try:
            f = api.Frame(s, dev)
            if cfg is not None:
                f.set_config(cfg)
            f.capture(stream)
            for rep in range(3):
                for dr, dh, any_hit in dev:  # scribble over the outputs between replays
                    dh.arrays["prim"].upload(np.full(dh.n, 777, np.int32))
                f.replay(stream)
                api.sync(stream)
                for (dr, dh, any_hit), r in zip(dev, ref):
                    got = dh.download()
                    assert np.array_equal(got["prim"], r["prim"]), rep
                    assert np.array_equal(_bits(got["t"]), _bits(r["t"])), rep
                # eager work and a plain launch between replays
                dr0, dh0, a0 = dev[4]
                s.trace(dr0, dh0, any_hit=a0)
                api.sync()
            # rewrite the table: the same batches in reverse order, replayed with no re-capture
            f.set_batches(list(reversed(dev)), stream)
            for dr, dh, any_hit in dev:
                dh.arrays["prim"].upload(np.full(dh.n, 777, np.int32))
            f.replay(stream)
            api.sync(stream)
            for (dr, dh, any_hit), r in zip(dev, ref):
                assert np.array_equal(dh.download()["prim"], r["prim"])
            f.close()
        finally:
            api.stream_destroy(stream)


def test_autotune_cache(tmp_path, small):
    v, i, batches = small
    with api.Scene(v, i) as s:
        s.build()
        dev = _device(batches)
        f = api.Frame(s, dev)
        path = tmp_path / "tune.json"
        best, ms = f.autotune(iters=1, cache_path=str(path))
        assert ms > 0 and path.exists()
        cache = json.loads(path.read_text())
        assert len(cache) == 1
        entry = next(iter(cache.values()))
        assert {k: entry[k] for k in best} == best
        best2, ms2 = f.autotune(iters=1, cache_path=str(path))  # served from the cache
        assert best2 == best and abs(ms2 - entry["ms"]) < 1e-3
        assert f.config() == best
        f.launch()
        api.sync()
        f.close()


@pytest.mark.parametrize("cfg", [
    {"layout": 0, "stack_lds": 8, "block": 256, "blocks_per_cu": 4, "refill": 64, "chunk": 0},
    {"layout": 0, "stack_lds": 8, "block": 128, "blocks_per_cu": 2, "refill": 16, "chunk": 64},
    {"layout": 1, "stack_lds": 16, "block": 64, "blocks_per_cu": 4, "refill": 64, "chunk": 64},
])
def test_frame_wald_matches_oracle(oracle, small, cfg):
    """The Wald-triangle batch kernels (PAIR and REF layouts, all three schedules)."""
    from rfmanta_amd import _lib as L
    v, i, batches = small
    nodes, ids, _ = oracle.build_bvh(v, i)
    with api.Scene(v, i) as s:
        s.load_dynbvh(ids, nodes, prim_kind=L.MH_PRIM_WALD)
        dev = _device(batches)
        f = api.Frame(s, dev)
        f.set_config(cfg)
        f.launch()
        api.sync()
        for (k, rays), (dr, dh, any_hit) in zip(batches, dev):
            g = dh.download()
            o = oracle.trace(v, i, nodes, ids, rays, any_hit=any_hit, packet_size=1, prim_kind=1)
            assert np.array_equal(g["prim"], o["prim"]), (k, len(rays["ox"]))
            assert np.array_equal(_bits(g["t"]), _bits(o["t"])), k
            if any_hit:
                assert np.array_equal(g["occluded"], o["occluded"])
            else:
                assert np.array_equal(_bits(g["u"]), _bits(o["u"]))
                assert np.array_equal(_bits(g["v"]), _bits(o["v"]))
        f.close()


def test_graph_refused_after_rebuild(small):
    """A captured frame holds the scene's device view by value: after a rebuild with the
    Wald test (another triangle kernel, new buffers) replaying it is refused until the
    frame is captured again, and the new capture replays correctly."""
    from rfmanta_amd import _lib as L
    v, i, batches = small
    with api.Scene(v, i) as s:
        s.build()
        dev = _device(batches)
        stream = api.stream_create()
        try:
            f = api.Frame(s, dev)
            f.capture(stream)
            f.replay(stream)
            api.sync(stream)
            s.build(prim_kind=L.MH_PRIM_WALD)
            with pytest.raises(L.MhError, match="capture the frame again"):
                f.replay(stream)
            with pytest.raises(L.MhError, match="capture the frame again"):
                f.time(1, use_graph=True, stream=stream)
            f.launch()  # an eager launch reads the scene as it is now
            api.sync()
            ref = [dh.download() for _, dh, _ in dev]
            f.capture(stream)
            for _, dh, _ in dev:
                dh.arrays["prim"].upload(np.full(dh.n, 777, np.int32))
            f.replay(stream)
            api.sync(stream)
            for (_, dh, _), r in zip(dev, ref):
                assert np.array_equal(dh.download()["prim"], r["prim"])
            f.close()
        finally:
            api.stream_destroy(stream)


def test_soa_layout_follows_refits(oracle):
    """The SOA layout reads the scene's copy of its PairNodes: after a refit moves the boxes,
    a launch and a graph replay both see the new ones (results equal the oracle on the
    refitted tree, frame after frame)."""
    v, i = scenes.heightfield(120, 60, seed=4)
    nodes, ids, _ = oracle.build_bvh(v, i)
    rays = scenes.pinhole_rays((0.0, 0.9, -1.1), (0, 0, 0), (0, 1, 0), 60, 60, 96, 64)
    cfg = {"layout": 2, "stack_lds": 8, "block": 256, "blocks_per_cu": 2, "refill": 64,
           "chunk": 0}
    stream = api.stream_create()
    with api.Scene(v, i) as s:
        s.load_dynbvh(ids, nodes)
        dr = api.DeviceRays.upload(rays)
        dh = api.DeviceHits.alloc(dr.n)
        f = api.Frame(s, [(dr, dh, False)])
        f.set_config(cfg)
        f.launch()
        api.sync()
        f.capture(stream)
        for frame in range(3):
            vd = v.copy()
            vd[:, 1] += np.float32(0.05) * np.sin(np.float32(3.0) * v[:, 0] + np.float32(frame))
            s.set_vertices(vd)
            s.refit()
            api.sync()
            ref = nodes.copy()
            oracle.refit(vd, i, ref, ids)
            o = oracle.trace(vd, i, ref, ids, rays, packet_size=1)
            for replay in (False, True):
                dh.arrays["prim"].upload(np.full(dr.n, 777, np.int32))
                api.sync()
                if replay:
                    f.replay(stream)
                    api.sync(stream)
                else:
                    f.launch()
                    api.sync()
                g = dh.download()
                assert np.array_equal(g["prim"], o["prim"]), (frame, replay)
                assert np.array_equal(_bits(g["t"]), _bits(o["t"])), (frame, replay)
            assert (o["prim"] >= 0).sum() > 1000
        f.close()
    api.stream_destroy(stream)
