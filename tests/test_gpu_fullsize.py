"""Parity at the workloads' own sizes: every ray set, table and scene the benchmark times,
through the launches it times, against the oracle.

* 1M-triangle heightfield (1000 x 500 vertices, 997,002 triangles): the host SAH tree
  equals the oracle's byte for byte; 1920x1080 primary rays, their diffuse-bounce rays and
  their shadow rays, traced by the plain kernel and by the persistent batch kernel under
  the launch configurations the benchmark's autotuner picks on this workload (and under
  the autotuner's live pick), equal the oracle's single-ray traversal bit for bit; against
  the reference's 8x8 packets, prim ids differ only inside the tie classes, whose counts
  are printed.
* 64 uneven batches (1k-2M rays, mixed kinds, 19.4M rays) and the same sizes all of bounce
  rays, in one persistent launch under the whole-wave and the per-lane refill schedules.
* Serving: requests of the serving workload over the heightfield and the 8M-triangle soup,
  in flight together, against the oracle (camera rays and shadow rays generated on the
  device from the oracle's own primary hits).
* 8M-triangle heightfield (4000 x 1000 vertices): three frames of device displacement
  (checked against 0.01 sin(4x + t)) and the rotating update, canonical trees equal to
  the oracle's update, and the 1080p frame traced on each.

The oracle runs multithreaded on the host; each fixture is sized to finish in seconds.
"""

import os

import numpy as np
import pytest

from conftest import tie_classes

pytestmark = pytest.mark.gpu

from rfmanta_amd import api, scenes  # noqa: E402

# Launch configurations the benchmark's autotuner picks on these workloads; C1_FRAME is the
# one bench.py pins for its timed C1 frame.
C1_FRAME = scenes.C1_LAUNCH
C1_STATIC = scenes.C1_LAUNCH_STATIC
C1_COHERENT = {"layout": 0, "stack_lds": 8, "block": 256, "blocks_per_cu": 16, "refill": 64,
               "chunk": 0, "xcd_remap": 0}
C1_SHADOW = {"layout": 0, "stack_lds": 8, "block": 256, "blocks_per_cu": 24, "refill": 64,
             "chunk": 0, "xcd_remap": 1}
C2_MIXED = {"layout": 0, "stack_lds": 8, "block": 256, "blocks_per_cu": 16, "refill": 64,
            "chunk": 256, "xcd_remap": 0}
C2_LANE = {"layout": 0, "stack_lds": 8, "block": 128, "blocks_per_cu": 12, "refill": 32,
           "chunk": 64, "xcd_remap": 0}

THREADS = min(16, len(os.sched_getaffinity(0)))


def _bits(a):
    return np.asarray(a).view(np.uint32)


def _same(g, o, any_hit, what):
    assert np.array_equal(g["prim"], o["prim"]), (what, int((g["prim"] != o["prim"]).sum()))
    assert np.array_equal(_bits(g["t"]), _bits(o["t"])), what
    if any_hit:
        assert np.array_equal(g["occluded"], o["occluded"]), what
    else:
        assert np.array_equal(_bits(g["u"]), _bits(o["u"])), what
        assert np.array_equal(_bits(g["v"]), _bits(o["v"])), what


@pytest.fixture(scope="module")
def c1(oracle):
    print("c1: heightfield 1000x500", flush=True)
    v, i = scenes.config1_scene()
    s = api.Scene(v, i).build()
    gi, gn = s.export_dynbvh()
    nodes, ids, _ = oracle.build_bvh(v, i)
    tree_equal = np.array_equal(gi, ids) and gn.tobytes() == nodes.tobytes()
    prim = scenes.config1_camera()
    o_prim = oracle.trace(v, i, nodes, ids, prim, threads=THREADS)
    bounce = scenes.bounce_rays(v, i, prim, o_prim, seed=1)
    shadow = scenes.shadow_rays(v, i, prim, o_prim, scenes.CONFIG1_LIGHT, compact=True)
    sets = {"primary": (prim, False), "bounce": (bounce, False), "shadow": (shadow, True)}
    ref = {"primary": o_prim}
    for k in ("bounce", "shadow"):
        r, a = sets[k]
        ref[k] = oracle.trace(v, i, nodes, ids, r, any_hit=a, threads=THREADS)
    dev = {k: (api.DeviceRays.upload(r), api.DeviceHits.alloc(len(r["ox"]), any_hit=a), a)
           for k, (r, a) in sets.items()}
    yield dict(v=v, i=i, scene=s, nodes=nodes, ids=ids, tree_equal=tree_equal, sets=sets,
               ref=ref, dev=dev)
    s.close()


def test_c1_tree_matches_oracle(c1):
    assert c1["tree_equal"]
    assert len(c1["i"]) == 997_002


@pytest.mark.parametrize("name", ["primary", "bounce", "shadow"])
def test_c1_plain_matches_oracle(c1, name):
    print(f"c1 plain {name}", flush=True)
    dr, dh, any_hit = c1["dev"][name]
    c1["scene"].trace(dr, dh, any_hit=any_hit, opts=api.trace_opts(1))
    api.sync()
    _same(dh.download(), c1["ref"][name], any_hit, name)
    if name == "primary":
        assert (c1["ref"][name]["prim"] >= 0).sum() > 1_500_000


def _scribble(dh):
    dh.arrays["prim"].upload(np.full(dh.n, 777, np.int32))
    dh.arrays["t"].upload(np.full(dh.n, -1.0, np.float32))


@pytest.mark.parametrize("cfg", [C1_FRAME, C1_STATIC, "autotune"],
                         ids=["bench_pick", "static_stride", "live_pick"])
def test_c1_frame_matches_oracle(c1, cfg):
    """The benchmark's timed launch: one k_trace_batch over the primary + bounce batches."""
    print(f"c1 frame {cfg}", flush=True)
    names = ("primary", "bounce")
    batches = [c1["dev"][k] for k in names]
    for _, dh, _ in batches:
        _scribble(dh)
    f = api.Frame(c1["scene"], batches)
    if cfg == "autotune":
        picked, _ = f.autotune(iters=1, force=True)
        print(f"autotuner picked {picked}", flush=True)
    else:
        f.set_config(cfg)
    for _ in range(2):  # the second launch runs on counters the first reset
        f.launch()
    api.sync()
    for k, (_, dh, a) in zip(names, batches):
        _same(dh.download(), c1["ref"][k], a, k)
    f.close()


@pytest.mark.parametrize("name,cfg", [("primary", C1_COHERENT), ("bounce", C1_STATIC),
                                      ("bounce", C1_FRAME), ("shadow", C1_SHADOW)])
def test_c1_single_set_frames_match_oracle(c1, name, cfg):
    print(f"c1 single-set frame {name} {cfg}", flush=True)
    dr, dh, a = c1["dev"][name]
    _scribble(dh)
    f = api.Frame(c1["scene"], [(dr, dh, a)])
    f.set_config(cfg)
    f.launch()
    api.sync()
    _same(dh.download(), c1["ref"][name], a, name)
    f.close()


@pytest.mark.parametrize("name", ["primary", "bounce", "shadow"])
def test_c1_single_ray_vs_reference_packets(oracle, c1, name):
    """GPU single-ray traversal against the reference's 64-ray packets (8x8 tiles for the
    primary rays; consecutive secondary rays): prim ids differ only in the tie classes."""
    r, a = c1["sets"][name]
    o = oracle.trace(c1["v"], c1["i"], c1["nodes"], c1["ids"], r, any_hit=a, packet_size=64,
                     threads=THREADS)
    g = c1["ref"][name]  # equal to the GPU bit for bit (tests above)
    if a:
        # any-hit: the occlusion flag does not depend on the visit order
        assert np.array_equal(g["occluded"], o["occluded"])
        print(f"{name}: occluded {int(o['occluded'].sum())} of {len(o['occluded'])}", flush=True)
        return
    h1, h2, real = tie_classes(o["prim"], o["t"], g["prim"], g["t"])
    print(f"{name} tie classes vs 8x8 packets: H1={h1} H2={h2} other={real} "
          f"of {len(g['prim'])}", flush=True)
    assert real == 0
    same = (g["prim"] == o["prim"]) & (o["prim"] >= 0)
    for k in ("t", "u", "v"):
        x, y = g[k][same].astype(np.float64), o[k][same].astype(np.float64)
        assert np.all(np.abs(x - y) <= 1e-5 * np.maximum(np.abs(y), 1e-6)), k


@pytest.fixture(scope="module")
def c2(c1):
    print("c2: 64 batches", flush=True)
    mixed = scenes.config2_batches(c1["v"], c1["i"])
    start = 0
    inc = []
    for _, r in mixed:
        n = len(r["ox"])
        inc.append(("bounce", scenes.batch_rays("bounce", n, c1["v"], c1["i"], 1, start)))
        start += n
    return {"mixed": mixed, "incoherent": inc}


@pytest.mark.parametrize("table,cfg", [("mixed", C2_MIXED), ("mixed", C2_LANE),
                                       ("incoherent", C2_LANE)])
def test_c2_table_matches_oracle(oracle, c1, c2, table, cfg):
    batches = c2[table]
    total = sum(len(r["ox"]) for _, r in batches)
    print(f"c2 {table} {cfg}: {len(batches)} batches, {total} rays", flush=True)
    assert len(batches) == 64 and total > 15_000_000
    dev = []
    for kind, r in batches:
        a = kind == "shadow"
        dev.append((api.DeviceRays.upload(r), api.DeviceHits.alloc(len(r["ox"]), any_hit=a), a))
    f = api.Frame(c1["scene"], dev)
    f.set_config(cfg)
    f.launch()
    f.launch()
    api.sync()
    for b, ((kind, r), (dr, dh, a)) in enumerate(zip(batches, dev)):
        o = oracle.trace(c1["v"], c1["i"], c1["nodes"], c1["ids"], r, any_hit=a,
                         threads=THREADS)
        _same(dh.download(), o, a, (table, b, kind, len(r["ox"])))
    f.close()


def test_c3_served_requests_match_oracle(oracle, c1):
    """Requests of the serving workload, submitted together (they share rounds and
    launches), over the heightfield and the 8M-triangle soup."""
    print("c3: soup 8M", flush=True)
    vs, is_ = scenes.soup(8_000_000, seed=0, scale=4.0)
    soup = api.Scene(vs, is_).build()
    s_ids, s_nodes = soup.export_dynbvh()
    meshes = [(c1["v"], c1["i"], c1["nodes"], c1["ids"]), (vs, is_, s_nodes, s_ids)]
    scs = [c1["scene"], soup]
    reqs = scenes.c3_requests(40)
    # every resolution, both scenes, several spp: the first 12 of the stream
    pick = reqs[:12]
    assert {(r[1]["xres"], r[1]["yres"]) for r in pick} >= {(1920, 1080), (256, 256)}
    assert {r[0] for r in pick} == {0, 1}
    srv = api.Server(scs, hbm_budget=8 << 30, host_budget=2 << 30, chunk_rays=1 << 21,
                     max_active=64, shadows=True, light=scenes.C3_LIGHT)
    cams = [api.camera(**kw) for _, kw in pick]
    ids = [srv.submit(sc, cam) for (sc, _), cam in zip(pick, cams)]
    got = [srv.wait(rid) for rid in ids]
    srv.close()
    for (sc, kw), cam, g in zip(pick, cams, got):
        print(f"c3 request scene {sc} {kw['xres']}x{kw['yres']} spp {kw['spp']}", flush=True)
        v, i, nodes, ids_ = meshes[sc]
        n = cam.xres * cam.yres * cam.spp
        dp = api.DeviceRays.alloc(n, masked=False)
        api.gen_camera_rays(cam, dp)
        api.sync()
        rays = dp.download()
        o = oracle.trace(v, i, nodes, ids_, rays, threads=THREADS)
        assert np.array_equal(g["prim"], o["prim"]), (sc, kw)
        assert np.array_equal(_bits(g["t"]), _bits(o["t"])), (sc, kw)
        # shadow rays from the oracle's primary hits, generated by the device generator
        dh = api.DeviceHits.alloc(n)
        dh.arrays["prim"].upload(o["prim"])
        dh.arrays["t"].upload(o["t"])
        sh = api.DeviceRays.alloc(n)
        api.gen_shadow_rays(scs[sc], dp, dh, scenes.C3_LIGHT, sh)
        api.sync()
        so = oracle.trace(v, i, nodes, ids_, sh.download(), any_hit=True, threads=THREADS)
        assert np.array_equal(g["occluded"], so["occluded"]), (sc, kw)
        assert g["rays"] == 2 * n
    soup.close()


def _canonical_equal(ga, gi, oa, oi):
    """Both trees walked from the root a level at a time (numpy); (ok, message, records)."""
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
        if leaf.any():
            cnt = c[leaf]
            rel = np.arange(int(cnt.sum())) - np.repeat(np.cumsum(cnt) - cnt, cnt)
            sg = np.repeat(ga["child"][g[leaf]].astype(np.int64), cnt) + rel
            so = np.repeat(oa["child"][o[leaf]].astype(np.int64), cnt) + rel
            if not np.array_equal(gi[sg], oi[so]):
                return False, "leaf ids differ", n
        gc = ga["child"][g[~leaf]].astype(np.int64)
        oc = oa["child"][o[~leaf]].astype(np.int64)
        g = np.concatenate([gc, gc + 1])
        o = np.concatenate([oc, oc + 1])
    return True, "", n


def test_c4_update_frames_match_oracle(oracle):
    """Three frames of the dynamic-scene workload at 7,990,002 triangles."""
    print("c4: heightfield 4000x1000", flush=True)
    v, i = scenes.heightfield(4000, 1000, seed=0)
    assert len(i) == 7_990_002
    cam = api.camera(**scenes.C4_CAMERA)
    n = cam.xres * cam.yres
    with api.Scene(v, i) as s:
        s.build()
        ids, nodes = s.export_dynbvh()
        used = s.stats()["nodes"]
        st = oracle.RotState(nodes, ids, used)
        dp = api.DeviceRays.alloc(n, masked=False)
        api.gen_camera_rays(cam, dp)
        dh = api.DeviceHits.alloc(n)
        api.sync()
        rays = dp.download()
        x = v[:, 0].astype(np.float64)
        for f in range(3):
            print(f"c4 frame {f}", flush=True)
            phase = 0.1 * f
            s.displace_sine(0.01, 4.0, phase)
            api.sync()
            vd = s.device_vertices()
            # y = y0 + 0.01 sin(4x + t) in binary32: within the rounding of 4x + t, of the
            # single-precision sine (a few ulp), of the product and of the sum
            exact = v[:, 1].astype(np.float64) + 0.01 * np.sin(4.0 * x + phase)
            bound = (2.0 ** -23 * (np.abs(exact) + 0.01 * (np.abs(4.0 * x) + abs(phase) + 1.0))
                     + 0.01 * 2.0 ** -21)
            assert np.all(np.abs(vd[:, 1] - exact) <= bound)
            assert np.array_equal(vd[:, 0], v[:, 0]) and np.array_equal(vd[:, 2], v[:, 2])
            st.update(vd, i)
            s.update()
            api.sync()
            gi, gn = s.export_dynbvh()
            ok, msg, count = _canonical_equal(gn, gi, st.nodes, st.ids)
            assert ok, (f, msg)
            assert count == st.next_free.value
            s.trace(dp, dh)
            api.sync()
            g = dh.download()
            o = oracle.trace(vd, i, st.nodes, st.ids, rays, threads=THREADS)
            _same(g, o, False, f"c4 frame {f}")
            assert (o["prim"] >= 0).sum() > n // 2
