"""Serving loop (mh_server_*): results identical to the direct path, alone and in a mix.

Each request's hits are checked against the same camera traced through mh_gen_camera_rays
+ mh_trace (+ mh_gen_shadow_rays + an any-hit mh_trace, + mh_gen_bounce_rays with seed
camera.seed + 1 + mh_trace), and the first request against the oracle. Admission is
checked against the HBM budget (queued, never over budget; oversized requests rejected)
and against a host staging budget that only frees up as results are collected.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from rfmanta_amd import _lib as L  # noqa: E402
from rfmanta_amd import api, scenes  # noqa: E402

LIGHT = (2.0, 2.0, -2.0)


def direct(scene, cam, shadows, bounces=False):
    n = cam.xres * cam.yres * cam.spp
    dp = api.DeviceRays.alloc(n, masked=False)
    api.gen_camera_rays(cam, dp)
    dh = api.DeviceHits.alloc(n)
    scene.trace(dp, dh)
    out = {}
    if shadows:
        sh = api.DeviceRays.alloc(n)
        api.gen_shadow_rays(scene, dp, dh, LIGHT, sh)
        so = api.DeviceHits.alloc(n, any_hit=True)
        scene.trace(sh, so, any_hit=True)
        api.sync()
        out["occluded"] = so.download()["occluded"]
    if bounces:
        bo = api.DeviceRays.alloc(n)
        api.gen_bounce_rays(scene, dp, dh, cam.seed + 1, bo)
        bh = api.DeviceHits.alloc(n)
        scene.trace(bo, bh)
        api.sync()
        b = bh.download()
        out["bounce_prim"], out["bounce_t"] = b["prim"], b["t"]
    api.sync()
    h = dh.download()
    out["prim"], out["t"] = h["prim"], h["t"]
    out["rays"] = dp.download()
    return out


@pytest.fixture(scope="module")
def two_scenes():
    v0, i0 = scenes.heightfield(200, 100, seed=3)
    v1, i1 = scenes.soup(20000, seed=4)
    s0 = api.Scene(v0, i0).build()
    s1 = api.Scene(v1, i1).build()
    yield (s0, v0, i0), (s1, v1, i1)
    s0.close()
    s1.close()


def cams():
    c = [api.camera((0, 0.9, -1.1), (0, 0, 0), (0, 1, 0), 50.0, 96, 64),
         api.camera((0.2, 1.2, -0.8), (0, 0, 0.1), (0, 1, 0), 40.0, 160, 96, spp=2, seed=9),
         api.camera((0, 0, -3), (0, 0, 0), (0, 1, 0), 60.0, 128, 128),
         api.camera((0.5, 0.3, -2.5), (0, 0, 0), (0, 1, 0), 45.0, 40, 56, spp=3, seed=1)]
    return c


@pytest.mark.parametrize("shadows,bounces", [(False, False), (True, False), (True, True)])
def test_alone_and_in_mix_identical(two_scenes, shadows, bounces, oracle):
    (s0, v0, i0), (s1, v1, i1) = two_scenes
    sc = [s0, s1]
    c = cams()
    jobs = [(0, c[0]), (0, c[1]), (1, c[2]), (1, c[3]), (0, c[3]), (1, c[0])]
    ref = [direct(sc[k], cam, shadows, bounces) for k, cam in jobs]
    keys = ("prim", "t") + (("occluded",) if shadows else ()) + \
        (("bounce_prim", "bounce_t") if bounces else ())
    # alone: one request at a time
    with_alone = []
    srv = api.Server(sc, shadows=shadows, bounces=bounces, light=LIGHT, chunk_rays=4096)
    for k, cam in jobs:
        with_alone.append(srv.wait(srv.submit(k, cam)))
    srv.close()
    # in a mix: all submitted at once, small chunks so requests share several rounds
    srv = api.Server(sc, shadows=shadows, bounces=bounces, light=LIGHT, chunk_rays=3000,
                     max_active=4)
    ids = [srv.submit(k, cam) for k, cam in jobs]
    mixed = [srv.wait(i) for i in reversed(ids)][::-1]
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(jobs) and st["active"] == 0 and st["pending"] == 0
    for j, (r, a, m) in enumerate(zip(ref, with_alone, mixed)):
        for key in keys:
            assert np.array_equal(a[key], r[key]), (j, key, "alone")
            assert np.array_equal(m[key], r[key]), (j, key, "mix")
        assert m["status"] == 0 and m["rays"] == len(r["prim"]) * (1 + shadows + bounces)
        assert m["submit_ms"] <= m["admit_ms"] <= m["done_ms"]
    # anchor the first request on the oracle
    nodes, ids_, _ = oracle.build_bvh(v0, i0)
    o = oracle.trace(v0, i0, nodes, ids_, ref[0]["rays"], packet_size=1, any_hit=False)
    assert np.array_equal(o["prim"], ref[0]["prim"]) and np.array_equal(o["t"], ref[0]["t"])
    assert (ref[0]["prim"] >= 0).mean() > 0.3


def test_admission_respects_budget(two_scenes):
    (s0, _, _), _ = two_scenes
    cam = api.camera((0, 0.9, -1.1), (0, 0, 0), (0, 1, 0), 50.0, 128, 128)
    n = 128 * 128
    one = 11 * 4 * n  # primary rays + hits, no shadows (already 256-aligned)
    srv = api.Server([s0], hbm_budget=2 * one + 1024, chunk_rays=2048)
    ids = [srv.submit(0, cam) for _ in range(6)]
    peak = 0
    for _ in range(200):
        st = srv.stats()
        peak = max(peak, st["hbm_in_use"])
        assert st["hbm_in_use"] <= 2 * one + 1024 and st["active"] <= 2
        if st["completed"] == 6:
            break
    outs = [srv.wait(i) for i in ids]
    for o in outs[1:]:
        assert np.array_equal(o["prim"], outs[0]["prim"]) and np.array_equal(o["t"], outs[0]["t"])
    # oversized request: rejected at submit, counted
    big = api.camera((0, 0.9, -1.1), (0, 0, 0), (0, 1, 0), 50.0, 1024, 1024)
    with pytest.raises(L.MhError):
        srv.submit(0, big)
    # collected / unknown ids are errors, not waits
    with pytest.raises(L.MhError):
        srv.wait(ids[0])
    with pytest.raises(L.MhError):
        srv.wait(10_000)
    st = srv.stats()
    assert st["rejected"] == 1 and st["completed"] == 6 and st["hbm_in_use"] == 0
    assert st["rays"] == 6 * n
    srv.close()


def test_device_allocation_failure_is_clean(two_scenes):
    """A server asked for more device memory than the GPU has: its hipMalloc fails and
    mh_server_create reports MH_E_NOMEM with nothing left behind; the next server is created,
    launches and serves a request identical to the direct path (the failed allocation's HIP
    error does not surface in later launch checks)."""
    (s0, _, _), _ = two_scenes
    with pytest.raises(L.MhError) as e:
        api.Server([s0], hbm_budget=1 << 50)
    assert e.value.code == L.MH_E_NOMEM
    cam = api.camera((0, 0.9, -1.1), (0, 0, 0), (0, 1, 0), 50.0, 96, 64)
    srv = api.Server([s0], hbm_budget=64 << 20, shadows=True, light=LIGHT)
    o = srv.wait(srv.submit(0, cam))
    d = direct(s0, cam, shadows=True)
    assert np.array_equal(o["prim"], d["prim"]) and np.array_equal(o["t"], d["t"])
    assert np.array_equal(o["occluded"], d["occluded"])
    srv.close()


def test_host_staging_frees_on_collection(two_scenes):
    """Results wait in pinned host staging; with room for one request's results, the others
    finish on the device and complete one by one as results are collected."""
    (s0, _, _), _ = two_scenes
    cam = api.camera((0, 0.9, -1.1), (0, 0, 0), (0, 1, 0), 50.0, 64, 64)
    n = 64 * 64
    ref = direct(s0, cam, True, True)
    host_one = 2 * 4 * n + n + 2 * 4 * n  # prim t, occluded, bounce prim t (256-aligned)
    srv = api.Server([s0], shadows=True, bounces=True, light=LIGHT, host_budget=host_one)
    ids = [srv.submit(0, cam) for _ in range(4)]
    for i in ids:
        o = srv.wait(i)
        for key in ("prim", "t", "occluded", "bounce_prim", "bounce_t"):
            assert np.array_equal(o[key], ref[key]), key
        assert srv.stats()["host_in_use"] <= host_one
    st = srv.stats()
    assert st["completed"] == 4 and st["host_in_use"] == 0 and st["hbm_in_use"] == 0
    srv.close()


def test_cancel_queued_and_in_flight(two_scenes):
    """mh_server_cancel: a queued request ends at once, one in flight ends behind the work
    already queued for it (or finishes normally if it got there first); the others are
    unaffected and every buffer is returned."""
    (s0, _, _), _ = two_scenes
    cam = api.camera((0, 0.9, -1.1), (0, 0, 0), (0, 1, 0), 50.0, 128, 128)
    n = 128 * 128
    one = 11 * 4 * n  # primary rays + hits, no shadows (256-aligned)
    ref = direct(s0, cam, False)
    # one request fits the budget at a time; small chunks make many launches per request
    srv = api.Server([s0], hbm_budget=one + 1024, chunk_rays=1024)
    ids = [srv.submit(0, cam) for _ in range(4)]
    srv.cancel(ids[3])  # still queued behind the budget
    first_cancelled = True
    try:
        srv.cancel(ids[0])
    except L.MhError as e:  # it had already finished
        assert e.code == L.MH_E_STATE
        first_cancelled = False
    for rid in ids[1:3]:
        o = srv.wait(rid)
        assert np.array_equal(o["prim"], ref["prim"]) and np.array_equal(o["t"], ref["t"])
    for rid, cancelled in ((ids[0], first_cancelled), (ids[3], True)):
        if cancelled:
            with pytest.raises(L.MhError) as e:
                srv.wait(rid)
            assert e.value.code == L.MH_E_CANCELED
        else:
            o = srv.wait(rid)
            assert np.array_equal(o["prim"], ref["prim"])
    with pytest.raises(L.MhError):  # collected
        srv.cancel(ids[1])
    st = srv.stats()
    assert st["completed"] == 4 - first_cancelled - 1
    assert st["active"] == 0 and st["pending"] == 0
    assert st["hbm_in_use"] == 0 and st["host_in_use"] == 0
    srv.close()
