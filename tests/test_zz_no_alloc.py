"""mh_trace_noalloc: the allocation-free trace entry point against mh_trace.

Called through the C ABI (ctypes) on the same device-resident rays: results equal
mh_trace's bit for bit on two calls in a row (the second with fewer rays, on the array the
first left behind), closest-hit and any-hit, at LDS rings that make the traversal spill;
no device memory is taken by the calls; a tree too deep for the fixed spill area is
refused before anything is launched.
"""

import ctypes as C

import numpy as np
import pytest

from rfmanta_amd import _lib as L
from rfmanta_amd import api, scenes

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.asarray(a).view(np.uint32)


def _free_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def _call(entry, scene, dr, dh, any_hit, opts):
    r, h = dr.struct(), dh.struct()
    fn = getattr(L.load(), entry)
    args = [scene.h, C.byref(r), C.byref(h), dr.n, L.MH_ANY_HIT if any_hit else 0, C.byref(opts)]
    if entry != "mh_trace_host":
        args.append(None)
    return fn(*args)


def _same(a, b, any_hit, what):
    if any_hit:
        assert np.array_equal(a["occluded"], b["occluded"]), what
        return
    assert np.array_equal(a["prim"], b["prim"]), what
    for k in ("t", "u", "v"):
        assert np.array_equal(_bits(a[k]), _bits(b[k])), (what, k)


def _sets(name):
    if name == "soup10k":
        v, i = scenes.config0_scene()
        prim, light = scenes.config0_camera(512, 512), scenes.CONFIG0_LIGHT
    else:  # the 1000 x 500 heightfield (997,002 triangles) and its 1920x1080 camera
        v, i = scenes.config1_scene()
        prim, light = scenes.config1_camera(), scenes.CONFIG1_LIGHT
    s = api.Scene(v, i).build()
    o = s.trace_host(prim)
    shadow = scenes.shadow_rays(v, i, prim, o, light, compact=True)
    return s, prim, shadow


@pytest.fixture(scope="module")
def scene_sets():
    made = {}

    def get(name):
        if name not in made:
            made[name] = _sets(name)
        return made[name]

    yield get
    for s, _, _ in made.values():
        s.close()


# (scene, LDS ring): tree depth 16 and 22; with ring 8 on the heightfield a lane spills
# 14 entries, so the workgroup count is lowered to fit the fixed area (1,170 of 4,096)
@pytest.mark.parametrize("name,stack_lds", [("soup10k", 8), ("heightfield1M", 16),
                                            ("heightfield1M", 8)])
def test_noalloc_equals_trace_twice(scene_sets, name, stack_lds):
    s, prim, shadow = scene_sets(name)
    depth = s.stats()["depth"]
    print(f"noalloc: {name} depth {depth} ring {stack_lds}", flush=True)
    assert depth > stack_lds, "the traversal must spill for this test to use the scratch area"
    for kind, rays, any_hit in (("primary", prim, False), ("shadow", shadow, True)):
        n = len(rays["ox"])
        for count in (n, n // 5 + 3):  # the second call smaller, on the first call's leftovers
            sub = {k: np.ascontiguousarray(rays[k][:count]) for k in api.RAY_FIELDS}
            dr = api.DeviceRays.upload(sub)
            want = api.DeviceHits.alloc(count, any_hit=any_hit)
            got = api.DeviceHits.alloc(count, any_hit=any_hit)
            opts = api.trace_opts(1, stack_lds)
            L.check(_call("mh_trace", s, dr, want, any_hit, opts))
            L.check(_call("mh_trace_noalloc", s, dr, got, any_hit, opts))
            api.sync()
            _same(got.download(), want.download(), any_hit, (name, kind, count, stack_lds))


def test_noalloc_takes_no_device_memory():
    v, i = scenes.heightfield(200, 100, seed=3)
    with api.Scene(v, i).build() as s:
        rays = scenes.pinhole_rays((0.5, 0.6, -0.8), (0.5, 0.0, 0.5), (0, 1, 0), 60.0, 60.0,
                                   256, 256)
        dr = api.DeviceRays.upload(rays)
        dh = api.DeviceHits.alloc(dr.n)
        ref = api.DeviceHits.alloc(dr.n)
        opts = api.trace_opts(1, 8)
        L.check(_call("mh_trace_noalloc", s, dr, dh, False, opts))  # loads the kernels
        api.sync()
        before = _free_bytes()
        for _ in range(2):
            L.check(_call("mh_trace_noalloc", s, dr, dh, False, opts))
        api.sync()
        assert _free_bytes() == before
        # a fresh scene has no spill buffer: mh_trace grows one, mh_trace_noalloc did not
        L.check(_call("mh_trace", s, dr, ref, False, opts))
        api.sync()
        _same(dh.download(), ref.download(), False, "heightfield")


def test_noalloc_refuses_a_tree_too_deep():
    """A chain-shaped tree (each inner record holds one leaf and the next inner record),
    loaded from a record array, 80 levels deep. With an 8-entry LDS ring a lane spills up
    to 72 entries (576 bytes): 256 workgroups of 256 lanes need 36 MiB, more than the
    fixed 32 MiB, so a call with rays for every compute unit is refused before anything
    runs; with a 32-entry ring (48 entries spilled, 24 MiB) it is accepted."""
    depth = 80
    T = depth + 1
    v = np.zeros((3 * T, 3), np.float32)
    for k in range(T):
        v[3 * k:3 * k + 3] = [[k, 0, 0], [k + 0.5, 1, 0], [k + 1, 0, 0.5]]
    idx = np.arange(3 * T, dtype=np.uint32).reshape(T, 3)
    nodes = np.zeros(2 * T - 1, L.NODE_DTYPE)
    lo, hi = v.reshape(T, 3, 3).min(1), v.reshape(T, 3, 3).max(1)
    # record 0 is the root; inner record r at chain level k has children 2k+1 (leaf of
    # triangle k) and 2k+2 (the rest)
    for k in range(depth):
        r = 0 if k == 0 else 2 * k
        nodes[r]["lo"], nodes[r]["hi"] = lo[k:].min(0), hi[k:].max(0)
        nodes[r]["child"], nodes[r]["axis"], nodes[r]["children"] = 2 * k + 1, 0, 0
        leaf = 2 * k + 1
        nodes[leaf]["lo"], nodes[leaf]["hi"] = lo[k], hi[k]
        nodes[leaf]["child"], nodes[leaf]["children"] = k, 1
    last = 2 * depth
    nodes[last]["lo"], nodes[last]["hi"] = lo[depth], hi[depth]
    nodes[last]["child"], nodes[last]["children"] = depth, 1
    ids = np.arange(T, dtype=np.int32)
    with api.Scene(v, idx) as s:
        s.load_dynbvh(ids, nodes)
        assert depth <= s.stats()["depth"] <= depth + 1
        assert api.device_info()["cu_count"] == 256
        m = 256 * 256  # one workgroup's rays for each of 256 compute units
        rays = {"ox": np.linspace(0.2, T - 0.2, m, dtype=np.float32),
                "oy": np.full(m, 0.3, np.float32), "oz": np.full(m, -1, np.float32),
                "dx": np.zeros(m, np.float32), "dy": np.zeros(m, np.float32),
                "dz": np.ones(m, np.float32), "tmax": np.full(m, 1e19, np.float32)}
        dr = api.DeviceRays.upload(rays)
        dh = api.DeviceHits.alloc(m)
        ref = api.DeviceHits.alloc(m)
        rc = _call("mh_trace_noalloc", s, dr, dh, False, api.trace_opts(1, 8))
        assert rc == L.MH_E_UNSUPPORTED
        assert b"fixed area" in L.load().mh_last_error()
        # a 32-entry ring leaves at most 49 entries per lane to spill: it fits
        L.check(_call("mh_trace_noalloc", s, dr, dh, False, api.trace_opts(1, 32)))
        L.check(_call("mh_trace", s, dr, ref, False, api.trace_opts(1, 8)))
        api.sync()
        _same(dh.download(), ref.download(), False, "chain")
