"""GPU parity: libmh.so on the device against the oracle on the same inputs.

Bar: prim ids, occlusion flags and the bits of t, u, v identical to the oracle run with
the same packet size (the GPU reproduces the reference's binary32 arithmetic and visit
order). Across packet sizes (single-ray GPU traversal against the reference's 8x8
packets) prim ids may differ only inside the tie bands classified by tie_classes (H1
exact ties, H2 the 1e-4 Kensler-Shirley acceptance band), and t, u, v then agree to
1e-5 relative where the prim agrees.
"""

import numpy as np
import pytest

from conftest import tie_classes

pytestmark = pytest.mark.gpu

from rfmanta_amd import api, scenes  # noqa: E402
from rfmanta_amd import _lib as L  # noqa: E402


@pytest.fixture(scope="module")
def c0():
    v, i = scenes.config0_scene()
    rays = scenes.config0_camera(512, 512)
    return v, i, rays


@pytest.fixture(scope="module")
def c0_tree(oracle, c0):
    v, i, _ = c0
    nodes, ids, _ = oracle.build_bvh(v, i)
    return nodes, ids


@pytest.fixture(scope="module")
def c0_scene(c0):
    v, i, _ = c0
    s = api.Scene(v, i).build()
    yield s
    s.close()


def _same_bits(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


def test_device_is_gfx950():
    info = api.device_info()
    print("device", info, flush=True)
    assert info["arch"].startswith("gfx950")


def test_built_tree_matches_oracle(c0_scene, c0_tree):
    ids, nodes = c0_scene.export_dynbvh()
    o_nodes, o_ids = c0_tree
    assert np.array_equal(ids, o_ids)
    assert nodes.tobytes() == o_nodes.tobytes()


@pytest.mark.parametrize("packet", [1, 64])
def test_c0_primary_closest_hit_exact(oracle, c0, c0_tree, c0_scene, packet):
    v, i, rays = c0
    print(f"c0 primary packet={packet}", flush=True)
    g = c0_scene.trace_host(rays, packet_size=packet)
    o = oracle.trace(v, i, *c0_tree, rays, packet_size=packet)
    assert (g["prim"] >= 0).sum() > 20000
    assert np.array_equal(g["prim"], o["prim"])
    assert _same_bits(g["t"], o["t"])
    assert _same_bits(g["u"], o["u"])
    assert _same_bits(g["v"], o["v"])


@pytest.mark.parametrize("packet", [1, 64])
def test_c0_shadow_any_hit_exact(oracle, c0, c0_tree, c0_scene, packet):
    v, i, rays = c0
    print(f"c0 shadow packet={packet}", flush=True)
    prim = oracle.trace(v, i, *c0_tree, rays, packet_size=1)
    sh = scenes.shadow_rays(v, i, rays, prim, scenes.CONFIG0_LIGHT, compact=False)
    g = c0_scene.trace_host(sh, any_hit=True, packet_size=packet)
    o = oracle.trace(v, i, *c0_tree, sh, any_hit=True, packet_size=packet)
    assert np.array_equal(g["occluded"], o["occluded"])
    assert np.array_equal(g["prim"], o["prim"])  # same visit order: same first occluder
    unmasked = sh["masked"] == 0
    assert 0 < (g["occluded"][unmasked] == 1).sum() < unmasked.sum()


def test_c0_single_ray_gpu_vs_reference_packets(oracle, c0, c0_tree, c0_scene):
    """GPU single-ray traversal against the reference's 8x8 packet traversal."""
    v, i, rays = c0
    g = c0_scene.trace_host(rays, packet_size=1)
    o = oracle.trace(v, i, *c0_tree, rays, packet_size=64)
    h1, h2, real = tie_classes(o["prim"], o["t"], g["prim"], g["t"])
    print(f"tie band counts: H1={h1} H2={h2} other={real}", flush=True)
    assert real == 0
    same = g["prim"] == o["prim"]
    hit = same & (o["prim"] >= 0)
    for k in ("t", "u", "v"):
        a, b = g[k][hit].astype(np.float64), o[k][hit].astype(np.float64)
        assert np.all(np.abs(a - b) <= 1e-5 * np.maximum(np.abs(b), 1e-6))


def test_wald_exact(oracle, c0, c0_tree):
    v, i, rays = c0
    with api.Scene(v, i) as s:
        s.load_dynbvh(c0_tree[1], c0_tree[0], prim_kind=L.MH_PRIM_WALD)
        for packet in (1, 64):
            g = s.trace_host(rays, packet_size=packet)
            o = oracle.trace(v, i, *c0_tree, rays, packet_size=packet, prim_kind=1)
            assert np.array_equal(g["prim"], o["prim"])
            assert _same_bits(g["t"], o["t"])
            assert _same_bits(g["u"], o["u"])
            assert _same_bits(g["v"], o["v"])


@pytest.mark.parametrize("packet", [1, 64])
def test_counters_match_oracle(oracle, c0, c0_tree, c0_scene, packet):
    v, i, rays = c0
    cnt = api.DeviceArray.from_numpy(np.zeros(2, np.int64))
    dr = api.DeviceRays.upload(rays)
    dh = api.DeviceHits.alloc(len(rays["ox"]))
    c0_scene.trace(dr, dh, opts=api.trace_opts(packet, counters=cnt))
    api.sync()
    c = cnt.download(np.int64, 2)
    o = oracle.trace(v, i, *c0_tree, rays, packet_size=packet, counters=True)
    assert int(c[0]) == o["box_tests"]
    assert int(c[1]) == o["tri_tests"]


@pytest.mark.parametrize("stack", [8, 16, 32])
def test_stack_ring_and_spill(oracle, stack):
    """Deep heightfield tree with small LDS rings: spills to global memory stay exact."""
    v, i = scenes.heightfield(200, 100, seed=3)
    rays = scenes.pinhole_rays((0.0, 0.9, -1.1), (0, 0, 0), (0, 1, 0), 60, 60, 128, 128)
    nodes, ids, _ = oracle.build_bvh(v, i)
    o = oracle.trace(v, i, nodes, ids, rays, packet_size=1)
    with api.Scene(v, i) as s:
        s.load_dynbvh(ids, nodes)
        assert s.stats()["depth"] > 8
        g = s.trace_host(rays, packet_size=1, stack_lds=stack)
    assert np.array_equal(g["prim"], o["prim"])
    assert _same_bits(g["t"], o["t"])


def test_refit_bit_exact(oracle):
    v, i = scenes.heightfield(160, 80, seed=5)
    nodes, ids, _ = oracle.build_bvh(v, i)
    with api.Scene(v, i) as s:
        s.load_dynbvh(ids, nodes)
        for frame in range(3):
            s.displace_sine(0.01, 4.0, 0.7 * frame)
            s.refit()
            api.sync()
            vd = s.device_vertices()
            ref = nodes.copy()
            oracle.refit(vd, i, ref, ids)
            _, gn = s.export_dynbvh()
            assert gn.tobytes() == ref.tobytes(), f"frame {frame}"
            rays = scenes.pinhole_rays((0.0, 0.9, -1.1), (0, 0, 0), (0, 1, 0), 60, 60, 64, 64)
            o = oracle.trace(vd, i, ref, ids, rays, packet_size=1)
            g = s.trace_host(rays)
            assert np.array_equal(g["prim"], o["prim"]), frame
            assert _same_bits(g["t"], o["t"]), frame


def test_edge_cases(oracle):
    """Masked rays, axis-aligned rays, rays starting inside the mesh, parallel rays."""
    v, i = scenes.soup(2000, seed=11)
    nodes, ids, _ = oracle.build_bvh(v, i)
    n = 4096
    rays = scenes.random_rays(n, 12, (-1.2, -1.2, -1.2), (1.2, 1.2, 1.2))
    ax = np.arange(n) % 7 == 0
    for k in ("dx", "dy"):
        rays[k] = np.where(ax, np.float32(0.0), rays[k]).astype(np.float32)
    rays["dz"] = np.where(ax, np.float32(1.0), rays["dz"]).astype(np.float32)
    rays["masked"] = (np.arange(n) % 13 == 0).astype(np.uint8)
    rays["tmax"] = np.where(np.arange(n) % 5 == 0, np.float32(0.5), rays["tmax"]).astype(np.float32)
    with api.Scene(v, i) as s:
        s.load_dynbvh(ids, nodes)
        for any_hit in (False, True):
            for packet in (1, 64, 7):
                g = s.trace_host(rays, any_hit=any_hit, packet_size=packet)
                o = oracle.trace(v, i, nodes, ids, rays, any_hit=any_hit, packet_size=packet)
                assert np.array_equal(g["prim"], o["prim"]), (any_hit, packet)
                assert _same_bits(g["t"], o["t"]), (any_hit, packet)
                if any_hit:
                    assert np.array_equal(g["occluded"], o["occluded"])
    assert (g["prim"][rays["masked"] == 1] == L.MH_MASKED).all()


def test_single_triangle_and_tiny_meshes(oracle):
    for T in (1, 2, 3, 5):
        v, i = scenes.soup(T, seed=T, sigma=0.3)
        nodes, ids, _ = oracle.build_bvh(v, i)
        rays = scenes.random_rays(2048, 99, (-1.5, -1.5, -1.5), (1.5, 1.5, 1.5))
        with api.Scene(v, i) as s:
            s.build()
            gi, gn = s.export_dynbvh()
            assert gn.tobytes() == nodes.tobytes()
            for packet in (1, 64):  # T = 1: the root is a leaf
                g = s.trace_host(rays, packet_size=packet)
                o = oracle.trace(v, i, nodes, ids, rays, packet_size=packet)
                assert np.array_equal(g["prim"], o["prim"]), (T, packet)


def _write_tree(path, ids, nodes):
    """DynBVH::saveToFile layout: u32 n_ids, i32 ids[], u32 n_nodes, 32-byte nodes."""
    with open(path, "wb") as f:
        f.write(np.uint32(len(ids)).tobytes())
        f.write(np.ascontiguousarray(ids, np.int32).tobytes())
        f.write(np.uint32(len(nodes)).tobytes())
        f.write(np.ascontiguousarray(nodes).tobytes())


def test_file_load_recomputes_rtsah(oracle, tmp_path):
    """DynBVH::buildFromFile runs computeTraversalCost after reading (DynBVH.cc:2469-2472):
    a file whose isLeftCheaper bits are cleared loads with the bits RTSAH gives, and any-hit
    rays report the occluders the oracle's traversal of that tree finds."""
    v, i = scenes.soup(3000, seed=21)
    nodes, ids, _ = oracle.build_bvh(v, i)
    want = nodes.copy()
    oracle.rtsah(want)
    cleared = nodes.copy()
    cleared["bits"] &= np.uint8(0xFE)
    assert (cleared["bits"] != want["bits"]).any()
    p = tmp_path / "tree.bvh"
    _write_tree(p, ids, cleared)
    rays = scenes.random_rays(4096, 22, (-1.2, -1.2, -1.2), (1.2, 1.2, 1.2))
    with api.Scene(v, i) as s:
        s.load_dynbvh_file(str(p))
        _, gn = s.export_dynbvh()
        assert gn.tobytes() == want.tobytes()
        g = s.trace_host(rays, any_hit=True)
        o = oracle.trace(v, i, want, ids, rays, any_hit=True, packet_size=1)
        assert np.array_equal(g["prim"], o["prim"])
        assert np.array_equal(g["occluded"], o["occluded"])
        # the in-memory import takes the bits as given
        s.load_dynbvh(ids, cleared)
        _, gn = s.export_dynbvh()
        assert gn.tobytes() == cleared.tobytes()


def test_overlapping_leaves_load_and_trace(oracle):
    """A tree whose ten leaves each hold all four objects (more leaves than objects, so
    not a partition): it loads, traces like the oracle and refits, and the update refuses."""
    from rfmanta_amd import _lib as L
    v, i = scenes.soup(4, seed=23, sigma=0.3)
    lo, hi = v.min(0), v.max(0)
    nodes = np.zeros(19, L.NODE_DTYPE)
    nodes["lo"], nodes["hi"] = lo, hi
    # records 0..8 inner (children 2k+1, 2k+2), records 9..18 leaves over slots [0, 4)
    for k in range(9):
        nodes["child"][k] = 2 * k + 1
    for k in range(9, 19):
        nodes["child"][k] = 0
        nodes["children"][k] = 4
    ids = np.arange(4, dtype=np.int32)
    rays = scenes.random_rays(2048, 24, (-1.5, -1.5, -1.5), (1.5, 1.5, 1.5))
    with api.Scene(v, i) as s:
        s.load_dynbvh(ids, nodes)
        assert s.stats()["nodes"] == 19
        for packet in (1, 64):
            g = s.trace_host(rays, packet_size=packet)
            o = oracle.trace(v, i, nodes, ids, rays, packet_size=packet)
            assert np.array_equal(g["prim"], o["prim"]), packet
            assert _same_bits(g["t"], o["t"]), packet
        s.refit()
        api.sync()
        _, gn = s.export_dynbvh()
        ref = nodes.copy()
        oracle.refit(v, i, ref, ids)
        assert gn.tobytes() == ref.tobytes()
        with pytest.raises(L.MhError, match="refit"):
            s.update()


def test_non_finite_and_degenerate_rays(oracle):
    """Directions with +-0, +-inf and NaN components, zero directions, non-finite origins
    and tmax of inf, 0, negative and NaN: the device gives the oracle's results (the
    reference's scalar arithmetic, select-based min/max included) at packet sizes 1 and 64,
    closest-hit and any-hit."""
    v, i = scenes.soup(2000, seed=31)
    nodes, ids, _ = oracle.build_bvh(v, i)
    n = 64 * 12
    rays = scenes.random_rays(n, 32, (-1.2, -1.2, -1.2), (1.2, 1.2, 1.2))
    inf, nan = np.inf, np.nan
    dirs = [(0.0, 0.0, 1.0), (-0.0, 0.0, 1.0), (0.0, -0.0, -1.0), (inf, 0.0, 0.0),
            (-inf, 1.0, 0.0), (nan, 0.0, 1.0), (0.0, 0.0, 0.0), (-0.0, -0.0, -0.0),
            (1.0, inf, -inf), (nan, nan, nan), (1e-30, -1e-30, 1.0), (0.0, 1e38, 0.0)]
    orgs = [(inf, 0.0, 0.0), (-inf, 0.1, 0.0), (nan, 0.0, 0.0), (0.0, 1e30, 0.0),
            (0.0, 0.0, -inf)]
    tmaxs = [np.float32(np.inf), np.float32(0.0), np.float32(-1.0), np.float32(np.nan)]
    k = np.arange(n)
    for j in np.nonzero(k % 3 == 0)[0]:
        d = dirs[(j // 3) % len(dirs)]
        rays["dx"][j], rays["dy"][j], rays["dz"][j] = d
    for j in np.nonzero(k % 5 == 1)[0]:
        o = orgs[(j // 5) % len(orgs)]
        rays["ox"][j], rays["oy"][j], rays["oz"][j] = o
    for j in np.nonzero(k % 7 == 2)[0]:
        rays["tmax"][j] = tmaxs[(j // 7) % len(tmaxs)]
    with api.Scene(v, i) as s:
        s.load_dynbvh(ids, nodes)
        for any_hit in (False, True):
            for packet in (1, 64):
                print(f"non-finite rays any_hit={any_hit} packet={packet}", flush=True)
                g = s.trace_host(rays, any_hit=any_hit, packet_size=packet)
                o = oracle.trace(v, i, nodes, ids, rays, any_hit=any_hit, packet_size=packet)
                assert np.array_equal(g["prim"], o["prim"]), (any_hit, packet)
                assert _same_bits(g["t"], o["t"]), (any_hit, packet)
                if any_hit:
                    assert np.array_equal(g["occluded"], o["occluded"])
                else:
                    hit = o["prim"] >= 0
                    assert _same_bits(g["u"][hit], o["u"][hit])
                    assert _same_bits(g["v"][hit], o["v"][hit])
        # the regular rays still hit
        assert (o["occluded"][k % 3 == 1] == 1).any()
