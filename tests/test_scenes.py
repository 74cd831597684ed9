"""Workload generators: PCG32 known answers, sizes, determinism, camera closed form."""

import numpy as np

from rfmanta_amd import scenes


def test_pcg32_reference_vectors():
    # pcg32_srandom_r(&rng, 42u, 54u) first six outputs (pcg-c-basic demo, O'Neill 2014)
    out = scenes.pcg32(42, np.array([54], np.uint64), 6)[:, 0]
    assert [hex(x) for x in out] == ["0xa15c02b7", "0x7b47f409", "0xba1d3330", "0x83d2f293",
                                     "0xbfa4784b", "0xcbed606e"]


def test_config_sizes():
    v, i = scenes.config0_scene()
    assert v.shape == (30000, 3) and i.shape == (10000, 3)
    v, i = scenes.heightfield(1000, 500)
    assert i.shape == (997_002, 3)  # 2 * 999 * 499
    rays = scenes.config0_camera(512, 512)
    assert len(rays["ox"]) == 512 * 512


def test_determinism():
    a = scenes.soup(1000, seed=0)
    b = scenes.soup(1000, seed=0)
    c = scenes.soup(1000, seed=1)
    assert np.array_equal(a[0], b[0]) and not np.array_equal(a[0], c[0])
    h1 = scenes.heightfield(64, 32, seed=3)
    h2 = scenes.heightfield(64, 32, seed=3)
    assert np.array_equal(h1[0], h2[0])
    assert np.abs(h1[0][:, 1]).max() <= 0.2


def test_soup_statistics():
    v, i = scenes.soup(20000, seed=0)
    p = v[i]
    c = p.mean(1)
    assert np.all(np.abs(c) < 1.2)
    e = (p - c[:, None, :]).reshape(-1)
    # vertex offsets from the centroid have std 0.02*sqrt(2/3) per component
    assert abs(e.std() - 0.02 * np.sqrt(2 / 3)) < 5e-4


def test_pinhole_tile_order_and_closed_form():
    r = scenes.pinhole_rays((0, 0, -3), (0, 0, -2), (0, 1, 0), 60, 60, 16, 16)
    # first packet is the 8x8 tile at the origin, row-major
    assert list(r["px"][:10]) == [0, 1, 2, 3, 4, 5, 6, 7, 0, 1]
    assert list(r["py"][:10]) == [0] * 8 + [1, 1]
    d = np.stack([r["dx"], r["dy"], r["dz"]], 1)
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-6)
    # the four tiles' corner rays span +-tan(30 deg) * (1 - 1/16) around the z axis
    t = np.tan(np.radians(30)) * (1 - 1 / 16)
    dn = d / d[:, 2:3]
    assert np.isclose(np.abs(dn[:, 0]).max(), t, atol=1e-5)
    assert np.isclose(np.abs(dn[:, 1]).max(), t, atol=1e-5)
