"""Device ray generation against float32 / float64 restatements in numpy."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from rfmanta_amd import api, scenes  # noqa: E402


def camera_reference(eye, lookat, up, vfov, xres, yres):
    """Camera rays in float32 with the device's operation order (pixel centres)."""
    w = np.asarray(lookat, np.float64) - np.asarray(eye, np.float64)
    nearz = np.sqrt(np.sum(w * w))
    v = np.cross(w, np.asarray(up, np.float64))
    v /= np.sqrt(np.sum(v * v))
    u = np.cross(v, w)
    u /= np.sqrt(np.sum(u * u))
    tv = np.tan(0.5 * vfov * np.pi / 180.0)
    th = tv * xres / yres
    U = (u * nearz * tv).astype(np.float32)
    V = (v * nearz * th).astype(np.float32)
    W = w.astype(np.float32)
    r = scenes.pinhole_rays(eye, lookat, up, 60, 60, xres, yres)  # for the tile order only
    px, py = r["px"].astype(np.float32), r["py"].astype(np.float32)
    f = np.float32
    x = f(-1) + (px + f(0.5)) * (f(2) / f(xres))
    y = f(-1) + (py + f(0.5)) * (f(2) / f(yres))
    d = [V[k] * x + U[k] * y + W[k] for k in range(3)]
    n2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
    inv = f(1) / np.sqrt(n2)
    return [dk * inv for dk in d]


def test_camera_rays_match_closed_form():
    eye, look, up = (0.0, 1.6, -0.6), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0)
    cam = api.camera(eye, look, up, 36.0, 64, 48)
    out = api.DeviceRays.alloc(64 * 48, masked=False)
    api.gen_camera_rays(cam, out)
    api.sync()
    g = out.download()
    ref = camera_reference(eye, look, up, 36.0, 64, 48)
    for k, name in enumerate(("dx", "dy", "dz")):
        assert np.allclose(g[name], ref[k], rtol=0, atol=4e-7), name
    assert np.all(g["ox"] == np.float32(eye[0])) and np.all(g["tmax"] == np.float32(1e19))


def test_camera_multisample_jitter():
    cam = api.camera((0, 0, -3), (0, 0, -2), (0, 1, 0), 60.0, 16, 16, spp=4, seed=7)
    out = api.DeviceRays.alloc(16 * 16 * 4, masked=False)
    api.gen_camera_rays(cam, out)
    api.sync()
    g = out.download()
    d = np.stack([g["dx"], g["dy"], g["dz"]], 1)
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-6)
    # the four samples of a pixel differ but stay inside that pixel's cone
    s = d.reshape(4, 256, 3)
    assert not np.allclose(s[0], s[1])
    x = s[..., 0] / s[..., 2]
    assert np.all(np.ptp(x, axis=0) < 2 * np.tan(np.radians(30)) / 16 * 1.01)


def test_shadow_and_bounce_rays(oracle):
    v, i = scenes.heightfield(120, 60, seed=2)
    rays = scenes.pinhole_rays((0, 0.9, -1.1), (0, 0, 0), (0, 1, 0), 60, 60, 64, 64)
    with api.Scene(v, i) as s:
        s.build()
        dp = api.DeviceRays.upload(rays)
        dh = api.DeviceHits.alloc(len(rays["ox"]))
        s.trace(dp, dh)
        api.sync()
        hits = dh.download()
        sh = api.DeviceRays.alloc(dp.n)
        api.gen_shadow_rays(s, dp, dh, (2, 2, -2), sh)
        bo = api.DeviceRays.alloc(dp.n)
        api.gen_bounce_rays(s, dp, dh, 5, bo)
        api.sync()
        g_sh = sh.download()
        g_bo = bo.download()
    ref = scenes.shadow_rays(v, i, rays, hits, (2, 2, -2), compact=False)
    hit = hits["prim"] >= 0
    assert hit.sum() > 1000
    for k in ("ox", "oy", "oz", "dx", "dy", "dz"):
        assert np.allclose(g_sh[k][hit], ref[k][hit], atol=2e-6), k
    agree = g_sh["masked"] == ref["masked"]
    assert agree.mean() > 0.999
    assert np.all(g_sh["masked"][~hit] == 1)
    # bounce: unit directions in the upper hemisphere of the face-forward normal
    d = np.stack([g_bo["dx"], g_bo["dy"], g_bo["dz"]], 1)[hit]
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-5)
    rb = scenes.bounce_rays(v, i, rays, hits, 5)
    # the numpy generator draws from the same PCG32 streams (ray index), in float64
    ref_d = np.stack([rb["dx"], rb["dy"], rb["dz"]], 1)
    assert np.allclose(d, ref_d, atol=2e-5)
    assert np.all(g_bo["masked"][~hit] == 1) and np.all(g_bo["masked"][hit] == 0)
