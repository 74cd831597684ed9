"""Synthetic workloads: meshes, cameras and secondary rays.

Project-defined generators (the reference ships no scenes for this path). Every random
number comes from PCG32 (O'Neill, "PCG: A Family of Simple Fast Space-Efficient
Statistically Good Algorithms for Random Number Generation", 2014) with one stream per
element: element i draws from pcg32(seed, stream=i), which makes every array
reproducible in any order and vectorises in numpy.

* triangle soup: centroid c ~ U[-1,1]^3, vertex k = c + e_k with e_k ~ N(0, 0.02^2) per
  component (Box-Muller), optional uniform scale;
* heightfield: an nx x nz vertex grid over x in [-1,1] with square cells, heights a
  6-octave value-noise fBm of amplitude 0.2, two triangles per cell;
* pinhole camera rays after Manta's PinholeCamera (Model/Cameras/PinholeCamera.cc:
  167-230): dir = v*x + u*y + (lookat - eye), normalised, in 8x8 tiles (the square
  fragments of Engine/ImageTraversers/TiledImageTraverser.cc:255-296);
* shadow rays after PointLight/HardShadows (Model/Lights/PointLight.cc:31-75,
  Engine/Shadows/HardShadows.cc:87-124): origin at the hit point, dir = L - P
  unnormalised, minT = 1 - T_EPSILON, back-facing rays masked;
* cosine-weighted diffuse bounce rays about the face-forward geometric normal.
"""

from __future__ import annotations

import numpy as np

MULT = np.uint64(6364136223846793005)
T_EPSILON = np.float32(1.0e-4)
MAXT = np.float32(1.0e19)


def pcg32(seed: int, streams: np.ndarray, n_draws: int) -> np.ndarray:
    """(n_draws, len(streams)) uint32 outputs of PCG32 srandom(seed, stream)."""
    streams = np.asarray(streams, np.uint64)
    with np.errstate(over="ignore"):
        inc = (streams << np.uint64(1)) | np.uint64(1)
        state = np.zeros_like(streams)
        state = state * MULT + inc
        state = state + np.uint64(seed)
        state = state * MULT + inc
        out = np.empty((n_draws, len(streams)), np.uint32)
        for k in range(n_draws):
            old = state
            state = old * MULT + inc
            xs = (((old >> np.uint64(18)) ^ old) >> np.uint64(27)) & np.uint64(0xFFFFFFFF)
            rot = old >> np.uint64(59)
            r = ((xs >> rot) | (xs << ((np.uint64(32) - rot) & np.uint64(31)))) & np.uint64(0xFFFFFFFF)
            out[k] = r.astype(np.uint32)
    return out


def u01(x: np.ndarray) -> np.ndarray:
    """uint32 -> float64 in [0,1) with 24 random bits."""
    return (x >> np.uint32(8)).astype(np.float64) * (1.0 / (1 << 24))


def soup(n_tris: int, seed: int = 0, sigma: float = 0.02, scale: float = 1.0):
    """Triangle soup. Returns (verts (3T,3) f32, idx (T,3) u32)."""
    r = pcg32(seed, np.arange(n_tris, dtype=np.uint64), 13)
    c = 2.0 * u01(r[0:3]).T - 1.0  # (T,3)
    a = u01(r[3:8])
    b = u01(r[8:13])
    rad = np.sqrt(-2.0 * np.log(1.0 - a))
    g = np.concatenate([rad * np.cos(2 * np.pi * b), rad * np.sin(2 * np.pi * b)], 0)[:9].T
    e = (sigma * g).reshape(n_tris, 3, 3)
    verts = (scale * (c[:, None, :] + e)).reshape(-1, 3).astype(np.float32)
    idx = np.arange(3 * n_tris, dtype=np.uint32).reshape(n_tris, 3)
    return verts, idx


def _value_noise_octave(gx: np.ndarray, gz: np.ndarray, seed: int, octave: int) -> np.ndarray:
    ix = np.floor(gx).astype(np.int64)
    iz = np.floor(gz).astype(np.int64)
    fx = gx - ix
    fz = gz - iz
    sx = fx * fx * (3 - 2 * fx)
    sz = fz * fz * (3 - 2 * fz)
    wx = int(ix.max()) + 2

    def lattice(i, j):
        sid = (np.uint64(octave) << np.uint64(40)) + (j.astype(np.uint64) * np.uint64(wx) +
                                                      i.astype(np.uint64))
        flat = sid.ravel()
        uniq, inv = np.unique(flat, return_inverse=True)
        vals = 2.0 * u01(pcg32(seed, uniq, 1)[0]) - 1.0
        return vals[inv].reshape(i.shape)

    v00 = lattice(ix, iz)
    v10 = lattice(ix + 1, iz)
    v01 = lattice(ix, iz + 1)
    v11 = lattice(ix + 1, iz + 1)
    a = v00 + (v10 - v00) * sx
    b = v01 + (v11 - v01) * sx
    return a + (b - a) * sz


def heightfield(nx: int, nz: int, seed: int = 0, octaves: int = 6, amplitude: float = 0.2,
                base_freq: float = 4.0):
    """Grid heightfield. Returns (verts (nx*nz,3) f32, idx (2(nx-1)(nz-1),3) u32)."""
    xs = np.linspace(-1.0, 1.0, nx)
    half = (nz - 1) / (nx - 1)
    zs = np.linspace(-half, half, nz)
    X, Z = np.meshgrid(xs, zs)  # (nz, nx)
    h = np.zeros_like(X)
    norm = 0.0
    for o in range(octaves):
        f = base_freq * (2 ** o)
        w = 0.5 ** o
        h += w * _value_noise_octave((X + 1) * 0.5 * f, (Z + half) * 0.5 * f, seed, o)
        norm += w
    Y = amplitude * h / norm
    verts = np.stack([X, Y, Z], -1).reshape(-1, 3).astype(np.float32)
    j, i = np.meshgrid(np.arange(nz - 1), np.arange(nx - 1), indexing="ij")
    v00 = (j * nx + i).ravel()
    v10 = v00 + 1
    v01 = v00 + nx
    v11 = v01 + 1
    t0 = np.stack([v00, v10, v11], -1)
    t1 = np.stack([v00, v11, v01], -1)
    idx = np.stack([t0, t1], 1).reshape(-1, 3).astype(np.uint32)
    return verts, idx


def pinhole_rays(eye, lookat, up, hfov_deg: float, vfov_deg: float, xres: int, yres: int,
                 tile: int = 8, normalize: bool = True) -> dict:
    """Camera rays in tile order (tile-major, then row-major inside a tile)."""
    eye = np.asarray(eye, np.float64)
    direction = np.asarray(lookat, np.float64) - eye
    nearz = np.linalg.norm(direction)
    upn = np.asarray(up, np.float64)
    upn = upn / np.linalg.norm(upn)
    v = np.cross(direction, upn)
    v /= np.linalg.norm(v)
    u = np.cross(v, direction)
    u /= np.linalg.norm(u)
    u *= nearz * np.tan(np.radians(0.5 * vfov_deg))
    v *= nearz * np.tan(np.radians(0.5 * hfov_deg))
    assert xres % tile == 0 and yres % tile == 0
    ty, tx, j, i = np.meshgrid(np.arange(yres // tile), np.arange(xres // tile), np.arange(tile),
                               np.arange(tile), indexing="ij")
    px = (tx * tile + i).ravel()
    py = (ty * tile + j).ravel()
    x = -1.0 + (px + 0.5) * (2.0 / xres)
    y = -1.0 + (py + 0.5) * (2.0 / yres)
    d = v[None, :] * x[:, None] + u[None, :] * y[:, None] + direction[None, :]
    if normalize:
        d /= np.linalg.norm(d, axis=1, keepdims=True)
    n = len(px)
    d = d.astype(np.float32)
    return {"ox": np.full(n, eye[0], np.float32), "oy": np.full(n, eye[1], np.float32),
            "oz": np.full(n, eye[2], np.float32), "dx": np.ascontiguousarray(d[:, 0]),
            "dy": np.ascontiguousarray(d[:, 1]), "dz": np.ascontiguousarray(d[:, 2]),
            "tmax": np.full(n, MAXT, np.float32), "px": px.astype(np.int32),
            "py": py.astype(np.int32)}


def geometric_normals(verts: np.ndarray, idx: np.ndarray, prim: np.ndarray) -> np.ndarray:
    """KenslerShirleyTriangle::computeGeometricNormal, Cross(p1-p0, p2-p0), in float32."""
    p = verts[idx[prim]]  # (n,3,3)
    return np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]).astype(np.float32)


def hit_points(rays: dict, t: np.ndarray) -> np.ndarray:
    o = np.stack([rays["ox"], rays["oy"], rays["oz"]], 1)
    d = np.stack([rays["dx"], rays["dy"], rays["dz"]], 1)
    return (o + d * t[:, None].astype(np.float32)).astype(np.float32)


def _face_forward(n: np.ndarray, d: np.ndarray) -> np.ndarray:
    s = np.sign(np.sum(n * d, axis=1, dtype=np.float32))
    return np.where((s > 0)[:, None], -n, n)


def shadow_rays(verts, idx, prim_rays: dict, hits: dict, light, compact: bool) -> dict:
    """One shadow ray per primary hit toward a point light (dir = L - P, unnormalised)."""
    hit = hits["prim"] >= 0
    n_all = len(hits["prim"])
    sel = np.nonzero(hit)[0] if compact else np.arange(n_all)
    prim = np.where(hit, hits["prim"], 0)[sel]
    P = hit_points({k: prim_rays[k][sel] for k in ("ox", "oy", "oz", "dx", "dy", "dz")},
                   hits["t"][sel])
    L = np.asarray(light, np.float32)[None, :]
    d = (L - P).astype(np.float32)
    dprim = np.stack([prim_rays["dx"][sel], prim_rays["dy"][sel], prim_rays["dz"][sel]], 1)
    ffn = _face_forward(geometric_normals(verts, idx, prim), dprim)
    facing = np.sum(d * ffn, axis=1, dtype=np.float32) > 0
    masked = ~(facing & hit[sel])
    n = len(sel)
    return {"ox": np.ascontiguousarray(P[:, 0]), "oy": np.ascontiguousarray(P[:, 1]),
            "oz": np.ascontiguousarray(P[:, 2]), "dx": np.ascontiguousarray(d[:, 0]),
            "dy": np.ascontiguousarray(d[:, 1]), "dz": np.ascontiguousarray(d[:, 2]),
            "tmax": np.full(n, np.float32(1) - T_EPSILON, np.float32),
            "masked": masked.astype(np.uint8), "src": sel.astype(np.int64)}


def bounce_rays(verts, idx, prim_rays: dict, hits: dict, seed: int) -> dict:
    """One cosine-weighted diffuse bounce ray per primary hit (compacted)."""
    sel = np.nonzero(hits["prim"] >= 0)[0]
    prim = hits["prim"][sel]
    P = hit_points({k: prim_rays[k][sel] for k in ("ox", "oy", "oz", "dx", "dy", "dz")},
                   hits["t"][sel])
    dprim = np.stack([prim_rays["dx"][sel], prim_rays["dy"][sel], prim_rays["dz"][sel]], 1)
    n = _face_forward(geometric_normals(verts, idx, prim), dprim).astype(np.float64)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    r = pcg32(seed, sel.astype(np.uint64), 2)
    u1, u2 = u01(r[0]), u01(r[1])
    rad = np.sqrt(u1)
    lx, ly, lz = rad * np.cos(2 * np.pi * u2), rad * np.sin(2 * np.pi * u2), np.sqrt(1.0 - u1)
    a = np.where((np.abs(n[:, 0]) > 0.9)[:, None], [[0.0, 1.0, 0.0]], [[1.0, 0.0, 0.0]])
    t1 = np.cross(n, a)
    t1 /= np.linalg.norm(t1, axis=1, keepdims=True)
    t2 = np.cross(n, t1)
    d = (t1 * lx[:, None] + t2 * ly[:, None] + n * lz[:, None]).astype(np.float32)
    m = len(sel)
    return {"ox": np.ascontiguousarray(P[:, 0]), "oy": np.ascontiguousarray(P[:, 1]),
            "oz": np.ascontiguousarray(P[:, 2]), "dx": np.ascontiguousarray(d[:, 0]),
            "dy": np.ascontiguousarray(d[:, 1]), "dz": np.ascontiguousarray(d[:, 2]),
            "tmax": np.full(m, MAXT, np.float32), "src": sel.astype(np.int64)}


def random_rays(n: int, seed: int, box_lo, box_hi, tmax=MAXT) -> dict:
    """Rays with origins uniform in a box and uniformly distributed unit directions."""
    r = pcg32(seed, np.arange(n, dtype=np.uint64), 5)
    lo = np.asarray(box_lo, np.float64)
    hi = np.asarray(box_hi, np.float64)
    o = lo + (hi - lo) * u01(r[0:3]).T
    z = 2.0 * u01(r[3]) - 1.0
    phi = 2 * np.pi * u01(r[4])
    s = np.sqrt(np.maximum(0.0, 1 - z * z))
    d = np.stack([s * np.cos(phi), s * np.sin(phi), z], 1)
    o = o.astype(np.float32)
    d = d.astype(np.float32)
    return {"ox": np.ascontiguousarray(o[:, 0]), "oy": np.ascontiguousarray(o[:, 1]),
            "oz": np.ascontiguousarray(o[:, 2]), "dx": np.ascontiguousarray(d[:, 0]),
            "dy": np.ascontiguousarray(d[:, 1]), "dz": np.ascontiguousarray(d[:, 2]),
            "tmax": np.full(n, tmax, np.float32)}


# Named workloads ---------------------------------------------------------------------

def config0_scene():
    """10,000-triangle soup, seed 0."""
    return soup(10_000, seed=0)


def config0_camera(xres: int = 512, yres: int = 512) -> dict:
    return pinhole_rays((0, 0, -3), (0, 0, -2), (0, 1, 0), 60.0, 60.0, xres, yres)


CONFIG0_LIGHT = (2.0, 2.0, -2.0)


def config1_scene(nx: int = 1000, nz: int = 500, seed: int = 0):
    """1000 x 500 vertex-grid heightfield: 997,002 triangles."""
    return heightfield(nx, nz, seed=seed)


# Launch configuration of the C1 frame (primary + bounce batches in one k_trace_batch):
# the round refill (a lane takes a new ray when 32 of its wave's lanes are idle, rays
# handed out 256 at a time) from the end of the table, so the incoherent bounce rays go
# first and the frame ends on short primary rays; 10 workgroups of 128 lanes per CU (five
# waves per SIMD), LDS ring of 8, PAIR nodes. In table order the same schedule at 12
# workgroups per CU was 1.5-1.7% ahead of the static stride; handed out in reverse it is
# 6-7% faster, and 2% more at 10 workgroups per CU.
C1_LAUNCH = {"layout": 0, "stack_lds": 8, "block": 128, "blocks_per_cu": 10, "refill": 32,
             "chunk": 256, "xcd_remap": 0, "reverse": 1}
# The static stride (24 workgroups of 256 lanes per CU), in table order.
C1_LAUNCH_STATIC = {"layout": 0, "stack_lds": 8, "block": 256, "blocks_per_cu": 24,
                    "refill": 64, "chunk": 0, "xcd_remap": 0, "reverse": 0}


def config1_camera(xres: int = 1920, yres: int = 1080) -> dict:
    vfov = 36.0
    hfov = float(np.degrees(2 * np.arctan(np.tan(np.radians(vfov / 2)) * xres / yres)))
    return pinhole_rays((0.0, 1.6, -0.6), (0.0, 0.0, 0.0), (0, 1, 0), hfov, vfov, xres, yres)


CONFIG1_LIGHT = (2.0, 2.0, -2.0)


def surface_points(verts, idx, n: int, seed: int, stream0: int):
    """n uniformly chosen (triangle, barycentric) points; returns (P, tri, face-up normal)."""
    T = len(idx)
    r = pcg32(seed, np.arange(stream0, stream0 + n, dtype=np.uint64), 3)
    tri = (u01(r[0]) * T).astype(np.int64).clip(0, T - 1)
    a, b = u01(r[1]), u01(r[2])
    flip = a + b > 1
    a = np.where(flip, 1 - a, a)
    b = np.where(flip, 1 - b, b)
    p = verts[idx[tri]].astype(np.float64)
    P = p[:, 0] * (1 - a - b)[:, None] + p[:, 1] * a[:, None] + p[:, 2] * b[:, None]
    nrm = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    nrm = np.where((nrm[:, 1] < 0)[:, None], -nrm, nrm)  # heightfield: face the sky
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    return P, tri, nrm


def _soa(o, d, tmax):
    o = o.astype(np.float32)
    d = d.astype(np.float32)
    return {"ox": np.ascontiguousarray(o[:, 0]), "oy": np.ascontiguousarray(o[:, 1]),
            "oz": np.ascontiguousarray(o[:, 2]), "dx": np.ascontiguousarray(d[:, 0]),
            "dy": np.ascontiguousarray(d[:, 1]), "dz": np.ascontiguousarray(d[:, 2]),
            "tmax": np.full(len(o), tmax, np.float32)}


def batch_rays(kind: str, n: int, verts, idx, seed: int, stream0: int, light=CONFIG1_LIGHT,
               eye=(0.0, 1.6, -0.6), lookat=(0.0, 0.0, 0.0)) -> dict:
    """One ray batch of the batching workload.

    primary: pinhole rays from `eye` through uniform image positions (the camera of the
    1M-triangle frame); shadow: from uniform surface points toward the light, unnormalised,
    minT = 1 - T_EPSILON; bounce: cosine-weighted about the surface normal, minT = MAXT.
    """
    if kind == "primary":
        r = pcg32(seed, np.arange(stream0, stream0 + n, dtype=np.uint64), 2)
        x = 2 * u01(r[0]) - 1
        y = 2 * u01(r[1]) - 1
        e = np.asarray(eye, np.float64)
        direction = np.asarray(lookat, np.float64) - e
        nz = np.linalg.norm(direction)
        v = np.cross(direction, [0.0, 1.0, 0.0])
        v /= np.linalg.norm(v)
        u = np.cross(v, direction)
        u /= np.linalg.norm(u)
        vf = 36.0
        hf = np.degrees(2 * np.arctan(np.tan(np.radians(vf / 2)) * 16 / 9))
        u *= nz * np.tan(np.radians(vf / 2))
        v *= nz * np.tan(np.radians(hf / 2))
        d = v[None] * x[:, None] + u[None] * y[:, None] + direction[None]
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        return _soa(np.broadcast_to(e, d.shape), d, MAXT)
    P, tri, nrm = surface_points(verts, idx, n, seed, stream0)
    if kind == "shadow":
        d = np.asarray(light, np.float64)[None] - P
        return _soa(P, d, np.float32(1) - T_EPSILON)
    r = pcg32(seed, np.arange(stream0, stream0 + n, dtype=np.uint64) + np.uint64(1 << 40), 2)
    u1, u2 = u01(r[0]), u01(r[1])
    rad = np.sqrt(u1)
    lx, ly, lz = rad * np.cos(2 * np.pi * u2), rad * np.sin(2 * np.pi * u2), np.sqrt(1 - u1)
    a = np.where((np.abs(nrm[:, 0]) > 0.9)[:, None], [[0.0, 1.0, 0.0]], [[1.0, 0.0, 0.0]])
    t1 = np.cross(nrm, a)
    t1 /= np.linalg.norm(t1, axis=1, keepdims=True)
    t2 = np.cross(nrm, t1)
    d = t1 * lx[:, None] + t2 * ly[:, None] + nrm * lz[:, None]
    return _soa(P, d, MAXT)


def config2_batches(verts, idx, n_batches: int = 64, lo: int = 1_000, hi: int = 2_000_000,
                    seed: int = 1):
    """Batching workload: sizes log-uniform in [lo, hi], types 40% primary, 30% shadow
    (any-hit), 30% diffuse bounce. Returns [(kind, rays)]."""
    r = pcg32(seed, np.arange(n_batches, dtype=np.uint64) + np.uint64(1 << 50), 2)
    sizes = np.exp(np.log(lo) + (np.log(hi) - np.log(lo)) * u01(r[0])).astype(np.int64)
    tu = u01(r[1])
    kinds = np.where(tu < 0.4, "primary", np.where(tu < 0.7, "shadow", "bounce"))
    out = []
    start = 0
    for b in range(n_batches):
        out.append((str(kinds[b]), batch_rays(str(kinds[b]), int(sizes[b]), verts, idx, seed, start)))
        start += int(sizes[b])
    return out


def c3_requests(n: int, seed: int = 2):
    """Serving workload: n requests (scene, camera keyword arguments) from PCG32(seed).
    Scene 0 is the 1M-triangle heightfield, scene 1 the 8M-triangle soup in [-4,4]^3; each
    request has a random orbit camera at 256^2, 512^2, 1024x768 or 1920x1080 with 1-4
    samples per pixel."""
    r = pcg32(seed, np.arange(n, dtype=np.uint64), 6)
    u = [u01(r[k]) for k in range(6)]
    res = [(256, 256), (512, 512), (1024, 768), (1920, 1080)]
    out = []
    for i in range(n):
        sc = int(u[0][i] < 0.5)
        xr, yr = res[min(3, int(u[1][i] * 4))]
        spp = 1 + min(3, int(u[2][i] * 4))
        az = 2 * np.pi * u[3][i]
        el = 0.35 + 0.5 * u[4][i]
        if sc == 0:  # heightfield in [-1,1] x [-0.5,0.5]
            rad, look = 1.6, (0.0, 0.0, 0.0)
        else:        # soup in [-4,4]^3
            rad, look = 11.0, (0.0, 0.0, 0.0)
        eye = (rad * np.cos(el) * np.cos(az), rad * np.sin(el), rad * np.cos(el) * np.sin(az))
        out.append((sc, dict(eye=eye, lookat=look, up=(0, 1, 0), vfov_deg=40.0, xres=xr,
                             yres=yr, spp=spp, seed=i)))
    return out


C3_LIGHT = (2.0, 6.0, -2.0)
C4_CAMERA = dict(eye=(0.0, 0.7, -0.9), lookat=(0.0, 0.0, 0.0), up=(0, 1, 0), vfov_deg=30.0,
                 xres=1920, yres=1080, seed=4)
