"""Python view of the C ABI: scenes, device buffers and tracing.

Mirrors the reference's AccelerationStructure life cycle (Interface/AccelerationStructure.h:
11-66): create over a mesh (setGroup), build (rebuild with group_changed), refit (update),
load/save the DynBVH binary tree (buildFromFile/saveToFile) and intersect
(Object::intersect, Interface/Object.h:34) over batches of rays in SoA form.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L

RAY_FIELDS = ("ox", "oy", "oz", "dx", "dy", "dz", "tmax")


class DeviceArray:
    """One device allocation (mh_malloc) with numpy-style upload/download."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        L.check(L.load().mh_malloc(C.byref(p), max(1, self.nbytes)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        if a.nbytes:
            L.check(L.load().mh_memcpy_h2d(d.ptr, a.ctypes.data, a.nbytes))
        return d

    def download(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            L.check(L.load().mh_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        L.check(L.load().mh_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes))

    def free(self) -> None:
        if self.ptr:
            L.check(L.load().mh_free(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@dataclass
class DeviceRays:
    """SoA rays resident in device memory."""

    n: int
    arrays: dict

    @classmethod
    def upload(cls, rays: dict) -> "DeviceRays":
        n = len(rays["ox"])
        arrs = {k: DeviceArray.from_numpy(np.asarray(rays[k], np.float32)) for k in RAY_FIELDS}
        if rays.get("masked") is not None:
            arrs["masked"] = DeviceArray.from_numpy(np.asarray(rays["masked"], np.uint8))
        return cls(n, arrs)

    @classmethod
    def alloc(cls, n: int, masked: bool = True) -> "DeviceRays":
        arrs = {k: DeviceArray(4 * n) for k in RAY_FIELDS}
        if masked:
            arrs["masked"] = DeviceArray(n)
        return cls(n, arrs)

    def struct(self) -> L.Rays:
        r = L.Rays()
        for k in RAY_FIELDS:
            setattr(r, k, self.arrays[k].ptr)
        r.masked = self.arrays["masked"].ptr if "masked" in self.arrays else None
        return r

    def buffers(self) -> L.RayBuffers:
        b = L.RayBuffers()
        for k in RAY_FIELDS:
            setattr(b, k, self.arrays[k].ptr)
        b.masked = self.arrays["masked"].ptr if "masked" in self.arrays else None
        return b

    def download(self) -> dict:
        out = {k: self.arrays[k].download(np.float32, self.n) for k in RAY_FIELDS}
        if "masked" in self.arrays:
            out["masked"] = self.arrays["masked"].download(np.uint8, self.n)
        return out


@dataclass
class DeviceHits:
    n: int
    arrays: dict

    @classmethod
    def alloc(cls, n: int, any_hit: bool = False) -> "DeviceHits":
        arrs = {"prim": DeviceArray(4 * n), "t": DeviceArray(4 * n), "u": DeviceArray(4 * n),
                "v": DeviceArray(4 * n)}
        if any_hit:
            arrs["occluded"] = DeviceArray(n)
        return cls(n, arrs)

    def struct(self) -> L.Hits:
        h = L.Hits()
        for k in ("prim", "t", "u", "v", "occluded"):
            setattr(h, k, self.arrays[k].ptr if k in self.arrays else None)
        return h

    def download(self) -> dict:
        out = {"prim": self.arrays["prim"].download(np.int32, self.n),
               "t": self.arrays["t"].download(np.float32, self.n),
               "u": self.arrays["u"].download(np.float32, self.n),
               "v": self.arrays["v"].download(np.float32, self.n)}
        if "occluded" in self.arrays:
            out["occluded"] = self.arrays["occluded"].download(np.uint8, self.n)
        return out


def trace_opts(packet_size: int = 1, stack_lds: int = 0, blocks_per_cu: int = 0,
               counters: DeviceArray | None = None) -> L.TraceOpts:
    o = L.TraceOpts()
    o.packet_size = packet_size
    o.stack_lds = stack_lds
    o.blocks_per_cu = blocks_per_cu
    o.counters = counters.ptr if counters is not None else None
    return o


class Scene:
    """A mesh plus its DynBVH on the device (one mh_scene)."""

    def __init__(self, verts: np.ndarray, idx: np.ndarray):
        lib = L.load()
        self.verts = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
        self.idx = np.ascontiguousarray(idx, np.uint32).reshape(-1, 3)
        h = C.c_void_p()
        L.check(lib.mh_scene_create(self.verts.ctypes.data, len(self.verts), self.idx.ctypes.data,
                                    len(self.idx), C.byref(h)))
        self.h = h.value

    def close(self) -> None:
        if getattr(self, "h", None):
            L.load().mh_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # AccelerationStructure::rebuild
    def build(self, prim_kind: int = L.MH_PRIM_KENSLER_SHIRLEY, rtsah: bool = True,
              threads: int = 0, approximate: bool = False) -> "Scene":
        """DynBVH::rebuild: the full-sweep SAH on the host, or with approximate=True the
        reference's approximate (binned) SAH build, run on the device."""
        o = L.BuildOpts(prim_kind, int(rtsah), threads, int(approximate))
        L.check(L.load().mh_scene_build(self.h, C.byref(o)))
        return self

    # Mesh::vertexNormals/normal_indices, Mesh::texCoords/texture_indices
    def set_attributes(self, normals=None, normal_idx=None, texcoords=None, tex_idx=None):
        keep = []

        def arr(a, dt):
            if a is None:
                return None, 0
            a = np.ascontiguousarray(a, dt)
            keep.append(a)
            return a.ctypes.data, (len(a) if a.ndim > 1 else len(a) // 3)

        n, nn = arr(normals, np.float32)
        ni, _ = arr(normal_idx, np.uint32)
        t, nt = arr(texcoords, np.float32)
        ti, _ = arr(tex_idx, np.uint32)
        L.check(L.load().mh_scene_set_attributes(self.h, n, nn, ni, t, nt, ti))

    def shade(self, rays: "DeviceRays", hits: "DeviceHits", stream=None) -> dict:
        """Hit positions, shading and geometric normals and texture coordinates (host copies)."""
        keys = ("px", "py", "pz", "nx", "ny", "nz", "gx", "gy", "gz", "tu", "tv", "tw")
        bufs = {k: DeviceArray(4 * rays.n) for k in keys}
        o = L.Shading(*[bufs[k].ptr for k in keys])
        r, h = rays.struct(), hits.struct()
        L.check(L.load().mh_shade(self.h, C.byref(r), C.byref(h), rays.n, C.byref(o), stream))
        sync()
        out = {k: bufs[k].download(np.float32, rays.n) for k in keys}
        for b in bufs.values():
            b.free()
        return out

    # AccelerationStructure::buildFromFile (from memory)
    def load_dynbvh(self, ids: np.ndarray, nodes: np.ndarray,
                    prim_kind: int = L.MH_PRIM_KENSLER_SHIRLEY) -> "Scene":
        ids = np.ascontiguousarray(ids, np.int32)
        nodes = np.ascontiguousarray(nodes).view(np.uint8)
        assert nodes.nbytes % 32 == 0
        L.check(L.load().mh_scene_load_dynbvh(self.h, ids.ctypes.data, len(ids), nodes.ctypes.data,
                                              nodes.nbytes // 32, prim_kind))
        return self

    def load_dynbvh_file(self, path: str, prim_kind: int = 0) -> "Scene":
        L.check(L.load().mh_scene_load_dynbvh_file(self.h, str(path).encode(), prim_kind))
        return self

    def save_dynbvh_file(self, path: str) -> None:
        L.check(L.load().mh_scene_save_dynbvh_file(self.h, str(path).encode()))

    def export_dynbvh(self) -> tuple[np.ndarray, np.ndarray]:
        lib = L.load()
        ni, nn = C.c_uint32(), C.c_uint32()
        L.check(lib.mh_scene_export_dynbvh(self.h, None, C.byref(ni), None, C.byref(nn)))
        ids = np.empty(ni.value, np.int32)
        nodes = np.empty(nn.value, L.NODE_DTYPE)
        L.check(lib.mh_scene_export_dynbvh(self.h, ids.ctypes.data, C.byref(ni), nodes.ctypes.data,
                                           C.byref(nn)))
        return ids, nodes

    def bounds(self) -> tuple[np.ndarray, np.ndarray]:
        lo = np.empty(3, np.float32)
        hi = np.empty(3, np.float32)
        L.check(L.load().mh_scene_bounds(self.h, lo.ctypes.data, hi.ctypes.data))
        return lo, hi

    def stats(self) -> dict:
        nn, ni, d, b = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_size_t()
        L.check(L.load().mh_scene_stats(self.h, C.byref(nn), C.byref(ni), C.byref(d), C.byref(b)))
        return {"nodes": nn.value, "inner": ni.value, "depth": d.value, "device_bytes": b.value}

    # AccelerationStructure::update
    def set_vertices(self, verts: np.ndarray, stream=None) -> None:
        v = np.ascontiguousarray(verts, np.float32)
        assert v.size == self.verts.size
        L.check(L.load().mh_scene_set_vertices(self.h, v.ctypes.data, 0, stream))
        self.verts = v.reshape(-1, 3)

    def set_workers(self, workers: int) -> None:
        """Render threads of the Manta instance (numProcs): the isLargeSubtree threshold."""
        L.check(L.load().mh_scene_set_workers(self.h, int(workers)))

    def refit(self, stream=None) -> None:
        """Bounds-only refit (no rotations)."""
        L.check(L.load().mh_scene_refit(self.h, stream))

    def update(self, stream=None) -> None:
        """DynBVH::update: refit with tree rotations and leaf splits (synchronises)."""
        L.check(L.load().mh_scene_update(self.h, stream))

    def displace_sine(self, amp: float, freq: float, phase: float, stream=None) -> None:
        L.check(L.load().mh_scene_displace_sine(self.h, amp, freq, phase, stream))

    def device_vertices(self) -> np.ndarray:
        p = C.c_void_p()
        L.check(L.load().mh_scene_vertices_device(self.h, C.byref(p)))
        out = np.empty(self.verts.size, np.float32)
        L.check(L.load().mh_memcpy_d2h(out.ctypes.data, p.value, out.nbytes))
        return out.reshape(-1, 3)

    # Object::intersect
    def trace(self, rays: DeviceRays, hits: DeviceHits, any_hit: bool = False,
              opts: L.TraceOpts | None = None, stream=None) -> None:
        r, h = rays.struct(), hits.struct()
        L.check(L.load().mh_trace(self.h, C.byref(r), C.byref(h), rays.n,
                                  L.MH_ANY_HIT if any_hit else L.MH_CLOSEST_HIT,
                                  C.byref(opts) if opts is not None else None, stream))

    def trace_host(self, rays: dict, any_hit: bool = False, packet_size: int = 1,
                   stack_lds: int = 0) -> dict:
        n = len(rays["ox"])
        arr = {k: np.ascontiguousarray(rays[k], np.float32) for k in RAY_FIELDS}
        m = rays.get("masked")
        m = np.ascontiguousarray(m, np.uint8) if m is not None else None
        out = {"prim": np.empty(n, np.int32), "t": np.empty(n, np.float32),
               "u": np.empty(n, np.float32), "v": np.empty(n, np.float32)}
        if any_hit:
            out["occluded"] = np.empty(n, np.uint8)
        r = L.Rays(*[arr[k].ctypes.data for k in RAY_FIELDS], L.ptr(m))
        h = L.Hits(out["prim"].ctypes.data, out["t"].ctypes.data, out["u"].ctypes.data,
                   out["v"].ctypes.data, out["occluded"].ctypes.data if any_hit else None)
        o = trace_opts(packet_size, stack_lds)
        L.check(L.load().mh_trace_host(self.h, C.byref(r), C.byref(h), n,
                                       L.MH_ANY_HIT if any_hit else L.MH_CLOSEST_HIT, C.byref(o)))
        return out


class Frame:
    """A fixed table of ray batches traced by one persistent launch (mh_frame_*)."""

    def __init__(self, scene: Scene, batches: list):
        """batches: list of (DeviceRays, DeviceHits, any_hit)."""
        self.scene = scene
        self._keep = list(batches)
        arr = self._table(batches)
        h = C.c_void_p()
        L.check(L.load().mh_frame_create(scene.h, arr, len(batches), C.byref(h)))
        self.h = h.value

    @staticmethod
    def _table(batches):
        arr = (L.Batch * len(batches))()
        for i, (dr, dh, any_hit) in enumerate(batches):
            arr[i].rays = dr.struct()
            arr[i].hits = dh.struct()
            arr[i].count = dr.n
            arr[i].flags = L.MH_ANY_HIT if any_hit else L.MH_CLOSEST_HIT
        return arr

    def set_batches(self, batches: list, stream=None) -> None:
        self._keep = list(batches)
        L.check(L.load().mh_frame_set_batches(self.h, self._table(batches), len(batches), stream))

    def set_config(self, cfg: dict | None) -> None:
        c = None
        if cfg is not None:
            c = L.LaunchConfig.from_dict(cfg)
        L.check(L.load().mh_frame_set_config(self.h, C.byref(c) if c is not None else None))

    def config(self) -> dict:
        c = L.LaunchConfig()
        L.check(L.load().mh_frame_get_config(self.h, C.byref(c)))
        return c.as_dict()

    def launch(self, stream=None) -> None:
        L.check(L.load().mh_frame_launch(self.h, stream))

    def capture(self, stream) -> None:
        L.check(L.load().mh_frame_capture(self.h, stream))

    def replay(self, stream) -> None:
        L.check(L.load().mh_frame_replay(self.h, stream))

    def time(self, iters: int, use_graph: bool = False, stream=None) -> float:
        ms = C.c_float()
        L.check(L.load().mh_frame_time(self.h, iters, int(use_graph), stream, C.byref(ms)))
        return ms.value

    def autotune(self, iters: int = 3, cache_path: str | None = None, force: bool = False):
        c = L.LaunchConfig()
        ms = C.c_float()
        L.check(L.load().mh_frame_autotune(self.h, iters, cache_path.encode() if cache_path else None,
                                           int(force), C.byref(c), C.byref(ms)))
        return c.as_dict(), ms.value

    def close(self) -> None:
        if getattr(self, "h", None):
            L.load().mh_frame_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def camera(eye, lookat, up, vfov_deg, xres, yres, spp=1, seed=0, hfov_deg=0.0) -> L.Camera:
    c = L.Camera()
    c.eye[:] = [float(x) for x in eye]
    c.lookat[:] = [float(x) for x in lookat]
    c.up[:] = [float(x) for x in up]
    c.vfov_deg, c.hfov_deg = vfov_deg, hfov_deg
    c.xres, c.yres, c.spp, c.seed = xres, yres, spp, seed
    return c


def gen_camera_rays(cam: L.Camera, out: DeviceRays, stream=None) -> None:
    assert out.n == cam.xres * cam.yres * cam.spp
    b = out.buffers()
    L.check(L.load().mh_gen_camera_rays(C.byref(cam), C.byref(b), stream))


def gen_shadow_rays(scene: "Scene", primary: DeviceRays, hits: DeviceHits, light, out: DeviceRays,
                    stream=None) -> None:
    lt = (C.c_float * 3)(*[float(x) for x in light])
    r, h, b = primary.struct(), hits.struct(), out.buffers()
    L.check(L.load().mh_gen_shadow_rays(scene.h, C.byref(r), C.byref(h), primary.n, C.byref(lt),
                                        C.byref(b), stream))


def gen_bounce_rays(scene: "Scene", primary: DeviceRays, hits: DeviceHits, seed: int,
                    out: DeviceRays, stream=None) -> None:
    r, h, b = primary.struct(), hits.struct(), out.buffers()
    L.check(L.load().mh_gen_bounce_rays(scene.h, C.byref(r), C.byref(h), primary.n, seed,
                                        C.byref(b), stream))


class Server:
    """Serving loop over resident scenes (mh_server_*): submit camera requests, wait for hits."""

    def __init__(self, scenes: list, hbm_budget: int = 0, host_budget: int = 0,
                 chunk_rays: int = 0, max_active: int = 0, shadows: bool = False,
                 bounces: bool = False, light=(0.0, 0.0, 0.0), launch: dict | None = None):
        self.scenes = list(scenes)
        o = L.ServerOpts()
        o.hbm_budget, o.host_budget, o.chunk_rays = hbm_budget, host_budget, chunk_rays
        o.max_active, o.shadows, o.bounces = max_active, int(shadows), int(bounces)
        o.light[:] = [float(x) for x in light]
        if launch is not None:
            o.launch = L.LaunchConfig.from_dict(launch)
        self.shadows, self.bounces = bool(shadows), bool(bounces)
        arr = (C.c_void_p * len(self.scenes))(*[sc.h for sc in self.scenes])
        h = C.c_void_p()
        L.check(L.load().mh_server_create(arr, len(self.scenes), C.byref(o), C.byref(h)))
        self.h = h.value
        self._sizes = {}

    def submit(self, scene: int, cam: L.Camera) -> int:
        r = L.Request(scene, cam)
        rid = C.c_uint64()
        L.check(L.load().mh_server_submit(self.h, C.byref(r), C.byref(rid)))
        self._sizes[rid.value] = cam.xres * cam.yres * cam.spp
        return rid.value

    def wait(self, rid: int, results: bool = True) -> dict:
        """Waits for request `rid`; returns its info and (results=True) its result arrays."""
        n = self._sizes.pop(rid, 0) if results else 0
        out = {}
        res = L.RequestResult()
        if results:
            out["prim"], out["t"] = np.empty(n, np.int32), np.empty(n, np.float32)
            res.prim, res.t = out["prim"].ctypes.data, out["t"].ctypes.data
            if self.shadows:
                out["occluded"] = np.empty(n, np.uint8)
                res.occluded = out["occluded"].ctypes.data
            if self.bounces:
                out["bounce_prim"] = np.empty(n, np.int32)
                out["bounce_t"] = np.empty(n, np.float32)
                res.bounce_prim = out["bounce_prim"].ctypes.data
                res.bounce_t = out["bounce_t"].ctypes.data
        else:
            self._sizes.pop(rid, None)
        info = L.RequestInfo()
        L.check(L.load().mh_server_wait(self.h, rid, C.byref(res), C.byref(info)))
        out.update(id=info.id, status=info.status, rays=info.rays, submit_ms=info.submit_ms,
                   admit_ms=info.admit_ms, done_ms=info.done_ms)
        return out

    def cancel(self, rid: int) -> None:
        """Cancels request `rid` (queued: at once; in flight: behind its queued work)."""
        L.check(L.load().mh_server_cancel(self.h, rid))

    def stats(self) -> dict:
        st = L.ServerStats()
        L.check(L.load().mh_server_get_stats(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in L.ServerStats._fields_}

    def clock_ms(self) -> float:
        return L.load().mh_server_clock_ms(self.h)

    def close(self) -> None:
        if getattr(self, "h", None):
            L.load().mh_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def stream_create():
    s = C.c_void_p()
    L.check(L.load().mh_stream_create(C.byref(s)))
    return s.value


def stream_destroy(s) -> None:
    L.check(L.load().mh_stream_destroy(s))


def build_dynbvh_host(verts: np.ndarray, idx: np.ndarray, rtsah: bool = True,
                      threads: int = 0) -> tuple[np.ndarray, np.ndarray, int]:
    """Host SAH build (no device): (nodes[2T], ids[T], records in use)."""
    v = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
    i = np.ascontiguousarray(idx, np.uint32).reshape(-1, 3)
    T = len(i)
    ids = np.empty(T, np.int32)
    nodes = np.empty(2 * T, L.NODE_DTYPE)
    used = C.c_uint32()
    o = L.BuildOpts(0, int(rtsah), threads, 0)
    L.check(L.load().mh_dynbvh_build_host(v.ctypes.data, len(v), i.ctypes.data, T, C.byref(o),
                                          ids.ctypes.data, nodes.ctypes.data, C.byref(used)))
    return nodes, ids, used.value


def read_ply(path: str, transform=None):
    """readPlyFile: (verts float32[V, 3], idx uint32[T, 3]) of a PLY file, as Manta builds
    the Mesh (fan triangulation, degenerate triangles removed). transform: 3x4 or None."""
    lib = L.load()
    xf = None if transform is None else np.ascontiguousarray(transform, np.float32).reshape(3, 4)
    nv, nt = C.c_uint32(), C.c_uint32()
    p = str(path).encode()
    L.check(lib.mh_ply_read(p, None if xf is None else xf.ctypes.data, None, 0, None, 0,
                            C.byref(nv), C.byref(nt)))
    v = np.empty((nv.value, 3), np.float32)
    i = np.empty((nt.value, 3), np.uint32)
    L.check(lib.mh_ply_read(p, None if xf is None else xf.ctypes.data, v.ctypes.data, nv.value,
                            i.ctypes.data, nt.value, C.byref(nv), C.byref(nt)))
    return v, i


def read_obj(path: str) -> dict:
    """ObjGroup's Mesh of a Wavefront OBJ file: verts float32[V, 3], idx uint32[T, 3],
    normals float32[N, 3] and normal_idx uint32[T, 3] (None without normals), texcoords
    float32[M, 3] and tex_idx uint32[T, 3] (None without texture coordinates),
    face_material uint32[T] (0 = the default material, k = material_names[k])."""
    lib = L.load()
    m = C.POINTER(L.ObjMesh)()
    L.check(lib.mh_obj_read(str(path).encode(), C.byref(m)))
    try:
        o = m.contents

        def arr(p, n, cols, dt):
            shape = (n, cols) if cols > 1 else (n,)
            if n == 0:
                return np.empty(shape, dt)
            return np.ctypeslib.as_array(p, shape=shape).astype(dt, copy=True)

        out = {
            "verts": arr(o.verts, o.n_verts, 3, np.float32),
            "idx": arr(o.idx, o.n_tris, 3, np.uint32),
            "face_material": arr(o.face_material, o.n_tris, 1, np.uint32),
            "material_names": [o.material_names[k].decode() for k in range(o.n_materials)],
            "normals": None, "normal_idx": None, "texcoords": None, "tex_idx": None,
        }
        if o.normals:
            out["normals"] = arr(o.normals, o.n_normals, 3, np.float32)
            out["normal_idx"] = arr(o.normal_idx, o.n_tris, 3, np.uint32)
        if o.texcoords:
            out["texcoords"] = arr(o.texcoords, o.n_texcoords, 3, np.float32)
            out["tex_idx"] = arr(o.tex_idx, o.n_tris, 3, np.uint32)
        return out
    finally:
        lib.mh_obj_free(m)


def sync(stream=None) -> None:
    L.check(L.load().mh_stream_sync(stream))


def device_info() -> dict:
    name = C.create_string_buffer(256)
    cu = C.c_int()
    hbm = C.c_size_t()
    L.check(L.load().mh_device_info(name, 256, C.byref(cu), C.byref(hbm)))
    return {"arch": name.value.decode(), "cu_count": cu.value, "hbm_bytes": hbm.value}
