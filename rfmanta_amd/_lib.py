"""ctypes bindings of libmh.so (include/mh.h).

The library is the only compute path: if libmh.so is missing this module raises at
import time instead of falling back to anything else.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libmh.so"

MH_OK = 0
MH_E_INVALID, MH_E_HIP, MH_E_NOMEM, MH_E_STATE, MH_E_UNSUPPORTED, MH_E_IO = -1, -2, -3, -4, -5, -6
MH_E_CANCELED = -7
MH_CLOSEST_HIT = 0
MH_ANY_HIT = 1
MH_PRIM_KENSLER_SHIRLEY = 0
MH_PRIM_WALD = 1
MH_NO_HIT = -1
MH_MASKED = -2
T_EPSILON = np.float32(1.0e-4)
MAXT = np.float32(1.0e19)

# Reference BVHNode record (Model/Groups/DynBVH.h:37-77), 32 bytes.
NODE_DTYPE = np.dtype([("lo", "<f4", (3,)), ("hi", "<f4", (3,)), ("child", "<i4"),
                       ("axis", "u1"), ("bits", "u1"), ("children", "<i2")])
assert NODE_DTYPE.itemsize == 32


class MhError(RuntimeError):
    """A libmh entry point returned a nonzero status."""

    def __init__(self, code: int, message: str):
        super().__init__(f"libmh error {code}: {message}")
        self.code = code


class Rays(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("ox", "oy", "oz", "dx", "dy", "dz", "tmax", "masked")]


class Hits(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("prim", "t", "u", "v", "occluded")]


class TraceOpts(C.Structure):
    _fields_ = [("packet_size", C.c_int), ("stack_lds", C.c_int), ("blocks_per_cu", C.c_int),
                ("counters", C.c_void_p)]


class Batch(C.Structure):
    _fields_ = [("rays", Rays), ("hits", Hits), ("count", C.c_int64), ("flags", C.c_uint32)]


class LaunchConfig(C.Structure):
    _fields_ = [("layout", C.c_int), ("stack_lds", C.c_int), ("block", C.c_int),
                ("blocks_per_cu", C.c_int), ("refill", C.c_int), ("chunk", C.c_int),
                ("xcd_remap", C.c_int), ("reverse", C.c_int)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}

    @classmethod
    def from_dict(cls, d: dict) -> "LaunchConfig":
        return cls(d["layout"], d["stack_lds"], d["block"], d["blocks_per_cu"], d["refill"],
                   d.get("chunk", 64), d.get("xcd_remap", 0), d.get("reverse", 0))


MH_LAYOUT_PAIR = 0
MH_LAYOUT_REF = 1
MH_LAYOUT_SOA = 2


class RayBuffers(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("ox", "oy", "oz", "dx", "dy", "dz", "tmax", "masked")]


class Camera(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("lookat", C.c_float * 3), ("up", C.c_float * 3),
                ("vfov_deg", C.c_float), ("hfov_deg", C.c_float), ("xres", C.c_int),
                ("yres", C.c_int), ("spp", C.c_int), ("seed", C.c_uint64)]


class Shading(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("px", "py", "pz", "nx", "ny", "nz", "gx", "gy", "gz",
                                         "tu", "tv", "tw")]


class ObjMesh(C.Structure):
    _fields_ = [("n_verts", C.c_uint32), ("n_tris", C.c_uint32), ("n_normals", C.c_uint32),
                ("n_texcoords", C.c_uint32), ("n_materials", C.c_uint32),
                ("verts", C.POINTER(C.c_float)), ("idx", C.POINTER(C.c_uint32)),
                ("normals", C.POINTER(C.c_float)), ("normal_idx", C.POINTER(C.c_uint32)),
                ("texcoords", C.POINTER(C.c_float)), ("tex_idx", C.POINTER(C.c_uint32)),
                ("face_material", C.POINTER(C.c_uint32)),
                ("material_names", C.POINTER(C.c_char_p))]


class BuildOpts(C.Structure):
    _fields_ = [("prim_kind", C.c_int), ("rtsah", C.c_int), ("threads", C.c_int),
                ("approximate", C.c_int)]


class ServerOpts(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("host_budget", C.c_size_t),
                ("chunk_rays", C.c_int64), ("max_active", C.c_int), ("shadows", C.c_int),
                ("bounces", C.c_int), ("light", C.c_float * 3), ("launch", LaunchConfig)]


class Request(C.Structure):
    _fields_ = [("scene", C.c_int), ("camera", Camera)]


class RequestResult(C.Structure):
    _fields_ = [("prim", C.c_void_p), ("t", C.c_void_p), ("occluded", C.c_void_p),
                ("bounce_prim", C.c_void_p), ("bounce_t", C.c_void_p)]


class RequestInfo(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_int), ("rays", C.c_int64),
                ("submit_ms", C.c_double), ("admit_ms", C.c_double), ("done_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("rejected", C.c_uint64),
                ("rays", C.c_int64), ("busy_ms", C.c_double), ("active", C.c_int),
                ("pending", C.c_int), ("hbm_in_use", C.c_size_t), ("host_in_use", C.c_size_t)]


# Every symbol include/mh.h declares, with its ctypes signature.
_P = C.c_void_p
_SIGS = {
    "mh_init": (C.c_int, [C.c_int]),
    "mh_last_error": (C.c_char_p, []),
    "mh_abi_version": (C.c_int, []),
    "mh_device_info": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_size_t)]),
    "mh_scene_create": (C.c_int, [_P, C.c_uint32, _P, C.c_uint32, C.POINTER(C.c_void_p)]),
    "mh_scene_destroy": (C.c_int, [_P]),
    "mh_scene_build": (C.c_int, [_P, C.POINTER(BuildOpts)]),
    "mh_scene_set_attributes": (C.c_int, [_P, _P, C.c_uint32, _P, _P, C.c_uint32, _P]),
    "mh_ply_read": (C.c_int, [C.c_char_p, _P, _P, C.c_uint32, _P, C.c_uint32,
                              C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "mh_obj_read": (C.c_int, [C.c_char_p, C.POINTER(C.POINTER(ObjMesh))]),
    "mh_obj_free": (None, [C.POINTER(ObjMesh)]),
    "mh_shade": (C.c_int, [_P, C.POINTER(Rays), C.POINTER(Hits), C.c_int64,
                           C.POINTER(Shading), _P]),
    "mh_dynbvh_build_host": (C.c_int, [_P, C.c_uint32, _P, C.c_uint32, C.POINTER(BuildOpts),
                                       _P, _P, C.POINTER(C.c_uint32)]),
    "mh_scene_load_dynbvh": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_uint32, C.c_int]),
    "mh_scene_load_dynbvh_file": (C.c_int, [_P, C.c_char_p, C.c_int]),
    "mh_scene_save_dynbvh_file": (C.c_int, [_P, C.c_char_p]),
    "mh_scene_export_dynbvh": (C.c_int, [_P, _P, C.POINTER(C.c_uint32), _P, C.POINTER(C.c_uint32)]),
    "mh_scene_bounds": (C.c_int, [_P, _P, _P]),
    "mh_scene_stats": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_uint32), C.POINTER(C.c_size_t)]),
    "mh_scene_set_vertices": (C.c_int, [_P, _P, C.c_int, _P]),
    "mh_scene_refit": (C.c_int, [_P, _P]),
    "mh_scene_update": (C.c_int, [_P, _P]),
    "mh_scene_set_workers": (C.c_int, [_P, C.c_int]),
    "mh_scene_vertices_device": (C.c_int, [_P, C.POINTER(C.c_void_p)]),
    "mh_scene_displace_sine": (C.c_int, [_P, C.c_float, C.c_float, C.c_float, _P]),
    "mh_trace": (C.c_int, [_P, C.POINTER(Rays), C.POINTER(Hits), C.c_int64, C.c_uint32,
                           C.POINTER(TraceOpts), _P]),
    "mh_trace_host": (C.c_int, [_P, C.POINTER(Rays), C.POINTER(Hits), C.c_int64, C.c_uint32,
                                C.POINTER(TraceOpts)]),
    "mh_trace_noalloc": (C.c_int, [_P, C.POINTER(Rays), C.POINTER(Hits), C.c_int64, C.c_uint32,
                                   C.POINTER(TraceOpts), _P]),
    "mh_frame_create": (C.c_int, [_P, C.POINTER(Batch), C.c_int, C.POINTER(C.c_void_p)]),
    "mh_frame_set_batches": (C.c_int, [_P, C.POINTER(Batch), C.c_int, _P]),
    "mh_frame_set_config": (C.c_int, [_P, C.POINTER(LaunchConfig)]),
    "mh_frame_get_config": (C.c_int, [_P, C.POINTER(LaunchConfig)]),
    "mh_frame_launch": (C.c_int, [_P, _P]),
    "mh_frame_capture": (C.c_int, [_P, _P]),
    "mh_frame_replay": (C.c_int, [_P, _P]),
    "mh_frame_time": (C.c_int, [_P, C.c_int, C.c_int, _P, C.POINTER(C.c_float)]),
    "mh_frame_autotune": (C.c_int, [_P, C.c_int, C.c_char_p, C.c_int, C.POINTER(LaunchConfig),
                                    C.POINTER(C.c_float)]),
    "mh_frame_destroy": (C.c_int, [_P]),
    "mh_gen_camera_rays": (C.c_int, [C.POINTER(Camera), C.POINTER(RayBuffers), _P]),
    "mh_gen_shadow_rays": (C.c_int, [_P, C.POINTER(Rays), C.POINTER(Hits), C.c_int64,
                                     C.POINTER(C.c_float * 3), C.POINTER(RayBuffers), _P]),
    "mh_gen_bounce_rays": (C.c_int, [_P, C.POINTER(Rays), C.POINTER(Hits), C.c_int64, C.c_uint64,
                                     C.POINTER(RayBuffers), _P]),
    "mh_server_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.POINTER(ServerOpts),
                                   C.POINTER(C.c_void_p)]),
    "mh_server_submit": (C.c_int, [_P, C.POINTER(Request), C.POINTER(C.c_uint64)]),
    "mh_server_wait": (C.c_int, [_P, C.c_uint64, C.POINTER(RequestResult), C.POINTER(RequestInfo)]),
    "mh_server_cancel": (C.c_int, [_P, C.c_uint64]),
    "mh_server_get_stats": (C.c_int, [_P, C.POINTER(ServerStats)]),
    "mh_server_clock_ms": (C.c_double, [_P]),
    "mh_server_destroy": (C.c_int, [_P]),
    "mh_malloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
    "mh_free": (C.c_int, [_P]),
    "mh_memcpy_h2d": (C.c_int, [_P, _P, C.c_size_t]),
    "mh_memcpy_d2h": (C.c_int, [_P, _P, C.c_size_t]),
    "mh_device_copy": (C.c_int, [_P, _P, C.c_size_t, _P]),
    "mh_stream_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "mh_stream_destroy": (C.c_int, [_P]),
    "mh_stream_sync": (C.c_int, [_P]),
    "mh_time_trace": (C.c_int, [_P, C.POINTER(Rays), C.POINTER(Hits), C.c_int64, C.c_uint32,
                                C.POINTER(TraceOpts), _P, C.c_int, C.POINTER(C.c_float)]),
}

_lib = None


def load() -> C.CDLL:
    """Loads libmh.so once and attaches signatures; raises if it is not built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is not built; run `python -m rfmanta_amd.build`")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(code: int) -> None:
    if code != MH_OK:
        msg = load().mh_last_error()
        raise MhError(code, msg.decode() if msg else "")


def ptr(a: np.ndarray | None) -> int | None:
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data
