"""ctypes front end of the oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this module, as
the checker. It loads oracle/liboracle.so, built by oracle/Makefile.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

NODE_DTYPE = np.dtype([("lo", "<f4", (3,)), ("hi", "<f4", (3,)), ("child", "<i4"),
                       ("axis", "u1"), ("bits", "u1"), ("children", "<i2")])

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        lib = C.CDLL(str(LIB_PATH))
        P = C.c_void_p
        lib.oracle_build.restype = C.c_int
        lib.oracle_build.argtypes = [P, C.c_int, P, C.c_int, C.c_int, P, P]
        lib.oracle_build_approx.restype = C.c_int
        lib.oracle_build_approx.argtypes = [P, C.c_int, P, C.c_int, C.c_int, P, P]
        lib.oracle_shade.restype = None
        lib.oracle_shade.argtypes = [P, C.c_int, P, C.c_int, C.c_int, P, P, P, P, P, P, P, P,
                                     P, C.c_int64, P]
        lib.oracle_partition_objects.restype = C.c_int
        lib.oracle_partition_objects.argtypes = [P, C.c_int, C.c_float, P]
        lib.oracle_rtsah.restype = None
        lib.oracle_rtsah.argtypes = [P, C.c_int]
        lib.oracle_rtsah_costs.restype = None
        lib.oracle_rtsah_costs.argtypes = [P, C.c_int, P]
        lib.oracle_trace.restype = C.c_int
        lib.oracle_trace.argtypes = [P, C.c_int, P, C.c_int, P, P, C.c_int] + [P] * 8 + [
            C.c_int64, C.c_int, C.c_int, C.c_int, P, P, P, P, P, P]
        lib.oracle_trace_mode.restype = C.c_int
        lib.oracle_trace_mode.argtypes = lib.oracle_trace.argtypes + [C.c_int]
        lib.oracle_bruteforce.restype = C.c_int
        lib.oracle_bruteforce.argtypes = [P, C.c_int, P, C.c_int, C.c_int] + [P] * 8 + [
            C.c_int64, C.c_int, C.c_int, P, P, P, P, P]
        lib.oracle_refit.restype = None
        lib.oracle_refit.argtypes = [P, C.c_int, P, C.c_int, P, P]
        lib.oracle_rot_init.restype = None
        lib.oracle_rot_init.argtypes = [P, C.c_int, P, P]
        lib.oracle_update.restype = None
        lib.oracle_update.argtypes = [P, C.c_int, P, C.c_int, P, P, P, P, P]
        lib.oracle_update_workers.restype = None
        lib.oracle_update_workers.argtypes = [P, C.c_int, P, C.c_int, P, P, P, P, P, C.c_int]
        lib.oracle_mark_large.restype = None
        lib.oracle_mark_large.argtypes = [P, C.c_int, C.c_int]
        lib.oracle_wald_setup.restype = None
        lib.oracle_wald_setup.argtypes = [P, P, C.c_int, P]
        lib.oracle_node_size.restype = C.c_int
        lib.oracle_ttp_probs.restype = None
        lib.oracle_ttp_probs.argtypes = [P, P, P, P]
        _lib = lib
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _mesh(verts, idx):
    v = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
    i = np.ascontiguousarray(idx, np.uint32).reshape(-1, 3)
    return v, i


def build_bvh(verts, idx, rtsah: bool = True):
    """DynBVH::rebuild(0,1) + RTSAH. Returns (nodes[2T] NODE_DTYPE, ids int32[T], num_nodes)."""
    v, i = _mesh(verts, idx)
    T = len(i)
    nodes = np.zeros(2 * T, NODE_DTYPE)
    ids = np.zeros(T, np.int32)
    used = load().oracle_build(_p(v), len(v), _p(i), T, int(rtsah), _p(nodes), _p(ids))
    if used < 0:
        raise ValueError("oracle build failed")
    return nodes, ids, used


def build_bvh_approx(verts, idx, rtsah: bool = True):
    """DynBVH::rebuild(0,1) built with the approximate (binned) SAH at every node, + RTSAH.
    Returns (nodes[2T] NODE_DTYPE, ids int32[T], num_nodes)."""
    v, i = _mesh(verts, idx)
    T = len(i)
    nodes = np.zeros(2 * T, NODE_DTYPE)
    ids = np.zeros(T, np.int32)
    used = load().oracle_build_approx(_p(v), len(v), _p(i), T, int(rtsah), _p(nodes), _p(ids))
    if used < 0:
        raise ValueError("oracle build failed")
    return nodes, ids, used


def shade(verts, idx, rays, hits, prim_kind=0, normals=None, normal_idx=None, tex=None,
          tex_idx=None):
    """Hit position, shading normal, geometric normal and texture coordinates per ray
    (12 floats; zeros for misses)."""
    v, i = _mesh(verts, idx)
    arr = [np.ascontiguousarray(rays[k], np.float32) for k in ("ox", "oy", "oz", "dx", "dy", "dz")]
    ptrs = (C.c_void_p * 6)(*[a.ctypes.data for a in arr])
    prim = np.ascontiguousarray(hits["prim"], np.int32)
    t = np.ascontiguousarray(hits["t"], np.float32)
    u = np.ascontiguousarray(hits["u"], np.float32)
    w = np.ascontiguousarray(hits["v"], np.float32)
    keep = [np.ascontiguousarray(x, np.float32 if k % 2 == 0 else np.uint32) if x is not None else None
            for k, x in enumerate((normals, normal_idx, tex, tex_idx))]
    ps = [None if x is None else _p(x) for x in keep]
    out = np.zeros((len(prim), 12), np.float32)
    load().oracle_shade(_p(v), len(v), _p(i), len(i), int(prim_kind), ps[0], ps[1], ps[2], ps[3],
                        C.cast(ptrs, C.c_void_p), _p(prim), _p(t), _p(u), _p(w), len(prim), _p(out))
    return out


def partition_objects(keys, position):
    """partitionObjects on objects 0..n-1 with centroids `keys`: (order after, split)."""
    k = np.ascontiguousarray(keys, np.float32)
    perm = np.zeros(len(k), np.int32)
    split = load().oracle_partition_objects(_p(k), len(k), float(np.float32(position)), _p(perm))
    return perm, split


def rtsah(nodes):
    load().oracle_rtsah(_p(nodes), len(nodes))


def rtsah_costs(nodes):
    """DynBVH::computeTraversalCost in place on nodes; returns float32[n, 4] with
    (c0, c1, lCost, rCost) per record (leaves: c0, c1, 0, 0)."""
    out = np.zeros((len(nodes), 4), np.float32)
    load().oracle_rtsah_costs(_p(nodes), len(nodes), _p(out))
    return out


def _rays(rays):
    arr = {k: np.ascontiguousarray(rays[k], np.float32) for k in
           ("ox", "oy", "oz", "dx", "dy", "dz", "tmax")}
    m = rays.get("masked")
    m = None if m is None else np.ascontiguousarray(m, np.uint8)
    return arr, m


def trace(verts, idx, nodes, ids, rays, any_hit=False, packet_size=1, prim_kind=0,
          threads=None, counters=False, sse=False):
    """Packet traversal over consecutive groups of packet_size rays. sse=True runs the
    reference's SSE build (rcpps reciprocals of this CPU; Kensler-Shirley only)."""
    v, i = _mesh(verts, idx)
    arr, m = _rays(rays)
    n = len(arr["ox"])
    out = {"prim": np.empty(n, np.int32), "t": np.empty(n, np.float32),
           "u": np.empty(n, np.float32), "v": np.empty(n, np.float32),
           "occluded": np.empty(n, np.uint8)}
    cnt = np.zeros(2, np.int64) if counters else None
    th = threads if threads is not None else (os.cpu_count() or 1)
    rc = load().oracle_trace_mode(_p(v), len(v), _p(i), len(i), _p(np.ascontiguousarray(nodes)),
                                  _p(np.ascontiguousarray(ids, np.int32)), prim_kind,
                                  *[_p(arr[k]) for k in ("ox", "oy", "oz", "dx", "dy", "dz",
                                                         "tmax")],
                                  _p(m), n, packet_size, int(any_hit), th, _p(out["prim"]),
                                  _p(out["t"]), _p(out["u"]), _p(out["v"]), _p(out["occluded"]),
                                  _p(cnt), int(bool(sse)))
    if rc != 0:
        raise ValueError("oracle trace failed")
    if counters:
        out["box_tests"] = int(cnt[0])
        out["tri_tests"] = int(cnt[1])
    return out


def bruteforce(verts, idx, rays, any_hit=False, prim_kind=0, threads=None):
    v, i = _mesh(verts, idx)
    arr, m = _rays(rays)
    n = len(arr["ox"])
    out = {"prim": np.empty(n, np.int32), "t": np.empty(n, np.float32),
           "u": np.empty(n, np.float32), "v": np.empty(n, np.float32),
           "occluded": np.empty(n, np.uint8)}
    th = threads if threads is not None else (os.cpu_count() or 1)
    load().oracle_bruteforce(_p(v), len(v), _p(i), len(i), prim_kind,
                             *[_p(arr[k]) for k in ("ox", "oy", "oz", "dx", "dy", "dz", "tmax")],
                             _p(m), n, int(any_hit), th, _p(out["prim"]), _p(out["t"]),
                             _p(out["u"]), _p(out["v"]), _p(out["occluded"]))
    return out


def refit(verts, idx, nodes, ids):
    """updateBounds<false> without rotations, in place on nodes."""
    v, i = _mesh(verts, idx)
    assert nodes.flags["C_CONTIGUOUS"]
    load().oracle_refit(_p(v), len(v), _p(i), len(i), _p(nodes),
                        _p(np.ascontiguousarray(ids, np.int32)))


class RotState:
    """DynBVH state for updates with tree rotations: records (2T), object ids, costs,
    subtree sizes and the next free record pair (DynBVH.h:377-385)."""

    def __init__(self, nodes, ids, num_nodes):
        self.nodes = np.ascontiguousarray(nodes).copy()
        self.ids = np.ascontiguousarray(ids, np.int32).copy()
        n = len(self.nodes)
        self.costs = np.zeros(n, np.float32)
        self.ssize = np.zeros(n, np.uint32)
        self.next_free = C.c_int(int(num_nodes))
        load().oracle_rot_init(_p(self.nodes), n, _p(self.costs), _p(self.ssize))

    def update(self, verts, idx, workers=None):
        """One DynBVH::update with rotations and leaf splits, in place. workers=None runs
        the one-thread update; workers=N the N-worker schedule with its large-subtree
        threshold."""
        v, i = _mesh(verts, idx)
        if workers is None:
            load().oracle_update(_p(v), len(v), _p(i), len(i), _p(self.nodes), _p(self.ids),
                                 _p(self.costs), _p(self.ssize), C.byref(self.next_free))
        else:
            load().oracle_update_workers(_p(v), len(v), _p(i), len(i), _p(self.nodes),
                                         _p(self.ids), _p(self.costs), _p(self.ssize),
                                         C.byref(self.next_free), int(workers))


def mark_large(nodes, workers):
    """isLargeSubtree bits for `workers` render threads (in place)."""
    load().oracle_mark_large(_p(nodes), len(nodes), int(workers))


def ttp_probs(parent, left, right):
    """(lr, jl, jr) of TreeTraversalProb::computeProbs; boxes as (lo.xyz, hi.xyz)."""
    a = [np.ascontiguousarray(x, np.float32) for x in (parent, left, right)]
    out = np.empty(3, np.float32)
    load().oracle_ttp_probs(*[_p(x) for x in a], _p(out))
    return out


def wald_setup(verts, idx):
    v, i = _mesh(verts, idx)
    out = np.empty((len(i), 12), np.float32)
    load().oracle_wald_setup(_p(v), _p(i), len(i), _p(out))
    return out
