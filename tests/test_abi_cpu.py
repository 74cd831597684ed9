"""C ABI checks that need no GPU: exports, the host SAH build, error paths."""

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from rfmanta_amd import _lib as L
from rfmanta_amd import api, scenes

HEADER = Path(__file__).resolve().parent.parent / "include" / "mh.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mh_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(mh):
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(mh, n), f"libmh.so does not export {n}"
        assert n in L._SIGS, f"no ctypes signature for {n}"


def test_abi_version(mh):
    assert mh.mh_abi_version() == 2


def test_node_record_layout():
    # BVHNode (Model/Groups/DynBVH.h:37-77): bounds @0, child @24, axis @28, bits @29,
    # children @30, 32 bytes.
    d = L.NODE_DTYPE
    assert d.itemsize == 32
    assert d.fields["child"][1] == 24 and d.fields["axis"][1] == 28
    assert d.fields["bits"][1] == 29 and d.fields["children"][1] == 30


MESHES = {
    "soup10k": lambda: scenes.config0_scene(),
    "heightfield": lambda: scenes.heightfield(120, 60, seed=4),
    "soup_wide": lambda: scenes.soup(3000, seed=2, sigma=0.3),
    "duplicates": lambda: (np.tile(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), (30, 1)),
                           np.arange(90, dtype=np.uint32).reshape(30, 3)),
    "flat": lambda: (lambda v, i: (v * np.array([1, 0, 1], np.float32), i))(*scenes.soup(500, seed=6)),
    "tiny3": lambda: scenes.soup(3, seed=3, sigma=0.4),
}


@pytest.mark.parametrize("name", sorted(MESHES))
@pytest.mark.parametrize("threads", [1, 8])
def test_host_build_matches_oracle(oracle, name, threads):
    v, i = MESHES[name]()
    on, oi, ou = oracle.build_bvh(v, i)
    pn, pi, pu = api.build_dynbvh_host(v, i, threads=threads)
    assert ou == pu
    assert np.array_equal(oi, pi)
    assert on.tobytes() == pn.tobytes()


def test_host_build_without_rtsah(oracle):
    v, i = scenes.soup(1000, seed=12)
    on, oi, _ = oracle.build_bvh(v, i, rtsah=False)
    pn, pi, _ = api.build_dynbvh_host(v, i, rtsah=False)
    assert on.tobytes() == pn.tobytes()
    assert not np.any(pn["bits"] & 1)


def test_host_build_errors(mh):
    v, i = scenes.soup(4, seed=1)
    ids = np.empty(4, np.int32)
    nodes = np.empty(8, L.NODE_DTYPE)
    used = C.c_uint32()
    o = L.BuildOpts(0, 1, 0)
    rc = mh.mh_dynbvh_build_host(v.ctypes.data, len(v), i.ctypes.data, 0, C.byref(o),
                                 ids.ctypes.data, nodes.ctypes.data, C.byref(used))
    assert rc == -1 and b"empty" in mh.mh_last_error()
    bad = i.copy()
    bad[2, 1] = 999
    rc = mh.mh_dynbvh_build_host(v.ctypes.data, len(v), bad.ctypes.data, 4, C.byref(o),
                                 ids.ctypes.data, nodes.ctypes.data, C.byref(used))
    assert rc == -1 and b"out of range" in mh.mh_last_error()
    with pytest.raises(L.MhError):
        L.check(rc)


def test_null_arguments_are_rejected(mh):
    assert mh.mh_trace(None, None, None, 10, 0, None, None) == -1
    assert mh.mh_scene_refit(None, None) == -1
    assert mh.mh_scene_destroy(None) == 0
    assert mh.mh_scene_create(None, 0, None, 0, None) == -1


def test_file_format_round_trip(tmp_path, oracle):
    """DynBVH::saveToFile layout: u32 n_ids, i32 ids[], u32 n_nodes, 32-byte nodes."""
    v, i = scenes.soup(300, seed=4)
    nodes, ids, _ = oracle.build_bvh(v, i)
    p = tmp_path / "tree.bvh"
    with open(p, "wb") as f:
        f.write(np.uint32(len(ids)).tobytes())
        f.write(ids.tobytes())
        f.write(np.uint32(len(nodes)).tobytes())
        f.write(nodes.tobytes())
    raw = p.read_bytes()
    assert len(raw) == 8 + 4 * len(ids) + 32 * len(nodes)
    n_ids = int(np.frombuffer(raw[:4], np.uint32)[0])
    back = np.frombuffer(raw[8 + 4 * n_ids:], L.NODE_DTYPE)
    assert back.tobytes() == nodes.tobytes()


def test_triangle_limit_is_rejected_before_any_work(mh):
    """2^26 triangles and more are refused (leaf codes and the pair table's descriptor reach)
    before the index array is read or a device is touched."""
    v = np.zeros(3, np.float32)
    i = np.zeros(3, np.uint32)
    out = C.c_void_p()
    assert mh.mh_scene_create(v.ctypes.data, 1, i.ctypes.data, 1 << 26, C.byref(out)) == L.MH_E_UNSUPPORTED
    assert out.value is None
    ids = np.zeros(1, np.int32)
    nodes = np.zeros(64, np.uint8)
    n = C.c_uint32()
    assert mh.mh_dynbvh_build_host(v.ctypes.data, 1, i.ctypes.data, 1 << 26, None, ids.ctypes.data,
                                   nodes.ctypes.data, C.byref(n)) == L.MH_E_UNSUPPORTED
    assert b"exceeds the limit" in mh.mh_last_error()
