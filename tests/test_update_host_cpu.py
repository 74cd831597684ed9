"""The rotating update's device pass on the host: rfmanta_amd/csrc/rotate.hip's own leaf pass,
leaf splits, rotations and PairNode stores (tests/hostrun/update_host.cpp, one lane at a
time, built with AddressSanitizer over buffers of exact size) give the oracle's updated tree
(DynBVH::update, DynBVH.cc:2068-2123 and rotateNode, DynBVH.cc:3051-3372) and stay inside
their buffers, on trees whose leaves hold up to 130 objects (so that leaves cross 64-lane
wave ends and the splits build deep subtrees) and on a built tree."""

import os
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from rfmanta_amd import scenes
from test_gpu_update import _tree_of_leaf_sizes, canonical_equal

REPO = Path(__file__).resolve().parents[1]
HOST = REPO / "tests" / "hostrun"


@pytest.fixture(scope="module")
def update_host(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no C++ compiler")
    exe = tmp_path_factory.mktemp("hostrun") / "update_host"
    subprocess.run([cxx, "-std=c++17", "-g", "-O1", "-fsanitize=address", "-fno-omit-frame-pointer",
                    "-ffp-contract=off", f"-I{HOST}", f"-I{REPO / 'rfmanta_amd' / 'csrc'}",
                    "-include", str(HOST / "shim.h"), str(HOST / "update_host.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    return exe


def _run(exe, d, oracle, nodes, ids, nn, vd, idx):
    """One update from (nodes, ids) with the new vertices vd: the host run's tree and the
    oracle's, from the same records, costs and subtree sizes."""
    st = oracle.RotState(nodes, ids, nn)
    d.mkdir()
    st.nodes.tofile(d / "nodes.bin")
    st.ids.tofile(d / "ids.bin")
    np.ascontiguousarray(vd, np.float32).tofile(d / "verts.bin")
    np.ascontiguousarray(idx, np.uint32).tofile(d / "idx.bin")
    st.costs.tofile(d / "costs.bin")
    st.ssize.tofile(d / "ssize.bin")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1")
    r = subprocess.run([str(exe), str(d), str(nn)], capture_output=True, text=True, env=env)
    assert r.returncode == 0, r.stderr[-3000:]
    st.update(vd, idx)
    gn = np.fromfile(d / "nodes_out.bin", oracle.NODE_DTYPE)
    gi = np.fromfile(d / "ids_out.bin", np.int32)
    ok, msg, n = canonical_equal(gn, gi, st.nodes, st.ids)
    assert ok, msg
    return r.stdout, n, st


def test_update_large_leaves_host(update_host, oracle, tmp_path):
    rng = np.random.default_rng(7)
    sizes = [int(x) for x in rng.integers(1, 9, 300)] + [64, 65, 130, 1, 63, 127, 2, 128]
    rng.shuffle(sizes)
    T = sum(sizes)
    v, i = scenes.soup(T, seed=17)
    ids = rng.permutation(T).astype(np.int32)
    nodes, nn = _tree_of_leaf_sizes(oracle, v, i, sizes, ids)
    vd = v.copy()
    vd[:, 1] += (0.2 * np.sin(3 * v[:, 0] + 1.8)).astype(np.float32)
    out, n, st = _run(update_host, tmp_path / "a", oracle, nodes, ids, nn, vd, i)
    assert st.next_free.value > nn  # the large leaves were split
    assert out.split()[1] == str(n) and int(out.split()[3]) == (n + 1) // 2


def test_update_built_tree_host(update_host, oracle, tmp_path):
    v, i = scenes.soup(3000, seed=7, sigma=0.05)
    nodes, ids, nn = oracle.build_bvh(v, i)
    vd = (v + np.random.default_rng(3).normal(0, 0.3, v.shape)).astype(np.float32)
    _run(update_host, tmp_path / "b", oracle, nodes, ids, nn, vd, i)


def test_update_heightfield_host(update_host, oracle, tmp_path):
    v, i = scenes.heightfield(120, 60, seed=4)
    nodes, ids, nn = oracle.build_bvh(v, i)
    vd = v.copy()
    vd[:, 1] += (0.05 * np.sin(4 * v[:, 0] + 0.3)).astype(np.float32)
    _run(update_host, tmp_path / "c", oracle, nodes, ids, nn, vd, i)
