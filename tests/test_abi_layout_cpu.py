"""The ctypes mirror of the C ABI (rfmanta_amd/_lib.py) against include/mh.h: every struct
the bindings pass by pointer has the size and field offsets the C compiler gives it."""

import ctypes as C
import shutil
import subprocess
from pathlib import Path

import pytest

from rfmanta_amd import _lib as L

REPO = Path(__file__).resolve().parents[1]

# ctypes class -> C type name; fields are compared by name.
PAIRS = [
    (L.Rays, "mh_rays"),
    (L.Hits, "mh_hits"),
    (L.TraceOpts, "mh_trace_opts"),
    (L.BuildOpts, "mh_build_opts"),
    (L.ObjMesh, "mh_obj_mesh"),
    (L.Shading, "mh_shading"),
    (L.Batch, "mh_batch"),
    (L.LaunchConfig, "mh_launch_config"),
    (L.RayBuffers, "mh_ray_buffers"),
    (L.Camera, "mh_camera"),
    (L.ServerOpts, "mh_server_opts"),
    (L.Request, "mh_request"),
    (L.RequestResult, "mh_request_result"),
    (L.RequestInfo, "mh_request_info"),
    (L.ServerStats, "mh_server_stats"),
]


def test_struct_layouts_match_header(tmp_path):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "mh.h"', "int main(void) {"]
    for cls, name in PAIRS:
        lines.append(f'  printf("{name} size %zu\\n", sizeof({name}));')
        for field, _ in cls._fields_:
            lines.append(f'  printf("{name} {field} %zu\\n", offsetof({name}, {field}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", "-I", str(REPO / "include"), str(src), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for ln in out.splitlines():
        name, key, val = ln.split()
        got[(name, key)] = int(val)
    for cls, name in PAIRS:
        assert got[(name, "size")] == C.sizeof(cls), name
        for field, _ in cls._fields_:
            assert got[(name, field)] == getattr(cls, field).offset, (name, field)
