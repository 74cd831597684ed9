"""Headline benchmark: closest-hit Mrays/s on the 1M-triangle heightfield (one MI355X).

Workload (one step = one frame of closest-hit work on resident inputs):
  scene   1000 x 500 vertex-grid heightfield, 997,002 triangles, DynBVH SAH build
  rays    1920 x 1080 pinhole primary rays (coherent, 8x8 tiles) and one cosine-weighted
          diffuse bounce ray per primary hit (incoherent, PCG32 seed 1); one shadow ray
          per hit toward a point light is traced and reported beside it (any-hit)
Timed: K steps of [primary closest-hit, bounce closest-hit] on one stream, inputs already
in HBM; kernel durations come from HIP events recorded on that stream.

Prints one JSON line (see the keys below). --gpus N > 1 exits: this build runs on one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))

METRIC = "Mrays/s closest-hit at 1M tris (coherent/incoherent); p50/p99 request latency"
HBM_PEAK_GBS = 8000.0
RAY_IN_BYTES = 32    # o.xyz, d.xyz, tmax, (masked byte rounded into the record)
HIT_OUT_CLOSEST = 16  # prim, t, u, v
HIT_OUT_ANY = 4
BOX_BYTES = 32        # one reference BVHNode per box test
TRI_BYTES = 48        # 3 x u32 index + 3 x 12-byte vertex per triangle test


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cores() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--stack", type=int, default=0)
    ap.add_argument("--bpc", type=int, default=0, help="blocks per CU for single-ray kernels")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="also run a device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard rays or requests across GPUs with a replicated BVH).",
              file=sys.stderr)
        return 2

    from rfmanta_amd import api, scenes

    info = api.device_info()
    log(f"device {info}")
    t0 = time.time()
    verts, idx = scenes.config1_scene()
    log(f"heightfield {len(idx)} triangles generated in {time.time() - t0:.1f}s")
    scene = api.Scene(verts, idx)
    t0 = time.time()
    scene.build()
    build_s = time.time() - t0
    st = scene.stats()
    log(f"SAH build {build_s:.2f}s, {st}")

    prim_rays = scenes.config1_camera()
    n_prim = len(prim_rays["ox"])
    d_prim = api.DeviceRays.upload(prim_rays)
    h_prim = api.DeviceHits.alloc(n_prim)
    scene.trace(d_prim, h_prim)
    api.sync()
    hp = h_prim.download()
    n_hit = int((hp["prim"] >= 0).sum())
    log(f"primary: {n_prim} rays, {n_hit} hits")
    bounce = scenes.bounce_rays(verts, idx, prim_rays, hp, seed=1)
    shadow = scenes.shadow_rays(verts, idx, prim_rays, hp, scenes.CONFIG1_LIGHT, compact=True)
    n_b, n_s = len(bounce["ox"]), len(shadow["ox"])
    d_b = api.DeviceRays.upload(bounce)
    h_b = api.DeviceHits.alloc(n_b)
    d_s = api.DeviceRays.upload(shadow)
    h_s = api.DeviceHits.alloc(n_s, any_hit=True)

    stream = None
    opts = api.trace_opts(1, args.stack, args.bpc)

    # Per-ray box and triangle tests of the single-ray traversal (untimed counting pass).
    def counts(dr, dh, any_hit):
        c = api.DeviceArray.from_numpy(np.zeros(2, np.int64))
        scene.trace(dr, dh, any_hit=any_hit, opts=api.trace_opts(1, counters=c))
        api.sync()
        v = c.download(np.int64, 2)
        return float(v[0]) / dr.n, float(v[1]) / dr.n

    nb_p, nt_p = counts(d_prim, h_prim, False)
    nb_b, nt_b = counts(d_b, h_b, False)
    nb_s, nt_s = counts(d_s, h_s, True)
    log(f"tests/ray: primary box {nb_p:.1f} tri {nt_p:.2f}; bounce box {nb_b:.1f} tri {nt_b:.2f}; "
        f"shadow box {nb_s:.1f} tri {nt_s:.2f}")

    def step():
        scene.trace(d_prim, h_prim, opts=opts, stream=stream)
        scene.trace(d_b, h_b, opts=opts, stream=stream)

    for _ in range(args.warmup):
        step()
        scene.trace(d_s, h_s, any_hit=True, opts=opts, stream=stream)
    api.sync(stream)

    # Whole-step wall clock (both launches back to back, synchronised on both sides).
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    api.sync(stream)
    wall = time.perf_counter() - t0
    ms_step = 1e3 * wall / args.steps

    # Kernel durations from HIP events on the launch stream.
    import ctypes as C
    from rfmanta_amd import _lib as L

    def kernel_ms(dr, dh, any_hit):
        ms = C.c_float()
        r, h = dr.struct(), dh.struct()
        L.check(L.load().mh_time_trace(scene.h, C.byref(r), C.byref(h), dr.n,
                                       L.MH_ANY_HIT if any_hit else 0, C.byref(opts), stream,
                                       args.steps, C.byref(ms)))
        return ms.value / args.steps

    k_p = kernel_ms(d_prim, h_prim, False)
    k_b = kernel_ms(d_b, h_b, False)
    k_s = kernel_ms(d_s, h_s, True)
    closest_rays = n_prim + n_b
    value = closest_rays / (ms_step * 1e-3) / 1e6
    mr_p = n_prim / (k_p * 1e-3) / 1e6
    mr_b = n_b / (k_b * 1e-3) / 1e6
    mr_s = n_s / (k_s * 1e-3) / 1e6
    log(f"step {ms_step:.3f} ms; kernels primary {k_p:.3f} ms ({mr_p:.0f} Mrays/s), bounce "
        f"{k_b:.3f} ms ({mr_b:.0f}), shadow {k_s:.3f} ms ({mr_s:.0f})")

    # Roofline of the dominant kernel (incoherent closest-hit bounce rays).
    bytes_ray_b = RAY_IN_BYTES + HIT_OUT_CLOSEST + BOX_BYTES * nb_b + TRI_BYTES * nt_b
    achieved_gbs = bytes_ray_b * n_b / (k_b * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved_gbs, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_trace_ray<closest,KS> on bounce rays",
                "bytes_per_ray": round(bytes_ray_b, 1), "rays_per_launch": n_b,
                "launch_ms": round(k_b, 4)}

    if args.copy_probe_mb > 0:
        # A copy kernel of known size in the same process: the reference for scaling
        # FETCH_SIZE / WRITE_SIZE counter readings to bytes.
        nbytes = args.copy_probe_mb << 20
        a, b = api.DeviceArray(nbytes), api.DeviceArray(nbytes)
        for _ in range(3):
            L.check(L.load().mh_device_copy(b.ptr, a.ptr, nbytes, stream))
        api.sync(stream)
        log(f"copy probe: 3 x {nbytes} bytes")
        a.free()
        b.free()
    traffic = load_traffic()
    if traffic is not None:
        roofline["traffic"] = traffic["bytes_per_launch"]
        roofline["traffic_source"] = traffic["source"]

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(verts, idx, scene, prim_rays, bounce)

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "heightfield 1000x500 (997,002 tris), 1920x1080 primary + "
                               "1 diffuse bounce per hit, closest-hit",
                   "primary_rays": n_prim, "bounce_rays": n_b, "shadow_rays": n_s,
                   "traversal": "single-ray", "parallelism": "1 GPU"},
        "coherent_mrays": round(mr_p, 1), "incoherent_mrays": round(mr_b, 1),
        "shadow_anyhit_mrays": round(mr_s, 1), "build_s": round(build_s, 2),
        "tests_per_ray": {"primary": [round(nb_p, 2), round(nt_p, 3)],
                          "bounce": [round(nb_b, 2), round(nt_b, 3)],
                          "shadow": [round(nb_s, 2), round(nt_s, 3)]},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    scene.close()
    return 0


def load_traffic():
    """HBM bytes per launch of the dominant kernel, from the committed counter summary."""
    p = REPO / "profiles" / "traffic_trace_ray_bounce_heightfield1M.json"
    if not p.exists():
        return None
    with open(p) as f:
        return json.load(f)


def cpu_baseline(verts, idx, scene, prim_rays, bounce) -> dict:
    """The oracle's packet traversal (8x8 packets, one thread per core) on a sample."""
    sys.path.insert(0, str(REPO / "oracle"))
    import pyoracle

    ids, nodes = scene.export_dynbvh()
    cores = host_cores()
    total_rays = 0
    total_s = 0.0
    reps = {}
    for name, rays, budget in (("primary", prim_rays, 5.0), ("bounce", bounce, 5.0)):
        sub = {k: rays[k] for k in ("ox", "oy", "oz", "dx", "dy", "dz", "tmax")}
        n = len(sub["ox"])
        done, spent = 0, 0.0
        while spent < budget or done < 2:
            t0 = time.perf_counter()
            pyoracle.trace(verts, idx, nodes, ids, sub, packet_size=64, threads=cores)
            spent += time.perf_counter() - t0
            done += 1
        reps[name] = done
        total_rays += n * done
        total_s += spent
        log(f"cpu baseline {name}: {done} x {n} rays in {spent:.2f}s on {cores} threads")
    return {"value": round(total_rays / total_s / 1e6, 3), "unit": "Mrays/s", "cores": cores,
            "kind": "port",
            "sample": f"the full frame's {len(prim_rays['ox'])} primary rays x{reps['primary']} "
                      f"and {len(bounce['ox'])} bounce rays x{reps['bounce']}, 64-ray packets "
                      f"in 8x8-tile order, closest-hit, ~10 s of CPU time"}


if __name__ == "__main__":
    sys.exit(main())
