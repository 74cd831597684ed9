"""Headline benchmark: closest-hit Mrays/s at 1M triangles (coherent / incoherent) and
request latency p50/p99, on one MI355X.

Sections (all on synthetic inputs, generated at start-up and resident in HBM when timed):
  C1  heightfield 1000 x 500 vertices (997,002 triangles), DynBVH SAH build. 1920 x 1080
      pinhole primary rays (coherent, 8x8 tiles), one cosine-weighted diffuse bounce ray
      per primary hit (incoherent, PCG32 seed 1) and one shadow ray per hit (any-hit).
      One step = one frame of closest-hit work, primary + bounce rays, traced by the
      batched form: one persistent launch over both batches (pinned launch configuration,
      the autotuner's pick reported beside it).
      `value` is that frame's closest-hit rays per second; the plain path (one launch per
      ray set) is timed beside it.
  C2  64 batches of 1k-2M rays (log-uniform; 40% primary, 30% shadow, 30% bounce; seed 1)
      against the C1 tree: 64 plain launches vs one persistent launch vs its graph replay.
  C3  request loop over two resident scenes (the C1 heightfield and an 8M-triangle soup in
      [-4,4]^3): Poisson arrivals swept over QPS, each request a random orbit camera at
      256^2 / 512^2 / 1024x768 / 1920x1080 with 1-4 spp (seed 2), primary + shadow rays;
      p50/p99 latency from submit to results in host memory; the offered rate doubles
      from 25 QPS until p99 exceeds the 100 ms objective (capacity).
  C4  4000 x 1000 heightfield (7,990,002 triangles): per frame y += 0.01 sin(4x + t), GPU
      update (refit with tree rotations and leaf splits), then 1920x1080 primary + shadow +
      bounce rays through the request loop; 100 frames.

Prints one JSON line. --gpus N > 1 exits: this build runs on one GPU.
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import threading
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))
from rfmanta_amd import scenes as scenes_mod  # noqa: E402  (numpy only; no library load)

METRIC = "Mrays/s closest-hit at 1M tris (coherent/incoherent); p50/p99 request latency"
HBM_PEAK_GBS = 8000.0
RAY_IN_BYTES = 32     # o.xyz, d.xyz, tmin, tmax
HIT_OUT_CLOSEST = 16  # prim, t, u, v
HIT_OUT_ANY = 4
BOX_BYTES = 32        # one reference BVHNode per box test
TRI_BYTES = 48        # 3 x u32 index + 3 x 12-byte vertex per triangle test
C3_SLO_MS = 100.0     # p99 objective for the sustained-load figure
# fp32 operations of one box test and one triangle test as the reference's scalar code
# states them. Slab test (DynBVH.cc:666-680): per axis two subtractions, two products, a
# min and a max for the interval, a max and a min into [tmin, tmax]: 8 x 3, plus the final
# comparison. Kensler-Shirley (KenslerShirleyTriangle.cc:336-377, edges and normal
# precomputed): det 5, reciprocal 1, p0 - o 3, t 6, its range 3, the cross product 9, u 6
# and its sign 1, v 6 and its checks 3 (u + v, > 1, v < 0), the closing interval test 2.
BOX_OPS = 25
TRI_OPS = 45
# VALU peaks of MI355X (256 CUs x 4 SIMDs of 16 lanes, 2.4 GHz peak engine clock): one
# fp32 operation per lane per cycle (not packed, not fused), and one wave64 instruction
# per SIMD every four cycles.
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12
VALU_ISSUE_PEAK = 256 * 4 * 2.4e9 / 4
PROFILE_JSON = "profiles/counters_trace_batch_heightfield1M.json"
# Launch configuration of the timed C1 frame (rfmanta_amd.scenes.C1_LAUNCH, which the
# full-size GPU tests check against the oracle): pinned so that the counter profile of the
# kernel is taken on the very launch that is timed.
C1_CONFIG = scenes_mod.C1_LAUNCH
KERNEL_SOURCES = ("batch_kernel.h", "traverse.h", "trace_common.h", "mh_device.h",
                  "batch_ks_pair.hip")


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cores() -> int:
    """Host cores this process may run on (all of them: no cap)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, n)


def mrays(n, ms):
    return n / (ms * 1e-3) / 1e6


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--skip", default="",
                    help="comma list of sections to skip: singles (per-set frames), dropin, "
                         "build, c2, c3, c4")
    ap.add_argument("--c1-config", default=json.dumps(C1_CONFIG),
                    help="JSON launch configuration of the timed C1 frame ('' = the "
                         "autotuner's pick)")
    ap.add_argument("--server-launch", default="",
                    help="JSON launch configuration of the C3/C4 request loops ('' = the "
                         "library's default)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner run reported beside the C1 frame")
    ap.add_argument("--c4-frames", type=int, default=100)
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="also run a device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard rays or requests across GPUs with a replicated BVH).",
              file=sys.stderr)
        return 2
    skip = set(s for s in args.skip.split(",") if s)

    from rfmanta_amd import _lib as L
    from rfmanta_amd import api, scenes

    info = api.device_info()
    log(f"device {info}")

    # ---------------------------------------------------------------- C1 set-up
    t0 = time.time()
    verts, idx = scenes.config1_scene()
    scene = api.Scene(verts, idx)
    t1 = time.time()
    scene.build()
    build_s = time.time() - t1
    log(f"C1 heightfield {len(idx)} tris generated+built in {time.time() - t0:.1f}s "
        f"(build {build_s:.2f}s), {scene.stats()}")
    prim_rays = scenes.config1_camera()
    n_prim = len(prim_rays["ox"])
    d_prim = api.DeviceRays.upload(prim_rays)
    h_prim = api.DeviceHits.alloc(n_prim)
    scene.trace(d_prim, h_prim)
    api.sync()
    hp = h_prim.download()
    log(f"primary: {n_prim} rays, {int((hp['prim'] >= 0).sum())} hits")
    bounce = scenes.bounce_rays(verts, idx, prim_rays, hp, seed=1)
    shadow = scenes.shadow_rays(verts, idx, prim_rays, hp, scenes.CONFIG1_LIGHT, compact=True)
    n_b, n_s = len(bounce["ox"]), len(shadow["ox"])
    d_b = api.DeviceRays.upload(bounce)
    h_b = api.DeviceHits.alloc(n_b)
    d_s = api.DeviceRays.upload(shadow)
    h_s = api.DeviceHits.alloc(n_s, any_hit=True)

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
    log(f"tests/ray: primary box {nb_p:.1f} tri {nt_p:.2f}; bounce box {nb_b:.1f} tri "
        f"{nt_b:.2f}; shadow box {nb_s:.1f} tri {nt_s:.2f}")
    bytes_p = RAY_IN_BYTES + HIT_OUT_CLOSEST + BOX_BYTES * nb_p + TRI_BYTES * nt_p
    bytes_b = RAY_IN_BYTES + HIT_OUT_CLOSEST + BOX_BYTES * nb_b + TRI_BYTES * nt_b
    bytes_s = RAY_IN_BYTES + HIT_OUT_ANY + BOX_BYTES * nb_s + TRI_BYTES * nt_s

    # ---------------------------------------------------------------- plain path
    opts = api.trace_opts(1)

    def kernel_ms(dr, dh, any_hit, iters):
        ms = C.c_float()
        r, h = dr.struct(), dh.struct()
        L.check(L.load().mh_time_trace(scene.h, C.byref(r), C.byref(h), dr.n,
                                       L.MH_ANY_HIT if any_hit else 0, C.byref(opts), None,
                                       iters, C.byref(ms)))
        return ms.value / iters

    for _ in range(args.warmup):
        scene.trace(d_prim, h_prim, opts=opts)
        scene.trace(d_b, h_b, opts=opts)
        scene.trace(d_s, h_s, any_hit=True, opts=opts)
    api.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        scene.trace(d_prim, h_prim, opts=opts)
        scene.trace(d_b, h_b, opts=opts)
    api.sync()
    plain_step_ms = 1e3 * (time.perf_counter() - t0) / args.steps
    k_p = kernel_ms(d_prim, h_prim, False, args.steps)
    k_b = kernel_ms(d_b, h_b, False, args.steps)
    k_s = kernel_ms(d_s, h_s, True, args.steps)
    plain = {"step_ms": round(plain_step_ms, 4),
             "closest_mrays": round(mrays(n_prim + n_b, plain_step_ms), 1),
             "coherent_mrays": round(mrays(n_prim, k_p), 1),
             "incoherent_mrays": round(mrays(n_b, k_b), 1),
             "shadow_anyhit_mrays": round(mrays(n_s, k_s), 1)}
    log(f"plain: step {plain_step_ms:.3f} ms; kernels primary {k_p:.3f} ms, bounce {k_b:.3f} "
        f"ms, shadow {k_s:.3f} ms -> {plain}")
    # PCIe-inclusive rate of the primary set (host rays in, host hits out).
    t0 = time.perf_counter()
    scene.trace_host(prim_rays)
    e2e_ms = 1e3 * (time.perf_counter() - t0)
    plain["primary_host_to_host_mrays"] = round(mrays(n_prim, e2e_ms), 1)

    # ---------------------------------------------------------------- batched C1 frame
    frame = api.Frame(scene, [(d_prim, h_prim, False), (d_b, h_b, False)])
    autotune = None
    if args.c1_config and not args.no_tune:
        # the autotuner's pick on this box, reported beside the pinned configuration
        t0 = time.time()
        pick, pick_ms = frame.autotune(iters=3, cache_path="/tmp/mh_tune.json", force=True)
        autotune = {"config": pick, "ms": round(pick_ms, 4), "s": round(time.time() - t0, 1)}
        log(f"C1 frame autotune {autotune}")
    if args.c1_config:
        frame.set_config(json.loads(args.c1_config))
        tuned = frame.config()
        log(f"C1 frame configuration {tuned}")
    else:
        t0 = time.time()
        tuned, tuned_ms = frame.autotune(iters=3, cache_path="/tmp/mh_tune.json", force=True)
        log(f"C1 frame autotune {time.time() - t0:.1f}s: {tuned} {tuned_ms:.3f} ms")
    for _ in range(args.warmup):
        frame.launch()
    api.sync()
    # Outputs overwritten before the timed region: the hits read back after it are the
    # timed launches' own results (checked against the oracle by the CPU leg).
    for dh in (h_prim, h_b):
        dh.arrays["prim"].upload(np.full(dh.n, 777, np.int32))
        dh.arrays["t"].upload(np.full(dh.n, -1.0, np.float32))
    api.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        frame.launch()
    api.sync()
    ms_step = 1e3 * (time.perf_counter() - t0) / args.steps
    frame_hits = {"primary": h_prim.download(), "bounce": h_b.download()}
    k_frame = frame.time(args.steps)  # HIP events on the launch stream
    value = mrays(n_prim + n_b, ms_step)
    single = {}
    for name, dr, dh, anyh, n in (("coherent", d_prim, h_prim, False, n_prim),
                                  ("incoherent", d_b, h_b, False, n_b),
                                  ("shadow_anyhit", d_s, h_s, True, n_s)):
        if "singles" in skip:
            single[name] = {"mrays": plain[f"{name}_mrays"], "config": "plain"}
            continue
        f1 = api.Frame(scene, [(dr, dh, anyh)])
        cfg1, ms1 = f1.autotune(iters=3, cache_path="/tmp/mh_tune.json", force=True)
        f1.launch()
        single[name] = {"mrays": round(mrays(n, f1.time(args.steps)), 1), "config": cfg1}
        f1.close()
    log(f"batched: step {ms_step:.3f} ms ({value:.0f} Mrays/s), kernel {k_frame:.3f} ms; "
        f"single-set frames {single}")

    # Roofline of the dominant kernel, the persistent batch kernel over the C1 frame. Its
    # binding resource is the VALUs (the HBM traffic it causes is a small fraction of the
    # HBM peak: node and triangle records are served from the L2s and the Infinity
    # Cache). achieved = the algorithm's fp32 operations per launch (per ray: box tests x
    # BOX_OPS + triangle tests x TRI_OPS, the tests counted live above) / the launch time,
    # against one operation per lane per cycle on every SIMD lane of the chip.
    alg_ops = (n_prim * (BOX_OPS * nb_p + TRI_OPS * nt_p) +
               n_b * (BOX_OPS * nb_b + TRI_OPS * nt_b))
    achieved_tops = alg_ops / (k_frame * 1e-3) / 1e12
    alg_bytes = n_prim * bytes_p + n_b * bytes_b
    roofline = {"bound": "valu", "achieved": round(achieved_tops, 3), "peak": VALU_PEAK_TOPS,
                "unit": "Tops/s", "frac": round(achieved_tops / VALU_PEAK_TOPS, 4),
                "traffic": None,
                "kernel": "k_trace_batch (primary + bounce batches, closest-hit, KS)",
                "ops_per_test": {"box": BOX_OPS, "triangle": TRI_OPS},
                "ops_per_launch": int(alg_ops), "rays_per_launch": n_prim + n_b,
                "launch_ms": round(k_frame, 4),
                # what an uncached traversal would read (32 B per box test, 48 B per
                # triangle test, ray in and hit out), as a rate over the same launch
                "uncached_demand_gbs": round(alg_bytes / (k_frame * 1e-3) / 1e9, 1),
                "uncached_bytes_per_ray": {"primary": round(bytes_p, 1),
                                           "bounce": round(bytes_b, 1)}}
    if args.copy_probe_mb > 0:
        # A copy kernel of known size in the same process: the reference for scaling
        # FETCH_SIZE / WRITE_SIZE counter readings to bytes.
        nbytes = args.copy_probe_mb << 20
        a, b = api.DeviceArray(nbytes), api.DeviceArray(nbytes)
        for _ in range(3):
            L.check(L.load().mh_device_copy(b.ptr, a.ptr, nbytes, None))
        api.sync()
        log(f"copy probe: 3 x {nbytes} bytes")
        a.free()
        b.free()
    prof = load_profile()
    roofline["profile"] = PROFILE_JSON
    if prof is None:
        roofline["profile_status"] = "missing"
    elif prof["sources_sha256"] != kernel_sources_sha():
        roofline["profile_status"] = "stale: the kernel sources changed after it was taken"
    elif prof["config"] != tuned:
        roofline["profile_status"] = f"taken at another launch configuration {prof['config']}"
    else:
        # Counters of this kernel, these sources and this launch configuration: HBM bytes
        # per launch (FETCH_SIZE / WRITE_SIZE, copy-calibrated), VALU instructions issued
        # and the lanes they had active. Rates use this run's launch time.
        sec = k_frame * 1e-3
        roofline["profile_status"] = "current"
        roofline["traffic"] = int(prof["bytes_per_launch"])
        roofline["hbm"] = {"achieved_gbs": round(prof["bytes_per_launch"] / sec / 1e9, 1),
                           "peak_gbs": HBM_PEAK_GBS,
                           "frac": round(prof["bytes_per_launch"] / sec / 1e9 / HBM_PEAK_GBS, 4)}
        roofline["valu_issue"] = {
            "insts_per_launch": int(prof["valu_insts_per_launch"]),
            "frac": round(prof["valu_insts_per_launch"] / sec / VALU_ISSUE_PEAK, 4),
            "lanes_active": round(prof["lanes_active"], 4)}
        roofline["l2_hit"] = prof.get("l2_hit")

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "heightfield 1000x500 (997,002 tris), 1920x1080 primary + "
                               "1 diffuse bounce per hit, closest-hit, one batched launch",
                   "primary_rays": n_prim, "bounce_rays": n_b, "shadow_rays": n_s,
                   "launch": tuned, "autotune": autotune, "parallelism": "1 GPU"},
        "coherent_mrays": single["coherent"]["mrays"],
        "incoherent_mrays": single["incoherent"]["mrays"],
        "shadow_anyhit_mrays": single["shadow_anyhit"]["mrays"],
        "p50_ms": None, "p99_ms": None,
        "plain": plain, "build_s": round(build_s, 2),
        "tests_per_ray": {"primary": [round(nb_p, 2), round(nt_p, 3)],
                          "bounce": [round(nb_b, 2), round(nt_b, 3)],
                          "shadow": [round(nb_s, 2), round(nt_s, 3)]},
        "roofline": roofline,
    }
    frame.close()

    if "dropin" not in skip:
        out["drop_in"] = bench_dropin(api, L, scene, verts, idx, prim_rays,
                                      {"primary": (d_prim, h_prim, n_prim),
                                       "bounce": (d_b, h_b, n_b)}, args)

    if "build" not in skip:
        out["build"] = bench_build(api, L, verts, idx, build_s,
                                   {"coherent": (d_prim, h_prim, False, n_prim),
                                    "incoherent": (d_b, h_b, False, n_b),
                                    "shadow_anyhit": (d_s, h_s, True, n_s)}, args)

    if "c2" not in skip:
        out["c2_batching"] = bench_c2(api, L, scenes, scene, verts, idx, args)
        out["c2_incoherent"] = bench_c2(api, L, scenes, scene, verts, idx, args,
                                        incoherent=True)
    if "c3" not in skip:
        c3 = bench_c3(api, scenes, scene, args)
        out["c3_serving"] = c3
        if c3.get("sustained"):
            out["p50_ms"] = c3["sustained"]["p50_ms"]
            out["p99_ms"] = c3["sustained"]["p99_ms"]
    del d_prim, h_prim, d_b, h_b, d_s, h_s
    if "c4" not in skip:
        out["c4_dynamic"] = bench_c4(api, scenes, args)

    if args.no_cpu:
        out["cpu_baseline"], out["parity"] = None, None
    else:
        out["cpu_baseline"], out["parity"] = cpu_leg(verts, idx, scene, prim_rays, bounce,
                                                     frame_hits)
    print(json.dumps(out), flush=True)
    scene.close()
    return 0


def bench_dropin(api, L, scene, verts, idx, prim_rays, sets, args) -> dict:
    """The reference-exact packet path and the Manta drop-in on the C1 rays.

    packet_kernel: k_trace_packet (64-ray packets on one wave each, DynBVH::intersectNode
    semantics) on the device-resident primary rays in 8x8-tile order and on the bounce rays
    in their order, timed with HIP events. hipdynbvh: HipDynBVH::intersect called by one
    render thread per host core with 64-ray RayPackets (the way Manta's workers call
    DynBVH), over the primary rays; the shim gathers concurrent packets into launches and
    copies rays and hits over PCIe; wall time of the intersect phase (the build in
    preprocess is not timed)."""
    res = {"packet_kernel": {}}
    opts = api.trace_opts(64)
    for name, (dr, dh, n) in sets.items():
        ms = C.c_float()
        r, h = dr.struct(), dh.struct()
        L.check(L.load().mh_time_trace(scene.h, C.byref(r), C.byref(h), dr.n, 0, C.byref(opts),
                                       None, args.steps, C.byref(ms)))
        res["packet_kernel"][f"{name}_mrays"] = round(mrays(n, ms.value / args.steps), 1)
    from rfmanta_amd import build as B
    lib = C.CDLL(str(B.HARNESS))
    lib.mh_manta_render.restype = C.c_int

    class Io(C.Structure):
        _fields_ = [(k, C.c_void_p) for k in ("ox", "oy", "oz", "dx", "dy", "dz", "tmax",
                                              "masked", "prim", "t", "u", "v", "occluded")]
    n = len(prim_rays["ox"])
    arr = {k: np.ascontiguousarray(prim_rays[k], np.float32)
           for k in ("ox", "oy", "oz", "dx", "dy", "dz", "tmax")}
    outp = {"prim": np.empty(n, np.int32), "t": np.empty(n, np.float32),
            "u": np.empty(n, np.float32), "v": np.empty(n, np.float32)}
    io = Io(*[arr[k].ctypes.data for k in ("ox", "oy", "oz", "dx", "dy", "dz", "tmax")], None,
            *[outp[k].ctypes.data for k in ("prim", "t", "u", "v")], None)
    v = np.ascontiguousarray(verts, np.float32)
    ii = np.ascontiguousarray(idx, np.uint32)
    err = C.create_string_buffer(512)
    threads = host_cores()
    sec = C.c_double()
    rc = lib.mh_manta_render(C.c_void_p(v.ctypes.data), C.c_uint32(len(v)),
                             C.c_void_p(ii.ctypes.data), C.c_uint32(len(ii)), 0, threads,
                             C.byref(io), C.c_int64(n), 0, 0, err, 512, C.byref(sec))
    if rc != 0:
        raise RuntimeError(err.value.decode())
    res["hipdynbvh"] = {"threads": threads, "rays": n, "intersect_s": round(sec.value, 4),
                        "mrays": round(n / sec.value / 1e6, 1)}
    # the shim's packets agree with the packet kernel on the device (same traversal)
    dp = sets["primary"][1].download()
    res["hipdynbvh"]["same_as_packet_kernel"] = bool(np.array_equal(outp["prim"], dp["prim"]))
    log(f"drop-in: {res}")
    return res


def bench_build(api, L, verts, idx, host_build_s, sets, args) -> dict:
    """Time to a traceable tree: the host full-sweep SAH build (the reference's default)
    against the device build of the reference's approximate (binned) SAH variant, and the
    plain kernel's rate on the approximate tree for the C1 ray sets."""
    sc = api.Scene(verts, idx)
    times = []
    for _ in range(3):
        api.sync()
        t = time.perf_counter()
        sc.build(approximate=True)
        api.sync()
        times.append(1e3 * (time.perf_counter() - t))
    st = sc.stats()
    opts = api.trace_opts(1)
    rates = {}
    for name, (dr, dh, any_hit, n) in sets.items():
        ms = C.c_float()
        r, h = dr.struct(), dh.struct()
        sc.trace(dr, dh, any_hit=any_hit, opts=opts)
        L.check(L.load().mh_time_trace(sc.h, C.byref(r), C.byref(h), dr.n,
                                       L.MH_ANY_HIT if any_hit else 0, C.byref(opts), None,
                                       args.steps, C.byref(ms)))
        rates[f"{name}_mrays"] = round(n / (ms.value / args.steps * 1e-3) / 1e6, 1)
    sc.close()
    res = {"tris": len(idx), "host_full_sah_s": round(host_build_s, 3),
           "device_approx_ms": round(float(np.median(times)), 2),
           "device_approx_ms_all": [round(t, 2) for t in times],
           "approx_nodes": st["nodes"], "approx_depth": st["depth"],
           "approx_tree_plain": rates}
    log(f"build: {res}")
    return res


def bench_c2(api, L, scenes, scene, verts, idx, args, incoherent=False) -> dict:
    """64 uneven batches against the 1M-triangle tree: 64 plain launches vs one frame.
    incoherent=True keeps the batch sizes but makes every batch diffuse-bounce rays."""
    t0 = time.time()
    batches = scenes.config2_batches(verts, idx)
    if incoherent:
        start = 0
        inc = []
        for _, r in batches:
            n = len(r["ox"])
            inc.append(("bounce", scenes.batch_rays("bounce", n, verts, idx, 1, start)))
            start += n
        batches = inc
    dev = []
    for kind, r in batches:
        dr = api.DeviceRays.upload(r)
        dev.append((dr, api.DeviceHits.alloc(dr.n, any_hit=(kind == "shadow")), kind == "shadow"))
    total = sum(d[0].n for d in dev)
    log(f"C2: {len(dev)} batches, {total} rays, generated in {time.time() - t0:.1f}s")
    opts = api.trace_opts(1)
    lib = L.load()
    structs = [(d[0].struct(), d[1].struct(), d[0].n, int(d[2])) for d in dev]

    def plain_all():
        for r, h, n, a in structs:
            L.check(lib.mh_trace(scene.h, C.byref(r), C.byref(h), n, a, C.byref(opts), None))

    for _ in range(args.warmup):
        plain_all()
    api.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plain_all()
    api.sync()
    plain_ms = 1e3 * (time.perf_counter() - t0) / args.steps
    f = api.Frame(scene, dev)
    t0 = time.time()
    cfg, _ = f.autotune(iters=3, cache_path="/tmp/mh_tune.json", force=True)
    tune_s = time.time() - t0
    for _ in range(args.warmup):
        f.launch()
    api.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        f.launch()
    api.sync()
    frame_ms = 1e3 * (time.perf_counter() - t0) / args.steps
    stream = api.stream_create()
    f.capture(stream)
    for _ in range(args.warmup):
        f.replay(stream)
    api.sync(stream)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        f.replay(stream)
    api.sync(stream)
    graph_ms = 1e3 * (time.perf_counter() - t0) / args.steps
    api.stream_destroy(stream)
    f.close()
    res = {"rays": total, "batches": len(dev), "plain_launches_ms": round(plain_ms, 3),
           "frame_ms": round(frame_ms, 3), "graph_ms": round(graph_ms, 3),
           "plain_mrays": round(mrays(total, plain_ms), 1),
           "frame_mrays": round(mrays(total, frame_ms), 1),
           "graph_mrays": round(mrays(total, graph_ms), 1),
           "speedup_vs_plain": round(plain_ms / min(frame_ms, graph_ms), 3),
           "autotune_s": round(tune_s, 1), "config": cfg}
    log(f"C2: {res}")
    return res


def bench_c3(api, scenes, scene_hf, args) -> dict:
    t0 = time.time()
    v, i = scenes.soup(8_000_000, seed=0, scale=4.0)
    soup = api.Scene(v, i)
    del v, i
    soup.build()
    log(f"C3: soup 8M tris generated+built in {time.time() - t0:.1f}s")
    srv = api.Server([scene_hf, soup], hbm_budget=8 << 30, host_budget=2 << 30,
                     chunk_rays=1 << 21, max_active=64, shadows=True, light=scenes.C3_LIGHT,
                     launch=json.loads(args.server_launch) if args.server_launch else None)
    per_point = 64
    qps_list = (25, 50, 100, 200, 400, 800, 1600, 3200)
    reqs = scenes.c3_requests(per_point * (len(qps_list) + 1))
    # warm-up: a few requests back to back
    for sc, kw in reqs[:8]:
        srv.wait(srv.submit(sc, api.camera(**kw)), results=False)
    points = []
    k = 8
    rng = np.random.default_rng(2)
    for qps in qps_list:
        batch = reqs[k:k + per_point]
        k += per_point
        gaps = rng.exponential(1.0 / qps, size=len(batch))
        ids = []
        done = []

        def waiter():
            while True:
                if not ids and submit_done.is_set():
                    return
                if not ids:
                    time.sleep(0.0002)
                    continue
                rid = ids.pop(0)
                done.append(srv.wait(rid, results=False))

        submit_done = threading.Event()
        th = threading.Thread(target=waiter)
        th.start()
        t_start = time.perf_counter()
        t_next = t_start
        for (sc, kw), g in zip(batch, gaps):
            t_next += g
            now = time.perf_counter()
            if t_next > now:
                time.sleep(t_next - now)
            ids.append(srv.submit(sc, api.camera(**kw)))
        submit_done.set()
        th.join()
        lat = np.array([d["done_ms"] - d["submit_ms"] for d in done])
        rays = sum(d["rays"] for d in done)
        span_ms = max(d["done_ms"] for d in done) - min(d["submit_ms"] for d in done)
        pt = {"qps": qps, "requests": len(done), "p50_ms": round(float(np.percentile(lat, 50)), 2),
              "p99_ms": round(float(np.percentile(lat, 99)), 2),
              "mean_ms": round(float(lat.mean()), 2),
              "mrays": round(rays / (span_ms * 1e-3) / 1e6, 1),
              "achieved_qps": round(len(done) / (span_ms * 1e-3), 1)}
        points.append(pt)
        log(f"C3 {pt}")
        if pt["p99_ms"] > C3_SLO_MS:
            break
    st = srv.stats()
    srv.close()
    soup.close()
    ok = [p for p in points if p["p99_ms"] <= C3_SLO_MS]
    sustained = max(ok, key=lambda p: p["qps"]) if ok else None
    over = [p for p in points if p["p99_ms"] > C3_SLO_MS]
    # capacity: the highest offered rate whose p99 meets the objective, and the requests
    # per second actually completed there (the sweep doubles the rate until p99 exceeds it)
    capacity = ({"offered_qps": sustained["qps"], "achieved_qps": sustained["achieved_qps"],
                 "p99_ms": sustained["p99_ms"], "first_over_slo_qps":
                 min(p["qps"] for p in over) if over else None} if sustained else None)
    return {"points": points, "slo_p99_ms": C3_SLO_MS, "sustained": sustained,
            "capacity": capacity,
            "busy_ms": round(st["busy_ms"], 1), "rays": st["rays"],
            "requests": "orbit cameras, {256^2, 512^2, 1024x768, 1920x1080} x 1-4 spp, "
                        "primary + shadow rays, 2 scenes (heightfield 1M, soup 8M)"}


def bench_c4(api, scenes, args) -> dict:
    t0 = time.time()
    v, i = scenes.heightfield(4000, 1000, seed=0)
    sc = api.Scene(v, i)
    nv, nt = len(v), len(i)
    del v, i
    approx_ms = []
    for _ in range(2):  # the device build of the approximate-SAH variant, for comparison
        api.sync()
        t = time.perf_counter()
        sc.build(approximate=True)
        api.sync()
        approx_ms.append(1e3 * (time.perf_counter() - t))
    t = time.perf_counter()
    sc.build()
    host_s = time.perf_counter() - t
    st = sc.stats()
    log(f"C4: {nt} tris generated+built in {time.time() - t0:.1f}s (host full SAH "
        f"{host_s:.2f}s, device approximate {approx_ms} ms), {st}")
    srv = api.Server([sc], hbm_budget=2 << 30, host_budget=1 << 30, chunk_rays=1 << 22,
                     launch=json.loads(args.server_launch) if args.server_launch else None,
                     shadows=True, bounces=True, light=(2.0, 2.0, -2.0))
    cam = api.camera(**scenes.C4_CAMERA)
    refit_ms, frame_ms, rays, hits = [], [], 0, 0
    # The bounds-only refit on its own, for the refit bandwidth figure.
    for f in range(3):
        sc.displace_sine(0.01, 4.0, -0.1 * f)
        api.sync()
        t_f = time.perf_counter()
        sc.refit()
        api.sync()
        refit_ms.append(1e3 * (time.perf_counter() - t_f))
    refit_ms = refit_ms[1:]
    update_ms = []
    for f in range(args.c4_frames):
        t_f = time.perf_counter()
        sc.displace_sine(0.01, 4.0, 0.1 * f)
        sc.update()  # DynBVH::update: refit with tree rotations and leaf splits
        api.sync()
        t_r = time.perf_counter()
        o = srv.wait(srv.submit(0, cam), results=(f == 0))
        t_e = time.perf_counter()
        if f == 0:
            hits = int((o["prim"] >= 0).sum())
        update_ms.append(1e3 * (t_r - t_f))
        frame_ms.append(1e3 * (t_e - t_f))
        rays += o["rays"]
    st_end = sc.stats()
    srv.close()
    sc.close()
    total_s = sum(frame_ms) * 1e-3
    # Algorithmic bytes of the timed refit (mh_scene_refit): per triangle its id, indices and
    # vertices read (4 + 12 + 36) and its 48-byte record written; per leaf its list entry and
    # record read (4 + 32) and its box stored in the record and in the PairNode (2 x 24);
    # per inner record its list entry, child index and two child boxes read (4 + 4 + 48) and
    # its box stored twice (2 x 24).
    n_leaf = st["nodes"] - st["inner"]
    refit_bytes = nt * (52 + 48) + n_leaf * (36 + 48) + st["inner"] * (56 + 48)
    res = {"frames": args.c4_frames, "tris": nt, "rays_per_frame": rays // args.c4_frames,
           "host_full_sah_build_s": round(host_s, 2),
           "device_approx_build_ms": round(min(approx_ms), 2),
           "primary_hits_frame0": hits,
           "frame_ms_mean": round(float(np.mean(frame_ms)), 3),
           "frame_ms_p99": round(float(np.percentile(frame_ms, 99)), 3),
           "displace_update_ms_mean": round(float(np.mean(update_ms)), 3),
           "displace_update_ms_p99": round(float(np.percentile(update_ms, 99)), 3),
           "refit_ms_mean": round(float(np.mean(refit_ms)), 3),
           "refit_bytes": int(refit_bytes),
           "refit_gbs": round(refit_bytes / (np.mean(refit_ms) * 1e-3) / 1e9, 1),
           "refit_hbm_frac": round(refit_bytes / (np.mean(refit_ms) * 1e-3) / 8e12, 3),
           "nodes_before": st["nodes"], "nodes_after": st_end["nodes"],
           "depth_before": st["depth"], "depth_after": st_end["depth"],
           "fps": round(args.c4_frames / total_s, 2), "mrays": round(rays / total_s / 1e6, 1)}
    log(f"C4: {res}")
    return res


def kernel_sources_sha() -> str:
    """SHA-256 over the sources of the batch traversal kernel, in a fixed order."""
    import hashlib

    h = hashlib.sha256()
    for name in KERNEL_SOURCES:
        h.update((REPO / "rfmanta_amd" / "csrc" / name).read_bytes())
    return h.hexdigest()


def load_profile():
    """Counter summary of the dominant kernel (rocprofv3 --pmc runs of this script)."""
    p = REPO / PROFILE_JSON
    if not p.exists():
        return None
    with open(p) as f:
        return json.load(f)


def tie_classes(p_ref, t_ref, p_gpu, t_gpu):
    """Prim-id differences split into H1 (|dt| <= 1e-6 t: exact ties such as shared
    edges), H2 (|dt| < 1e-4: the Kensler-Shirley acceptance band, where the first-visited
    triangle wins) and others."""
    mis = p_ref != p_gpu
    dt = np.abs(t_ref.astype(np.float64) - t_gpu.astype(np.float64))
    scale = np.maximum(np.abs(t_ref.astype(np.float64)), 1e-30)
    h1 = mis & (dt <= 1e-6 * scale)
    h2 = mis & ~h1 & (dt < 1e-4)
    return int(h1.sum()), int(h2.sum()), int((mis & ~h1 & ~h2).sum())


def cpu_leg(verts, idx, scene, prim_rays, bounce, frame_hits):
    """The CPU baseline and, as its checker, the parity of the timed frame.

    Baseline: Manta's multithreaded SSE path as the oracle restates it (the reference's
    x86 build: 64-ray packets in 8x8-tile order, 4-wide SSE groups, rcpps reciprocals),
    one thread per host core, on a bounded sample of the C1 frame; the oracle's scalar
    build is timed beside it. Parity: the host tree against the oracle's build, the timed
    frame's hits bit for bit against the oracle's single-ray traversal, the prim ids
    against the reference's 64-ray packets (differences only in the tie classes H1/H2,
    counted) and the SSE build's packets against the scalar build's (H3: decisions the
    rcpps low bits move, counted)."""
    sys.path.insert(0, str(REPO / "oracle"))
    import pyoracle

    cores = host_cores()
    ids, nodes = scene.export_dynbvh()
    t0 = time.perf_counter()
    o_nodes, o_ids, _ = pyoracle.build_bvh(verts, idx)
    log(f"oracle build {time.perf_counter() - t0:.1f}s")
    parity = {"tree_equal": bool(np.array_equal(ids, o_ids) and
                                 nodes.tobytes() == o_nodes.tobytes()),
              "reference": "oracle single-ray traversal (bits of prim, t, u, v); 64-ray "
                           "packets (prim ids; H1/H2 tie classes); SSE-build packets vs "
                           "scalar-build packets (H3)"}
    mism_total = 0
    sets = (("primary", prim_rays), ("bounce", bounce))
    subs = {name: {k: rays[k] for k in ("ox", "oy", "oz", "dx", "dy", "dz", "tmax")}
            for name, rays in sets}
    for name, _ in sets:
        sub = subs[name]
        g = frame_hits[name]
        o1 = pyoracle.trace(verts, idx, o_nodes, o_ids, sub, packet_size=1, threads=cores)
        bad = g["prim"] != o1["prim"]
        for k in ("t", "u", "v"):
            bad |= g[k].view(np.uint32) != o1[k].view(np.uint32)
        o64 = pyoracle.trace(verts, idx, o_nodes, o_ids, sub, packet_size=64, threads=cores)
        h1, h2, other = tie_classes(o64["prim"], o64["t"], g["prim"], g["t"])
        s64 = pyoracle.trace(verts, idx, o_nodes, o_ids, sub, packet_size=64, threads=cores,
                             sse=True)
        flips = s64["prim"] != o64["prim"]
        f1, f2, f3 = tie_classes(o64["prim"], o64["t"], s64["prim"], s64["t"])
        parity[name] = {"rays": int(len(g["prim"])), "mismatch": int(bad.sum()),
                        "packets_h1": h1, "packets_h2": h2, "packets_other": other,
                        "sse_h3_prim_flips": int(flips.sum()),
                        "sse_h3_flips_in_tie_bands": f1 + f2,
                        "sse_h3_t_lowbit_rays": int((s64["t"].view(np.uint32) !=
                                                     o64["t"].view(np.uint32)).sum())}
        mism_total += int(bad.sum()) + other
    parity["mismatch"] = mism_total
    log(f"parity of the timed frame: {parity}")

    def rate(sse, threads, budget):
        """Frame rate: (primary + bounce rays) / (time of one pass over each set), each
        pass time averaged over repeated passes for about `budget` seconds per set."""
        frame_rays, frame_s, reps = 0, 0.0, {}
        for name, _ in sets:
            sub = subs[name]
            n = len(sub["ox"])
            done, t_set = 0, 0.0
            while t_set < budget or done < 1:
                t0 = time.perf_counter()
                pyoracle.trace(verts, idx, nodes, ids, sub, packet_size=64, threads=threads,
                               sse=sse)
                t_set += time.perf_counter() - t0
                done += 1
            reps[name] = done
            frame_rays += n
            frame_s += t_set / done
            log(f"cpu {'sse' if sse else 'scalar'} {name}: {done} x {n} rays in {t_set:.2f}s "
                f"on {threads} threads")
        return frame_rays / frame_s / 1e6, reps

    value, reps = rate(True, cores, 5.0)
    one, _ = rate(True, 1, 1.0)
    scalar, _ = rate(False, cores, 2.0)
    log(f"cpu baseline: SSE {value:.1f} Mrays/s on {cores} threads ({one:.2f} on one), "
        f"scalar build {scalar:.1f}")
    base = {"value": round(value, 3), "unit": "Mrays/s", "cores": cores,
            "one_thread": round(one, 3), "thread_scaling": round(value / one, 2),
            "kind": "port",
            "build": "the reference's SSE build (MANTA_SSE) restated by the oracle",
            "scalar_build_mrays": round(scalar, 3),
            "sample": f"the C1 frame's {len(prim_rays['ox'])} primary rays x{reps['primary']} "
                      f"and {len(bounce['ox'])} bounce rays x{reps['bounce']}, 64-ray packets "
                      f"in 8x8-tile order, closest-hit, about 10 s of CPU time; the rate is "
                      f"the frame's rays over one pass of each set"}
    return base, parity


if __name__ == "__main__":
    sys.exit(main())
