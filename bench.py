"""Headline benchmark: closest-hit Mrays/s at 1M triangles (coherent / incoherent) and
request latency p50/p99, on one MI355X.

Sections (all on synthetic inputs, generated at start-up and resident in HBM when timed):
  C1  heightfield 1000 x 500 vertices (997,002 triangles), DynBVH SAH build. 1920 x 1080
      pinhole primary rays (coherent, 8x8 tiles), one cosine-weighted diffuse bounce ray
      per primary hit (incoherent, PCG32 seed 1) and one shadow ray per hit (any-hit).
      One step = one frame of closest-hit work, primary + bounce rays, traced by the
      batched form: one persistent launch over both batches (autotuned configuration).
      `value` is that frame's closest-hit rays per second; the plain path (one launch per
      ray set) is timed beside it.
  C2  64 batches of 1k-2M rays (log-uniform; 40% primary, 30% shadow, 30% bounce; seed 1)
      against the C1 tree: 64 plain launches vs one persistent launch vs its graph replay.
  C3  request loop over two resident scenes (the C1 heightfield and an 8M-triangle soup in
      [-4,4]^3): Poisson arrivals swept over QPS, each request a random orbit camera at
      256^2 / 512^2 / 1024x768 / 1920x1080 with 1-4 spp (seed 2), primary + shadow rays;
      p50/p99 latency from submit to results in host memory.
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

METRIC = "Mrays/s closest-hit at 1M tris (coherent/incoherent); p50/p99 request latency"
HBM_PEAK_GBS = 8000.0
RAY_IN_BYTES = 32     # o.xyz, d.xyz, tmin, tmax
HIT_OUT_CLOSEST = 16  # prim, t, u, v
HIT_OUT_ANY = 4
BOX_BYTES = 32        # one reference BVHNode per box test
TRI_BYTES = 48        # 3 x u32 index + 3 x 12-byte vertex per triangle test
C3_SLO_MS = 100.0     # p99 objective for the sustained-load figure


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cores() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def mrays(n, ms):
    return n / (ms * 1e-3) / 1e6


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--skip", default="",
                    help="comma list of sections to skip: singles (per-set frames), c2, c3, c4")
    ap.add_argument("--c1-config", default="",
                    help="JSON launch configuration for the C1 frame instead of autotuning")
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
    t0 = time.perf_counter()
    for _ in range(args.steps):
        frame.launch()
    api.sync()
    ms_step = 1e3 * (time.perf_counter() - t0) / args.steps
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

    # Roofline of the dominant kernel: the persistent batch kernel over the C1 frame.
    alg_bytes = n_prim * bytes_p + n_b * bytes_b
    achieved_gbs = alg_bytes / (k_frame * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved_gbs, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_trace_batch (primary + bounce batches, closest-hit, KS)",
                "bytes_per_ray": {"primary": round(bytes_p, 1), "bounce": round(bytes_b, 1)},
                "rays_per_launch": n_prim + n_b, "launch_ms": round(k_frame, 4)}
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
    traffic = load_traffic()
    if traffic is not None:
        roofline["traffic"] = traffic["bytes_per_launch"]
        roofline["traffic_source"] = traffic["source"]

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "heightfield 1000x500 (997,002 tris), 1920x1080 primary + "
                               "1 diffuse bounce per hit, closest-hit, one batched launch",
                   "primary_rays": n_prim, "bounce_rays": n_b, "shadow_rays": n_s,
                   "launch": tuned, "parallelism": "1 GPU"},
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

    out["cpu_baseline"] = None if args.no_cpu else cpu_baseline(verts, idx, scene, prim_rays,
                                                                bounce)
    print(json.dumps(out), flush=True)
    scene.close()
    return 0


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


def c3_requests(scenes_mod, n, seed=2):
    """Request stream: (scene, camera kwargs) from PCG32(seed)."""
    r = scenes_mod.pcg32(seed, np.arange(n, dtype=np.uint64), 6)
    u = [scenes_mod.u01(r[k]) for k in range(6)]
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


def bench_c3(api, scenes, scene_hf, args) -> dict:
    t0 = time.time()
    v, i = scenes.soup(8_000_000, seed=0, scale=4.0)
    soup = api.Scene(v, i)
    del v, i
    soup.build()
    log(f"C3: soup 8M tris generated+built in {time.time() - t0:.1f}s")
    srv = api.Server([scene_hf, soup], hbm_budget=8 << 30, host_budget=2 << 30,
                     chunk_rays=1 << 21, max_active=64, shadows=True, light=(2.0, 6.0, -2.0))
    per_point = 64
    qps_list = (25, 50, 100, 200, 400)
    reqs = c3_requests(scenes, per_point * (len(qps_list) + 1))
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
    st = srv.stats()
    srv.close()
    soup.close()
    ok = [p for p in points if p["p99_ms"] <= C3_SLO_MS]
    sustained = max(ok, key=lambda p: p["qps"]) if ok else None
    return {"points": points, "slo_p99_ms": C3_SLO_MS, "sustained": sustained,
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
                     shadows=True, bounces=True, light=(2.0, 2.0, -2.0))
    cam = api.camera((0.0, 0.7, -0.9), (0.0, 0.0, 0.0), (0, 1, 0), 30.0, 1920, 1080, seed=4)
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
    # Refit traffic per frame (leaf boxes from indexed vertices, node boxes written, inner
    # nodes reading two child boxes) plus the displacement pass (read base, write verts).
    refit_bytes = nt * 48 + st["nodes"] * 32 + st["inner"] * 48 + nv * 24
    res = {"frames": args.c4_frames, "tris": nt, "rays_per_frame": rays // args.c4_frames,
           "host_full_sah_build_s": round(host_s, 2),
           "device_approx_build_ms": round(min(approx_ms), 2),
           "primary_hits_frame0": hits,
           "frame_ms_mean": round(float(np.mean(frame_ms)), 3),
           "frame_ms_p99": round(float(np.percentile(frame_ms, 99)), 3),
           "displace_update_ms_mean": round(float(np.mean(update_ms)), 3),
           "displace_update_ms_p99": round(float(np.percentile(update_ms, 99)), 3),
           "refit_ms_mean": round(float(np.mean(refit_ms)), 3),
           "refit_gbs": round(refit_bytes / (np.mean(refit_ms) * 1e-3) / 1e9, 1),
           "nodes_before": st["nodes"], "nodes_after": st_end["nodes"],
           "depth_before": st["depth"], "depth_after": st_end["depth"],
           "fps": round(args.c4_frames / total_s, 2), "mrays": round(rays / total_s / 1e6, 1)}
    log(f"C4: {res}")
    return res


def load_traffic():
    """HBM bytes per launch of the dominant kernel, from the committed counter summary."""
    p = REPO / "profiles" / "traffic_trace_batch_heightfield1M.json"
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
    # Thread scaling: one pass over each ray set on one thread (about a second), weighted
    # like the sample above (the same number of passes of each set).
    one_rays, one_s = 0, 0.0
    for name, rays in (("primary", prim_rays), ("bounce", bounce)):
        sub = {k: rays[k] for k in ("ox", "oy", "oz", "dx", "dy", "dz", "tmax")}
        t0 = time.perf_counter()
        pyoracle.trace(verts, idx, nodes, ids, sub, packet_size=64, threads=1)
        one_s += (time.perf_counter() - t0) * reps[name]
        one_rays += len(sub["ox"]) * reps[name]
    value = total_rays / total_s / 1e6
    one = one_rays / one_s / 1e6
    log(f"cpu baseline one thread: {one:.2f} Mrays/s ({value / one:.1f}x on {cores} threads)")
    return {"value": round(value, 3), "unit": "Mrays/s", "cores": cores,
            "one_thread": round(one, 3), "thread_scaling": round(value / one, 2),
            "kind": "port",
            "sample": f"the full frame's {len(prim_rays['ox'])} primary rays x{reps['primary']} "
                      f"and {len(bounce['ox'])} bounce rays x{reps['bounce']}, 64-ray packets "
                      f"in 8x8-tile order, closest-hit, ~10 s of CPU time"}


if __name__ == "__main__":
    sys.exit(main())
