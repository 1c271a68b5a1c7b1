"""Benchmark of the RLE voxel-column raycaster on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-batch]

Workload (plain): a 4096x4096x1024 seeded terrain (8-octave value noise,
5.6-5.9 runs per column, mips 0-5) resident in HBM, rendered at 1920x1080 along a
seeded 100-frame flythrough.  One step = one frame: ray map (host) + raycast
+ compose + gap-fill on the device; inputs are resident before timing starts.
`value` is Mrays/s (one ray = one screen pixel) over the K timed frames.

The roofline object is for the dominant kernel (the raycast): its algorithmic
bytes per launch come from the exact work counts of every timed frame
(8 B per pointer-map entry read, 2 B per extra RLE element, 2 B per colour,
4 B per texel cleared and per texel written) of the walk the kernels make,
in which a line ends as soon as no later step can draw (the same counts over
the reference's walk to the far distance are reported beside them), divided by its mean duration
from HIP events recorded on the launch stream around each launch.  The
events are recorded in a second timed pass over the same K frames: a timed
event costs about 3 us of device time per record, which would otherwise
count against `value` (three per frame: 5% of a 1080p frame).

The batched object renders 64 views over 4 resident scenes of the same kind
(resolutions uniform over 320x240, 1280x720, 1920x1080, 3840x2160) in one
persistent launch and compares with 64 plain launches of the same views.

The cpu_baseline object times the oracle (the CPU restatement, OpenMP over
ray lines) on the host cores for a bounded sample of the same frames.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rle-based-voxel-raycasting_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def raycast_bytes(c):
    return (8 * c["ptr_loads"] + 2 * c["rle_loads"] + 2 * c["col_loads"]
            + 4 * (c["cleared"] + c["col_loads"]))


def traffic_profile(kernel):
    return os.path.join("profiles", "%s_c2_1080p_pmc.txt" % kernel)


def issue_profile(kernel):
    return os.path.join("profiles", "%s_c2_1080p_sq.txt" % kernel)


def profile_json(path):
    """The JSON summary line of a profiles/ file, or {}."""
    try:
        for line in open(os.path.join(ROOT, path)):
            if line.startswith("JSON "):
                return json.loads(line[5:])
    except (OSError, ValueError):
        pass
    return {}


def measured_traffic(kernel):
    """HBM bytes per raycast launch of this workload, measured by the
    rocprofv3 FETCH_SIZE/WRITE_SIZE passes summarised in profiles/."""
    return profile_json(traffic_profile(kernel)).get("traffic_bytes")


def post_bytes(W, H):
    # compose: read one texel, write colour16 + depth32; gap-fill: read
    # colour16 + depth32, write colour16 (taps hit cache).
    return W * H * (4 + 2 + 4) + W * H * (2 + 4 + 2)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--c3", action="store_true",
                    help="also time 3840x2160 grazing views over a 16384^2 x 1024 map")
    ap.add_argument("--serve-requests", type=int, default=1500)
    ap.add_argument("--serve-rates", type=int, nargs="*", default=[200, 500, 1000, 2000])
    ap.add_argument("--serve-batch", type=int, default=48)
    ap.add_argument("--serve-streams", type=int, default=3)
    ap.add_argument("--serve-budget-gb", type=int, default=48)
    ap.add_argument("--serve-c3", type=int, default=3, metavar="K",
                    help="make K of the 12 resident scenes device copies of the "
                         "16384^2 x 1024 terrain (~33 GB each) instead of 4096^2 ones "
                         "(default 3: ~105 GB resident)")
    ap.add_argument("--serve-large-lanes", type=int, default=0,
                    help="server lanes reserved for 3840x2160 requests (0: any lane takes any)")
    ap.add_argument("--serve-large-waves", type=int, default=8,
                    help="persistent raycast wavefronts per CU on the large lanes")
    ap.add_argument("--serve-batch-mpix", type=float, default=0.0,
                    help="pixel cap of one served batch, in millions (0: none)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--waves-per-line", type=int, choices=(0, 1, 2), default=0,
                    help="wavefronts walking each ray line (2: DDA and drawing overlap; "
                         "0: the library's choice, two for frames at most 2560 pixels wide)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first run a plain copy kernel over MIB MiB (for counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(ray lines and scenes would shard across GPUs without collectives)", file=sys.stderr)
        sys.exit(2)

    import rlevox
    import workloads

    W, H = 1920, 1080
    t0 = time.time()
    hs = workloads.gen("c2", seed=args.seed)
    log("generated c2 scene in %.1fs (%.2f GB host)" % (time.time() - t0, hs.nbytes / 1e9))
    poses = workloads.flythrough(hs.mips[0], frames=100, seed=7)

    ctx = rlevox.Context(0)
    ctx.set_tuning(waves_per_block=args.waves_per_line)
    if args.calibrate_copy:
        nb = args.calibrate_copy << 20
        src, dst = ctx.alloc(nb), ctx.alloc(nb)
        src.fill_u32(0x01020304)
        for _ in range(2):
            dst.copy_from(src)
        ctx.sync()
        log("copy calibration: 2 x k_copy_u128 over %d MiB (read %d B + write %d B each)"
            % (args.calibrate_copy, nb, nb))
        src.free()
        dst.free()
    scene = ctx.upload(hs.mips)
    log("uploaded %.2f GB to HBM" % (scene.device_bytes / 1e9))
    stream = ctx.stream()
    rms = [rlevox.raymap(p, W, H) for p in poses]
    max_lines = max(r.map_line_count for r in rms)
    # the kernel the library runs for these frames (MI355X: 256 CUs)
    two = args.waves_per_line == 2 or (args.waves_per_line == 0 and max_lines <= 14 * 256)
    kernel = "k_raycast_wave2" if two else "k_raycast_wave"
    d_rb = ctx.alloc(max_lines * W * 4)
    d_clip = ctx.alloc(max_lines * 8)
    d_c0 = ctx.alloc(W * H * 2)
    d_col = ctx.alloc(W * H * 2)
    d_dep = ctx.alloc(W * H * 4)

    # exact work counts per pose (counting kernel, untimed): the walk as the
    # kernels make it (a line ends once nothing more can draw), and the
    # reference's walk of every line to its far distance
    counts = [ctx.raycast_counters(scene, rm, d_rb, d_clip, walk=True) for rm in rms]
    counts_ref = [ctx.raycast_counters(scene, rm, d_rb, d_clip) for rm in rms]
    log("work counts done")

    def frame(i, ev=None):
        # the host ray map is part of every step
        rm = rlevox.raymap(poses[i % len(poses)], W, H)
        if ev is not None:
            ctx.record(ev[0], stream)
        ctx.render_raybuf(scene, rm, d_rb, d_clip, stream)
        if ev is not None:
            ctx.record(ev[1], stream)
        ctx.compose(rm, d_rb, d_clip, d_c0, d_dep, stream)
        ctx.gapfill(W, H, d_c0, d_dep, d_col, stream)
        if ev is not None:
            ctx.record(ev[2], stream)

    for i in range(args.warmup):
        frame(i)
    ctx.sync(stream)
    K = args.steps
    evs = [(ctx.event(), ctx.event(), ctx.event()) for _ in range(K)]
    # throughput: the K frames back to back, nothing else on the stream
    ctx.sync(stream)
    t0 = time.perf_counter()
    for i in range(K):
        frame(i)
    ctx.sync(stream)
    wall = time.perf_counter() - t0
    # kernel durations: the same K frames with events around each stage
    for i in range(K):
        frame(i, evs[i])
    ctx.sync(stream)
    ray_ms = [ctx.elapsed_ms(e[0], e[1]) for e in evs]
    post_ms = [ctx.elapsed_ms(e[1], e[2]) for e in evs]
    ms_step = wall * 1000.0 / K
    mrays = K * W * H / wall / 1e6
    bytes_per = [raycast_bytes(counts[i % len(poses)]) for i in range(K)]
    mean_bytes = float(np.mean(bytes_per))
    mean_bytes_ref = float(np.mean([raycast_bytes(counts_ref[i % len(poses)]) for i in range(K)]))
    steps_frac = (sum(counts[i % len(poses)]["dda_steps"] for i in range(K))
                  / max(1, sum(counts_ref[i % len(poses)]["dda_steps"] for i in range(K))))
    mean_ray_ms = float(np.mean(ray_ms))
    achieved = mean_bytes / (mean_ray_ms * 1e-3) / 1e9
    log("plain: %.2f ms/frame, %.1f Mrays/s; raycast kernel mean %.3f ms (p50 %.3f, p99 %.3f), "
        "post %.3f ms; %.1f MB algorithmic per raycast" % (
            ms_step, mrays, mean_ray_ms, float(np.median(ray_ms)),
            float(np.percentile(ray_ms, 99)), float(np.mean(post_ms)), mean_bytes / 1e6))

    result = {
        "metric": "Mrays/s @1080p plain vs batched; p50/p99 latency at fixed request rate",
        "value": round(mrays, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+u16",
        "data": "synthetic",
        "config": {"workload": "plain 1920x1080 flythrough (100 seeded poses) over a resident "
                               "4096x4096x1024 RLE terrain, mips 0-5",
                   "scene_seed": args.seed, "pose_seed": 7, "resolution": [W, H],
                   "frame_ms_p50": round(float(np.median(ray_ms) + np.median(post_ms)), 4),
                   "raycast_ms_mean": round(mean_ray_ms, 4),
                   "post_ms_mean": round(float(np.mean(post_ms)), 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": measured_traffic(kernel), "kernel": kernel,
                     "bytes_per_launch": round(mean_bytes),
                     # the same counted over the reference's walk (every line to
                     # its far distance); the kernels' walk takes this share of
                     # its DDA steps
                     "bytes_reference_walk": round(mean_bytes_ref),
                     "dda_steps_vs_reference": round(steps_frac, 4),
                     "traffic_source": traffic_profile(kernel) + " (rocprofv3 FETCH_SIZE + "
                                       "WRITE_SIZE, copy-kernel calibrated)",
                     # diagnostics from the same profiles: L2 hit rate (TCC_HIT /
                     # (TCC_HIT + TCC_MISS)) and the share of all SIMD-cycles of the
                     # launch spent issuing VALU instructions (SQ_ACTIVE_INST_VALU)
                     "l2_hit": profile_json(traffic_profile(kernel)).get("l2_hit"),
                     "valu_busy": profile_json(issue_profile(kernel)).get("valu_busy")},
    }

    if not args.no_batch:
        for b in (d_rb, d_clip, d_c0, d_col, d_dep):
            b.free()
        d_rb = d_clip = d_c0 = d_col = d_dep = None
        scenes_h = [hs] + [workloads.gen("c2", seed=s) for s in (2, 3, 4)]
        log("generated c2 scenes 2-4")
        scenes = [scene] + [ctx.upload(h.mips) for h in scenes_h[1:]]
        try:
            result["batched"] = bench_batched(ctx, rlevox, workloads, scenes_h, scenes)
        except Exception as e:  # reported, not fatal: the plain line stands
            log("batched leg failed: %s" % e)
            result["batched"] = {"error": str(e)}
        if not args.no_serve:
            try:
                result["served"] = bench_served(ctx, rlevox, workloads, scenes_h, scenes, args)
            except Exception as e:
                log("served leg failed: %s" % e)
                result["served"] = {"error": str(e)}
        for sc in scenes[1:]:
            sc.free()

    if args.c3:
        try:
            result["c3"] = bench_c3(ctx, rlevox, workloads)
        except Exception as e:
            log("c3 leg failed: %s" % e)
            result["c3"] = {"error": str(e)}

    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(hs, poses, W, H, args.cpu_seconds)

    for b in (d_rb, d_clip, d_c0, d_col, d_dep):
        if b is not None:
            b.free()
    scene.free()
    ctx.close()
    print(json.dumps(result), flush=True)


def bench_batched(ctx, rlevox, workloads, scenes_h, scenes):
    views_spec = workloads.c4_views([h.mips[0] for h in scenes_h], n=64, seed=4)
    bufs = []
    views = []
    rays = 0
    for si, p, Wv, Hv in views_spec:
        dc = ctx.alloc(Wv * Hv * 2)
        dd = ctx.alloc(Wv * Hv * 4)
        bufs += [dc, dd]
        views.append(rlevox.make_view(scenes[si], p, Wv, Hv, dc, dd))
        rays += Wv * Hv
    stream = ctx.stream()
    # plain: one launch sequence per view
    for _ in range(2):
        for v, (si, p, Wv, Hv) in zip(views, views_spec):
            ctx.render(scenes[si], p, Wv, Hv, _Buf(v.d_colour), _Buf(v.d_depth), stream=stream)
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        for v, (si, p, Wv, Hv) in zip(views, views_spec):
            ctx.render(scenes[si], p, Wv, Hv, _Buf(v.d_colour), _Buf(v.d_depth), stream=stream)
    ctx.sync(stream)
    plain_s = (time.perf_counter() - t0) / reps
    out = {"views": len(views), "scenes": 4, "rays_per_batch": rays,
           "plain_mrays_s": round(rays / plain_s / 1e6, 2),
           "plain_ms": round(plain_s * 1e3, 3)}
    for graph in (False, True):
        key = "graph" if graph else "persistent"
        for _ in range(2):
            ctx.render_batch(views, stream=stream, graph=graph)
        ctx.sync(stream)
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.render_batch(views, stream=stream, graph=graph)
        ctx.sync(stream)
        bs = (time.perf_counter() - t0) / reps
        out[key + "_mrays_s"] = round(rays / bs / 1e6, 2)
        out[key + "_ms"] = round(bs * 1e3, 3)
    best = max(out.get("persistent_mrays_s", 0), out.get("graph_mrays_s", 0))
    out["speedup_vs_plain"] = round(best / out["plain_mrays_s"], 3)
    log("batched: %s" % out)
    for b in bufs:
        b.free()
    return out


def bench_c3(ctx, rlevox, workloads, frames=8, reps=3):
    """Plain kernel on a 16384x16384x1024 map (mips 0-7, level 0
    beyond 2^32 slab words) at 3840x2160, grazing poses |pitch| 0.6-5 deg."""
    W, H = 3840, 2160
    t0 = time.time()
    hs = workloads.gen("c3", seed=1)
    log("generated c3 scene in %.1fs (%.2f GB host, level-0 offset shift %d)"
        % (time.time() - t0, hs.nbytes / 1e9, hs.mips[0]["shift"]))
    scene = ctx.upload(hs.mips)
    poses = workloads.grazing(hs.mips[0], n=frames, seed=11)
    rms = [rlevox.raymap(p, W, H) for p in poses]
    lines = max(r.map_line_count for r in rms)
    d_rb, d_clip = ctx.alloc(lines * W * 4), ctx.alloc(lines * 8)
    d_c0, d_col, d_dep = ctx.alloc(W * H * 2), ctx.alloc(W * H * 2), ctx.alloc(W * H * 4)
    counts = [ctx.raycast_counters(scene, rm, d_rb, d_clip, walk=True) for rm in rms]
    counts_ref = [ctx.raycast_counters(scene, rm, d_rb, d_clip) for rm in rms]
    stream = ctx.stream()
    ev = [(ctx.event(), ctx.event()) for _ in range(frames * reps)]
    for rm in rms:  # warm-up
        ctx.render_raybuf(scene, rm, d_rb, d_clip, stream)
    ctx.sync(stream)
    t0 = time.perf_counter()
    for r in range(reps):
        for i, p in enumerate(poses):
            rm = rlevox.raymap(p, W, H)
            ctx.record(ev[r * frames + i][0], stream)
            ctx.render_raybuf(scene, rm, d_rb, d_clip, stream)
            ctx.record(ev[r * frames + i][1], stream)
            ctx.compose(rm, d_rb, d_clip, d_c0, d_dep, stream)
            ctx.gapfill(W, H, d_c0, d_dep, d_col, stream)
    ctx.sync(stream)
    wall = time.perf_counter() - t0
    ray_ms = float(np.mean([ctx.elapsed_ms(a, b) for a, b in ev]))
    mean_bytes = float(np.mean([raycast_bytes(c) for c in counts]))
    out = {"workload": "3840x2160 grazing views (8 seeded poses, |pitch| 0.6-5 deg) over "
                       "a resident 16384x16384x1024 RLE terrain, mips 0-7",
           "scene_gb": round(scene.device_bytes / 1e9, 2), "offset_shift": hs.mips[0]["shift"],
           "mrays_s": round(frames * reps * W * H / wall / 1e6, 2),
           "ms_per_frame": round(wall * 1e3 / (frames * reps), 3),
           "raycast_ms_mean": round(ray_ms, 3),
           "raycast_bytes_per_launch": round(mean_bytes),
           "raycast_bytes_reference_walk": round(float(np.mean([raycast_bytes(c)
                                                                for c in counts_ref]))),
           "raycast_gbs": round(mean_bytes / (ray_ms * 1e-3) / 1e9, 2)}
    # measured HBM traffic and L2 hit rate of the same workload (rocprofv3
    # FETCH_SIZE / WRITE_SIZE / TCC passes, copy-kernel calibrated)
    prof = os.path.join("profiles", "k_raycast_wave_c3_4k_grazing.txt")
    pj = profile_json(prof)
    if pj.get("traffic_bytes"):
        out["hbm_traffic_per_launch"] = pj["traffic_bytes"]
        out["hbm_gbs"] = round(pj["traffic_bytes"] / (ray_ms * 1e-3) / 1e9, 2)
        out["hbm_frac"] = round(out["hbm_gbs"] / HBM_PEAK_GBS, 5)
        out["l2_hit"] = pj.get("l2_hit")
        out["traffic_source"] = prof
    log("c3: %s" % out)
    for b in (d_rb, d_clip, d_c0, d_col, d_dep):
        b.free()
    scene.free()
    hs.close()
    return out


C5_SIZES = [((640, 480), 0.5), ((1920, 1080), 0.35), ((3840, 2160), 0.15)]


def bench_served(ctx, rlevox, workloads, scenes_h, scenes, args):
    """Served requests: 12 resident scenes, Poisson arrivals, mixed resolutions.

    The 12 resident scenes are the 4 generated c2 scenes uploaded 3 times
    each (distinct device copies); with --serve-c3 K, K of the copies are
    replaced by device copies of the 16384^2 x 1024 terrain.  Requests draw a scene, a seeded pose and a
    resolution (640x480 / 1920x1080 / 3840x2160 at 0.5 / 0.35 / 0.15) and are
    rendered into device framebuffers.  Three runs over the same requests:
    plain (one synchronous rlevox_render per request, back to back), served
    open-loop at the given arrival rate (latency from submit to framebuffer
    ready), and served saturated (all submitted at once: peak throughput).
    """
    import threading
    resident = list(scenes) + [ctx.upload(h.mips) for h in scenes_h for _ in range(2)]
    resident_h = list(scenes_h) + [h for h in scenes_h for _ in range(2)]
    big_h = None
    if args.serve_c3 > 0:
        k = min(args.serve_c3, len(resident) - len(scenes))
        for s in resident[len(resident) - k:]:
            s.free()
        big_h = workloads.gen("c3", seed=1)
        resident = resident[:len(resident) - k] + [ctx.upload(big_h.mips) for _ in range(k)]
        resident_h = resident_h[:len(resident_h) - k] + [big_h] * k
    rng = np.random.default_rng(55)
    n = args.serve_requests
    reqs = []
    for i in range(n):
        u = rng.random()
        (W, H) = C5_SIZES[0][0] if u < 0.5 else (C5_SIZES[1][0] if u < 0.85 else C5_SIZES[2][0])
        si = int(rng.integers(len(resident)))
        pose = workloads.random_views(resident_h[si].mips[0], 1, int(rng.integers(1 << 30)))[0]
        reqs.append((si, pose, W, H))
    offs, total = [], 0
    for _, _, W, H in reqs:
        offs.append(total)
        total += (W * H * 6 + 255) // 256 * 256
    arena = ctx.alloc(total)
    base = arena.ptr.value
    dcol = [base + o for o in offs]
    ddep = [base + o + (W * H * 2 + 255) // 256 * 256 for o, (_, _, W, H) in zip(offs, reqs)]
    rays = sum(W * H for _, _, W, H in reqs)
    log("served: %d requests, %.1f Mrays, %d resident scenes (%.1f GB)" % (
        n, rays / 1e6, len(resident), sum(s.device_bytes for s in resident) / 1e9))

    # plain per-request path
    stream = ctx.stream()
    for k in range(min(20, n)):
        si, p, W, H = reqs[k]
        ctx.render(resident[si], p, W, H, _Buf(dcol[k]), _Buf(ddep[k]), stream=stream)
    t0 = time.perf_counter()
    lat_plain = []
    for k, (si, p, W, H) in enumerate(reqs):
        t1 = time.perf_counter()
        ctx.render(resident[si], p, W, H, _Buf(dcol[k]), _Buf(ddep[k]), stream=stream)
        lat_plain.append((time.perf_counter() - t1) * 1e3)
    plain_s = time.perf_counter() - t0
    out = {"requests": n, "resident_scenes": len(resident), "rays": rays,
           "resident_gb": round(sum(s.device_bytes for s in resident) / 1e9, 2),
           "resident_16384_copies": args.serve_c3,
           "resident_mix": "%d device copies of the 4096^2x1024 terrains (seeds 1-4), %d of "
                           "the 16384^2x1024 terrain (seed 1)" % (len(resident) - args.serve_c3,
                                                                  args.serve_c3),
           "mix": "640x480:0.5 1920x1080:0.35 3840x2160:0.15",
           "server": {"streams": args.serve_streams, "max_batch_views": args.serve_batch,
                      "large_lanes": args.serve_large_lanes,
                      "large_waves": args.serve_large_waves,
                      "batch_mpix": args.serve_batch_mpix},
           "plain_mrays_s": round(rays / plain_s / 1e6, 2)}

    def run(rate):
        srv = rlevox.Server(ctx, memory_budget=args.serve_budget_gb << 30,
                            max_batch_views=args.serve_batch, max_queue=1 << 20,
                            streams=args.serve_streams, large_lanes=args.serve_large_lanes,
                            large_pixels=3840 * 2160, large_waves=args.serve_large_waves,
                            batch_pixels=int(args.serve_batch_mpix * 1e6))
        tickets = [None] * n
        t_start = time.perf_counter()
        if rate is None:
            for k, (si, p, W, H) in enumerate(reqs):
                tickets[k] = srv.submit(resident[si], p, W, H, dcol[k], ddep[k])
        else:
            arrivals = np.cumsum(np.random.default_rng(77).exponential(1.0 / rate, n))

            def submitter():
                for k, (si, p, W, H) in enumerate(reqs):
                    d = arrivals[k] - (time.perf_counter() - t_start)
                    if d > 0:
                        time.sleep(d)
                    tickets[k] = srv.submit(resident[si], p, W, H, dcol[k], ddep[k])
            th = threading.Thread(target=submitter)
            th.start()
        lat = []
        for k in range(n):
            while tickets[k] is None:
                time.sleep(0.0002)
            lat.append(srv.wait(tickets[k])[2])
        wall = time.perf_counter() - t_start
        if rate is not None:
            th.join()
        st = srv.stats()
        srv.close()
        lat = np.array(lat)
        r1080 = np.array([W == 1920 for _, _, W, _ in reqs])
        return {"mrays_s": round(rays / wall / 1e6, 2), "wall_s": round(wall, 3),
                "p50_ms": round(float(np.percentile(lat, 50)), 3),
                "p99_ms": round(float(np.percentile(lat, 99)), 3),
                "p99_1080p_ms": round(float(np.percentile(lat[r1080], 99)), 3),
                "batches": int(st["batches"]), "busy_s": round(st["busy_seconds"], 3),
                "buffer_growths": int(st["buffer_growths"])}

    run(None)  # warm-up: lane graphs, scratch growth
    sat = run(None)
    out["saturated"] = sat
    out["speedup_vs_plain"] = round(sat["mrays_s"] / out["plain_mrays_s"], 3)
    out["open_loop"] = {}
    for rate in args.serve_rates:
        r = run(float(rate))
        out["open_loop"][str(rate)] = r
        log("served at %s req/s: %s" % (rate, r))
    out["plain_latency_p50_ms"] = round(float(np.percentile(lat_plain, 50)), 3)
    log("served: %s" % {k: v for k, v in out.items() if k != "open_loop"})
    arena.free()
    for s in resident[len(scenes):]:
        s.free()
    if big_h is not None:
        big_h.close()
    return out


class _Buf:
    def __init__(self, ptr):
        import ctypes
        self.ptr = ctypes.c_void_p(ptr)


def cpu_baseline(hs, poses, W, H, seconds):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle
    O = Oracle()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    n = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        O.render(hs.mips, poses[n % len(poses)], W, H, threads=threads)
        n += 1
    dt = time.perf_counter() - t0
    v = n * W * H / dt / 1e6
    log("cpu baseline: %d frames in %.1fs on %d threads -> %.2f Mrays/s" % (n, dt, threads, v))
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": round(v, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "cpu": model, "nproc": os.cpu_count(),
            "sample": "%d frames of the 100-pose flythrough at %dx%d (%.1f s: ray map + raycast "
                      "+ compose + gap-fill), OpenMP over ray lines" % (n, W, H, dt)}


if __name__ == "__main__":
    main()
