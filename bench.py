"""Benchmark of the MI355X path tracer (one GPU).

Default workload (the metric's configuration), ``--workload plain_1080p``:
1920x1080, 256 random spheres + a ground sphere, 8 point lights, 8 bounces,
1 sample per pixel per frame, accumulated over the timed frames; synthetic
scene from the seeded generator.  One "step" is one frame (one path sample
for every pixel).  `value` is the single-scene renderer's production path:
the persistent wavefront kernel (several frames per launch) with the sphere
BVH.  The plain one-lane-per-pixel kernel and the flat loop over all spheres
(the reference's algorithm) are timed on the same frames and reported in
`variants`.  All four produce bit-identical images.

The default run also carries, in the same JSON line:
  batch       the 64-scene mixed workload: batched persistent renderer
              (autotuned, hipGraph) vs the plain kernel rendering the scenes
              one by one with the same sphere acceleration; `speedup`
  serve       Poisson request streams through the serving loop at 50, 100,
              200 and 400 requests/s (samples/s, p50/p99 latency, HBM
              high-water mark per rate)
  cpu_baseline  the oracle on the host cores over frames of the same scene

Other workloads (one JSON line each):
  --workload batch_mix   the batch leg alone (with its own knobs)
  --workload serve       one Poisson rate (--rate, --requests)
  --workload precision   naive vs Kahan FP32 accumulation against a float64
                         host sum at 1024x1024 and 4096 spp
  --workload copy        a device-to-device copy of a known size: the
                         reference reading for rocprofv3's FETCH_SIZE and
                         WRITE_SIZE counters

The roofline entry is the VALU bound of the dominant kernel: 17 algorithmic
FLOP per ray-sphere test and 12 per BVH node box test (6 FMA; FMA = 2), both
counted exactly by the kernels, divided by that kernel's average launch
duration from HIP events recorded on the renderer's stream around the path
kernel alone (the persistent kernel's launches are each followed by a short
accumulate kernel, which is timed in `gpu_ms_per_step` but not here), against
MI355X's published FP32 vector peak.
"""
import argparse
import json
import os
import re
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_VALU_PEAK_TFLOPS = 157.3   # MI355X published FP32 vector peak
FLOP_PER_TEST = 17
FLOP_PER_BOX = 12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=256, help="timed frames")
    ap.add_argument("--warmup", type=int, default=8, help="untimed frames")
    ap.add_argument("--workload", choices=["plain_1080p", "batch_mix", "serve", "precision", "copy"],
                    default="plain_1080p")
    ap.add_argument("--precision-size", type=int, default=1024, help="precision: image side in pixels")
    ap.add_argument("--precision-spp", type=int, default=4096, help="precision: samples per pixel")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="wall time budget of the CPU baseline")
    ap.add_argument("--exact-loop", action="store_true", help="use the exact (slow) closest-hit loop")
    ap.add_argument("--kernel", choices=["plain", "wavefront", "both"], default="both")
    ap.add_argument("--accel", choices=["bvh", "none", "both"], default="both",
                    help="sphere BVH, flat loop over all spheres, or both")
    ap.add_argument("--scenes", type=int, default=64, help="batch_mix: number of scenes")
    ap.add_argument("--tuning", default=None,
                    help="batch_mix: fixed chunk_px,workgroups,frames_per_launch instead of the autotuner")
    ap.add_argument("--no-plain", action="store_true", help="batch_mix: skip the plain-kernel comparison")
    ap.add_argument("--spp-scale", type=float, default=1.0, help="batch_mix: scale every scene's spp")
    ap.add_argument("--rate", type=float, default=50.0, help="serve: Poisson arrival rate (requests/s)")
    ap.add_argument("--requests", type=int, default=64, help="serve: number of requests")
    ap.add_argument("--frames-per-tick", type=int, default=0,
                    help="serve: frames per request per launch (0: the server's default)")
    ap.add_argument("--tick-mpaths", type=int, default=0,
                    help="serve: million paths per launch before admitting more (0: the server's default)")
    ap.add_argument("--no-serve", action="store_true", help="plain_1080p: skip the serving runs")
    ap.add_argument("--no-batch", action="store_true", help="plain_1080p: skip the batch leg")
    ap.add_argument("--serve-rates", default="50,100,200,400", help="plain_1080p: Poisson rates of the serve leg")
    ap.add_argument("--serve-requests", type=int, default=128, help="plain_1080p: requests per serve rate")
    ap.add_argument("--copy-bytes", type=int, default=268435456, help="copy: bytes per copy")
    return ap.parse_args()


def traffic_from_profile(name, kernel, frames):
    """HBM bytes per launch of `kernel` for a launch of `frames` frames, from
    a committed rocprof summary (its FETCH_SIZE + WRITE_SIZE lines,
    calibrated), or None.  Summaries are named ..._f<frames>.txt after the
    launch length they profiled; the one of the same length is read as is,
    otherwise the closest one is scaled per frame (the path kernel's traffic
    is the samples it writes and the first-bounce records it reads, both
    proportional to the frames of the launch).  Returns (bytes, basis)."""
    best = None
    for fn in sorted(os.listdir(os.path.join(ROOT, "profiles"))):
        m = re.fullmatch(re.escape(name) + r"_f(\d+)\.txt", fn)
        if m:
            f = int(m.group(1))
            if best is None or abs(f - frames) < abs(best[0] - frames):
                best = (f, fn)
    if best is None:
        return None, None
    mb = [float(m) for line in open(os.path.join(ROOT, "profiles", best[1]))
          if re.search(rf"::{kernel}[(<]", line) for m in re.findall(r"= ([0-9.]+) MB \(calibrated\)", line)]
    if len(mb) < 2:
        return None, None
    pf = best[0]
    if pf == frames:
        return sum(mb) * 1e6, f"profiles/{best[1]} (a {pf}-frame launch, as timed here)"
    return sum(mb) * 1e6 * frames / pf, f"profiles/{best[1]} (a {pf}-frame launch) scaled to {frames} frames"


def host_cores():
    """Cores this process may run on, and the host's count.  A GPU box grants
    each GPU a share of 16 host cores, so the baseline uses at most 16."""
    visible = os.cpu_count() or 1
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = visible
    return max(1, min(16, allowed)), allowed, visible


def cpu_baseline(sc, target_s):
    """The oracle on the host: whole frames of the same scene, row-parallel over
    every core the process may use (at most 16, a GPU's share of the box),
    repeated until about `target_s` seconds elapse."""
    import numpy as np
    import oracle as O
    cam = O.camera_from(sc.camera)
    osc = O.Scene(sc.spheres, sc.materials, sc.lights)
    st = O.settings(sc.settings.max_bounces, sc.settings.sky_light, sc.settings.accumulation)
    threads, allowed, visible = host_cores()
    n = cam.width * cam.height
    acc = np.zeros((n, 4), np.float32)
    frames, rays = 0, 0
    t0 = time.perf_counter()
    while True:
        acc, _, _, r = O.render_frames_mt(cam, osc, st, frames + 1, 1, threads, accum=acc)
        frames += 1
        rays += r
        dt = time.perf_counter() - t0
        if dt >= target_s:
            break
    samples = n * frames
    # one more frame on a single thread: the per-core rate
    t1 = time.perf_counter()
    O.render_frames_mt(cam, osc, st, frames + 1, 1, 1, accum=acc)
    dt1 = time.perf_counter() - t1
    return {"value": samples / dt, "unit": "path samples/s", "cores": threads, "kind": "port",
            "sample": f"{frames} full frame(s) of the same 1080p workload ({samples} samples, {rays} rays) "
                      f"in {dt:.2f} s on {threads} threads; one more frame on 1 thread in {dt1:.2f} s",
            "single_thread_value": n / dt1, "cores_allowed": allowed, "nproc": visible,
            "cores_note": "threads = the cores this process may use, at most 16 (one GPU's share of the box)"}


def time_single_scene(sc, kernel, accel, steps, warmup):
    """Times `steps` frames of one render call.  The renderer reserves its
    buffers for calls of `steps` frames and the warm-up frames use a camera
    moved by 1e-3, so that the renderer's buffers exist but its per-camera
    work (primary directions, first-bounce cache) is redone inside the timed
    region, as a fresh camera would need it."""
    import dataclasses
    from ataraxia_amd import Renderer
    from ataraxia_amd._lib import lib
    cam = sc.camera
    warm_cam = dataclasses.replace(cam, position=(cam.position[0] + 1e-3,) + tuple(cam.position[1:]))
    r = Renderer()
    r.set_kernel(kernel)
    r.set_accel(accel)
    r.on_resize(cam.width, cam.height)
    r.set_settings(sc.settings)
    r.set_scene(sc)
    r.reserve(steps)
    if warmup:
        r.render_frames(warm_cam, warmup)
    r.reset_frame_index()
    lib.atx_device_synchronize()
    t0 = time.perf_counter()
    r.render_frames(cam, steps)
    lib.atx_device_synchronize()
    wall = time.perf_counter() - t0
    st = r.stats()
    r.close()
    return wall, st


def algo_flops(st):
    return FLOP_PER_TEST * st["sphere_tests"] + FLOP_PER_BOX * st["box_tests"]


def roofline(st, traffic, basis=None):
    launches = max(1, st["launches"])
    kernel_ms = st["path_ms"] / launches
    achieved = algo_flops(st) / launches / (kernel_ms * 1e-3) / 1e12
    return {"bound": "valu", "achieved": achieved, "peak": FP32_VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP32_VALU_PEAK_TFLOPS, "traffic": traffic, "traffic_basis": basis,
            "kernel_ms": kernel_ms,
            "launches": launches, "sphere_tests_per_launch": st["sphere_tests"] / launches,
            "box_tests_per_launch": st["box_tests"] / launches, "rays_per_launch": st["rays"] / launches,
            "reused_rays_per_launch": st["reused_rays"] / launches,
            "exact_retraces": st["exact_retraces"]}


def bench_plain_1080p(args):
    from ataraxia_amd import (ATX_ACCEL_BVH, ATX_ACCEL_NONE, ATX_KERNEL_PLAIN, ATX_KERNEL_WAVEFRONT,
                              plain_1080p_scene)
    sc = plain_1080p_scene()
    cam = sc.camera
    samples = cam.width * cam.height * args.steps
    kernels = [k for k in ("wavefront", "plain") if args.kernel in (k, "both")]
    accels = [a for a in ("bvh", "flat") if args.accel in ({"bvh": "bvh", "flat": "none"}[a], "both")]
    kid = {"plain": ATX_KERNEL_PLAIN, "wavefront": ATX_KERNEL_WAVEFRONT}
    aid = {"bvh": ATX_ACCEL_BVH, "flat": ATX_ACCEL_NONE}
    variants = {}
    for k in kernels:
        for a in accels:
            wall, st = time_single_scene(sc, kid[k], aid[a], args.steps, args.warmup)
            frames_per_launch = args.steps / max(1, st["launches"])
            traffic, basis = traffic_from_profile(f"plain_1080p_k_render_{k}_{a}", f"k_render_{k}",
                                                  int(round(frames_per_launch)))
            variants[f"{k}/{a}"] = {"value": samples / wall, "ms_per_step": wall * 1e3 / args.steps,
                                    "gpu_ms_per_step": st["gpu_ms"] / args.steps,
                                    "roofline": roofline(st, traffic, basis)}
    main_variant = next(iter(variants))   # wavefront/bvh unless restricted
    v = variants[main_variant]
    out = {
        "metric": "path samples/s (1080p, 8 bounces) + p50/p99 request latency in serve loop",
        "value": v["value"],
        "unit": "path samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": v["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded splitmix64 scene generator)",
        "config": {"workload": "plain_1080p: 1920x1080, 256 spheres + ground, 8 lights, 8 bounces, "
                               "1 spp/frame accumulated", "renderer": "single-scene",
                   "kernel": main_variant.split("/")[0], "accel": main_variant.split("/")[1],
                   "spheres": int(sc.num_spheres), "lights": int(len(sc.lights)), "frames": args.steps},
        "roofline": v["roofline"],
        "variants": variants,
    }
    if not args.no_batch:
        bm = bench_batch_mix(args)
        out["batch"] = {k: bm[k] for k in ("config", "batched", "plain", "speedup")}
    if not args.no_serve:
        out["serve"] = [bench_serve(args, rate=float(r), requests=args.serve_requests)
                        for r in args.serve_rates.split(",")]
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(sc, args.cpu_seconds)
    return out


def bench_batch_mix(args):
    """Batched persistent renderer vs the plain kernel on the same 64 scenes
    (both with the same sphere acceleration)."""
    from ataraxia_amd import ATX_ACCEL_BVH, ATX_ACCEL_NONE, ATX_KERNEL_PLAIN, Batch, Renderer, batch_mix_scenes
    from ataraxia_amd._lib import lib
    accel = ATX_ACCEL_NONE if args.accel == "none" else ATX_ACCEL_BVH
    scenes = batch_mix_scenes()[: args.scenes]
    for s in scenes:
        s.spp = max(1, int(round(s.spp * args.spp_scale)))
    samples = sum(s.camera.width * s.camera.height * s.spp for s in scenes)
    b = Batch()
    b.set_accel(accel)
    for s in scenes:
        b.add_scene(s)
    t0 = time.perf_counter()
    if args.tuning:
        chunk, wg, fpl = (int(v) for v in args.tuning.split(","))
        b.set_tuning(chunk_px=chunk, workgroups=wg, frames_per_launch=fpl)
        tuning = {"chunk_px": chunk, "workgroups": wg, "frames_per_launch": fpl, "fixed": True}
    else:
        cache = os.environ.get("ATX_AUTOTUNE_CACHE", os.path.join(tempfile.gettempdir(), "atx_autotune.txt"))
        tuning = b.autotune(cache or None)
    tune_s = time.perf_counter() - t0
    b.render(max_frames=2)       # warm-up: graph capture + first frames
    lib.atx_device_synchronize()
    t0 = time.perf_counter()
    bst = b.render()
    lib.atx_device_synchronize()
    bwall = time.perf_counter() - t0
    b.close()
    # plain kernel, scene after scene, one launch per frame
    r = Renderer()
    r.set_kernel(ATX_KERNEL_PLAIN)
    r.set_accel(accel)
    pwall, pflops, pms = 0.0, 0, 0.0
    for s in ([] if args.no_plain else scenes):
        r.on_resize(s.camera.width, s.camera.height)
        r.set_settings(s.settings)
        r.set_scene(s)
        r.reset_frame_index()
        lib.atx_device_synchronize()
        t0 = time.perf_counter()
        r.render_frames(s.camera, s.spp)
        lib.atx_device_synchronize()
        pwall += time.perf_counter() - t0
        st = r.stats()
        pflops += algo_flops(st)
        pms += st["gpu_ms"]
    r.close()
    bv, pv = samples / bwall, (samples / pwall if pwall > 0 else float("nan"))
    return {
        "metric": "path samples/s, batched vs plain on the mixed workload",
        "value": bv, "unit": "path samples/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": bwall * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded splitmix64 scene generator)",
        "config": {"workload": f"batch_mix: {len(scenes)} scenes, 256..1920 x 256..1080, 16..4096 spheres, "
                               f"1..16 lights, 2..10 bounces, spp x{args.spp_scale}", "samples": samples,
                   "accel": "flat" if accel == ATX_ACCEL_NONE else "bvh", "tuning": tuning, "autotune_s": tune_s},
        "batched": {"samples_per_s": bv, "wall_s": bwall, "gpu_ms": bst["gpu_ms"],
                    "tflops": algo_flops(bst) / (bst["gpu_ms"] * 1e-3) / 1e12},
        "plain": {"samples_per_s": pv, "wall_s": pwall, "gpu_ms": pms,
                  "tflops": pflops / (pms * 1e-3) / 1e12 if pms > 0 else float("nan")},
        "speedup": bv / pv,
    }


def bench_serve(args, rate=None, requests=None):
    """Poisson request stream through the continuous-batching server."""
    import math
    import numpy as np
    from ataraxia_amd import Server, SplitMix64, serve_request_scene
    rate = rate or args.rate
    requests = requests or args.requests
    rng = SplitMix64(4)
    scenes = [serve_request_scene(rng, i) for i in range(requests)]
    arr = SplitMix64(104)
    t, arrivals = 0.0, []
    for _ in scenes:
        t += -math.log(1.0 - arr.uniform()) / rate
        arrivals.append(t)
    from ataraxia_amd import ATX_ACCEL_BVH, ATX_ACCEL_NONE
    srv = Server(frames_per_tick=args.frames_per_tick, tick_mpaths=args.tick_mpaths,
                 accel=ATX_ACCEL_NONE if args.accel == "none" else ATX_ACCEL_BVH)
    t0 = time.perf_counter()
    ids = []
    for when, sc in zip(arrivals, scenes):
        delay = t0 + when - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        ids.append(srv.submit(sc))
    for rid in ids:
        srv.wait(rid, 600.0)
    wall = time.perf_counter() - t0
    infos = [srv.info(rid) for rid in ids]
    st = srv.stats()
    srv.close()
    lat = np.array([i["t_done"] - i["t_submit"] for i in infos])
    span = max(i["t_done"] for i in infos) - min(i["t_submit"] for i in infos)
    samples = sum(s.camera.width * s.camera.height * s.spp for s in scenes)
    return {"rate_per_s": rate, "requests": requests, "samples": samples, "samples_per_s": samples / span,
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(lat.mean()), "wall_s": wall, "hbm_high_water_bytes": st["hbm_high_water"],
            "device_used_high_water_bytes": st["device_used_high_water"], "ticks": st["ticks"],
            "gpu_busy_ms": st["gpu_busy_ms"], "worker_admit_ms": st["admit_ms"], "worker_retire_ms": st["retire_ms"],
            "worker_wait_ms": st["wait_ms"], "worker_release_ms": st["release_ms"], "gap_ms": st["gap_ms"],
            "frames_per_tick": args.frames_per_tick or "default",
            "tick_mpaths": args.tick_mpaths or "default",
            "requests_spec": "resolution from {512x512, 1280x720, 1920x1080}, spheres logU[4,1024] + ground, "
                             "spp logU[16,1024], 8 bounces, 1..8 lights"}


def bench_precision(args):
    """Naive vs Kahan FP32 accumulation against a float64 host sum of the
    identical per-sample radiance stream (captured from the GPU every frame)."""
    import numpy as np
    from ataraxia_amd import ATX_ACCUM_KAHAN, Renderer, precision_scene
    from ataraxia_amd.scene import scene_from_camera_overrides
    side, spp = args.precision_size, args.precision_spp
    sc = scene_from_camera_overrides(precision_scene(), side, side)
    cam = sc.camera
    n = side * side
    s64 = np.zeros((n, 4), np.float64)
    s32 = np.zeros((n, 4), np.float32)
    r = Renderer()
    r.on_resize(side, side)
    r.set_settings(sc.settings)
    r.set_scene(sc)
    r.set_accumulator(ATX_ACCUM_KAHAN)
    t0 = time.perf_counter()
    for f in range(spp):
        rad, _, _ = r.render_debug(cam, 1)
        s64 += rad
        s32 = s32 + rad
        if (f + 1) % 512 == 0:
            print(f"precision: {f + 1}/{spp} frames, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    kah = r.get_accumulation().reshape(n, 4)
    r.close()

    def err(x):
        m = s64[:, :3] > 0
        e = np.abs(x[:, :3].astype(np.float64)[m] - s64[:, :3][m]) / np.maximum(s64[:, :3][m], 1e-6)
        return {"max": float(e.max()), "mean": float(e.mean())}
    ek, en = err(kah), err(s32)
    return {"metric": "accumulator relative error vs float64 at 4096 spp", "value": ek["max"],
            "unit": "relative error (max over pixels and channels)", "n_gpus": 1, "steps": spp, "warmup": 0,
            "ms_per_step": wall * 1e3 / spp, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic (seeded splitmix64 scene generator)",
            "config": {"workload": f"precision: {side}x{side} of the 1024x1024 scene, 512 spheres + ground, 4 lights, "
                                   f"6 bounces, {spp} spp"},
            "kahan": ek, "naive": en, "kahan_below_1e-6": ek["max"] < 1e-6}


def bench_copy(args):
    """Device-to-device copies of --copy-bytes bytes through the HIP runtime:
    with rocprofv3 --pmc FETCH_SIZE (or WRITE_SIZE) each copy's reading is
    the counter's value for that many bytes read (or written)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    n = args.copy_bytes
    src, dst = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(src), n) != 0 or hip.hipMalloc(C.byref(dst), n) != 0:
        raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(src, 1, n) != 0 or hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipMemset failed")
        copies = max(1, args.steps)
        t0 = time.perf_counter()
        for _ in range(copies):
            if hip.hipMemcpy(dst, src, n, 3) != 0:   # hipMemcpyDeviceToDevice
                raise RuntimeError("hipMemcpy failed")
        hip.hipDeviceSynchronize()
        wall = time.perf_counter() - t0
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)
    return {"metric": "device-to-device copy (counter calibration)", "value": 2 * n * copies / wall / 1e9,
            "unit": "GB/s (read + write)", "n_gpus": 1, "steps": copies, "warmup": 0,
            "ms_per_step": wall * 1e3 / copies, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "u8", "data": "synthetic", "config": {"workload": f"copy: {n} bytes per copy"}}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.exact_loop:
        os.environ["ATX_EXACT_LOOP"] = "1"
    if args.workload == "plain_1080p":
        out = bench_plain_1080p(args)
    elif args.workload == "batch_mix":
        out = bench_batch_mix(args)
    elif args.workload == "precision":
        out = bench_precision(args)
    elif args.workload == "copy":
        out = bench_copy(args)
    else:
        sv = bench_serve(args)
        out = {"metric": "path samples/s + p50/p99 request latency in serve loop", "value": sv["samples_per_s"],
               "unit": "path samples/s", "n_gpus": 1, "steps": sv["requests"], "warmup": 0,
               "ms_per_step": sv["wall_s"] * 1e3 / sv["requests"], "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded splitmix64 request generator)",
               "config": {"workload": f"serve: Poisson {sv['rate_per_s']}/s, {sv['requests']} requests"},
               "serve": sv}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
