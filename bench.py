"""Benchmark: path samples/s of the 1080p workload on one MI355X.

Workload (the metric's configuration): 1920x1080, 256 random spheres + a
ground sphere, 8 point lights, 8 bounces, 1 sample per pixel per frame,
accumulated; synthetic scene from the seeded generator.  One "step" is one
frame (one path sample for every pixel).

Prints one JSON line with samples/s, the VALU roofline of the render kernel
(17 algorithmic FLOP per ray-sphere test, measured with HIP events on the
renderer's stream) and the CPU oracle timed on the host's cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_VALU_PEAK_TFLOPS = 157.3   # MI355X published FP32 vector peak
FLOP_PER_TEST = 17


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=256, help="timed frames")
    ap.add_argument("--warmup", type=int, default=8, help="untimed frames")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="wall time budget of the CPU baseline")
    ap.add_argument("--exact-loop", action="store_true", help="use the exact (slow) closest-hit loop")
    ap.add_argument("--kernel", choices=["plain", "wavefront"], default="plain")
    return ap.parse_args()


def cpu_baseline(sc, target_s):
    """The oracle on the host: whole frames of the same scene, row-parallel over
    min(16, cores) threads, repeated until about `target_s` seconds elapse."""
    import numpy as np
    import oracle as O
    cam = O.camera_from(sc.camera)
    osc = O.Scene(sc.spheres, sc.materials, sc.lights)
    st = O.settings(sc.settings.max_bounces, sc.settings.sky_light, sc.settings.accumulation)
    threads = max(1, min(16, os.cpu_count() or 1))
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
    return {"value": samples / dt, "unit": "path samples/s", "cores": threads, "kind": "port",
            "sample": f"{frames} full frame(s) of the same 1080p workload ({samples} samples, {rays} rays) "
                      f"in {dt:.2f} s on {threads} threads"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.exact_loop:
        os.environ["ATX_EXACT_LOOP"] = "1"
    from ataraxia_amd import ATX_KERNEL_PLAIN, ATX_KERNEL_WAVEFRONT, Renderer, plain_1080p_scene

    sc = plain_1080p_scene()
    cam = sc.camera
    r = Renderer()
    r.set_kernel(ATX_KERNEL_WAVEFRONT if args.kernel == "wavefront" else ATX_KERNEL_PLAIN)
    r.on_resize(cam.width, cam.height)
    r.set_settings(sc.settings)
    r.set_scene(sc)
    if args.warmup:
        r.render_frames(cam, args.warmup)
    r.reset_frame_index()
    # timed region: K frames queued back to back, synchronised on both sides
    from ataraxia_amd._lib import lib
    lib.atx_device_synchronize()
    t0 = time.perf_counter()
    r.render_frames(cam, args.steps)
    lib.atx_device_synchronize()
    wall = time.perf_counter() - t0
    st = r.stats()
    samples = cam.width * cam.height * args.steps
    kernel_ms = st["gpu_ms"] / args.steps
    tests_per_launch = st["sphere_tests"] / args.steps
    achieved = FLOP_PER_TEST * tests_per_launch / (kernel_ms * 1e-3) / 1e12
    out = {
        "metric": "path samples/s (1080p, 8 bounces)",
        "value": samples / wall,
        "unit": "path samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded splitmix64 scene generator)",
        "config": {"workload": "plain_1080p: 1920x1080, 256 spheres + ground, 8 lights, 8 bounces, "
                               "1 spp/frame accumulated", "kernel": args.kernel,
                   "spheres": int(sc.num_spheres), "lights": int(len(sc.lights)), "frames": args.steps},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_VALU_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP32_VALU_PEAK_TFLOPS, "traffic": None,
                     "kernel_ms": kernel_ms, "rays_per_launch": st["rays"] / args.steps,
                     "sphere_tests_per_launch": tests_per_launch,
                     "exact_retraces": st["exact_retraces"]},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(sc, args.cpu_seconds)
    r.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
