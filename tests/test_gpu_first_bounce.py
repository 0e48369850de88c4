"""First-bounce cache of the persistent kernel.

A pixel's primary ray has no jitter, so it is the same ray in every frame,
and so is the shadow ray from its hit towards each light.  The persistent
kernel takes bounce 0 of every path from a per-pixel cache (closest primary
hit + one shadow-decision bit per light, scenes of at most 32 lights) that
k_first_hit computes once per camera and scene, with the list of pixels
whose primary ray hits: the launch's jobs cover only those, and the others'
sample (sky colour or black) is added by the accumulate kernel (the renderer builds it for a
call of two or more frames, or once the camera has not moved since the
previous call; batches and the server for scenes of two or more frames).
Every case here renders
through that cache (not the debug capture, which traces every ray) and
compares with the plain kernel, which traces every ray as the reference's
loop does, and with the oracle: accumulation and image bit for bit, the same
ray count, and exactly the expected number of rays answered from the cache.
"""
import numpy as np
import pytest

import oracle as O
from ataraxia_amd import ATX_KERNEL_PLAIN, ATX_KERNEL_WAVEFRONT, Camera, Renderer, Settings, plain_cpu_scene
from ataraxia_amd.scene import (FlatScene, SplitMix64, make_lights, make_materials, make_spheres, random_scene,
                                scene_from_camera_overrides)
from _scenes import oracle_accumulate

pytestmark = pytest.mark.gpu


def _lights_scene(n_lights, seed, bounces=4):
    rng = SplitMix64(seed)
    sph, mat, lig = random_scene(rng, 30, n_lights, 8.0, (0.0, 3.0), (6.0, 12.0))
    return FlatScene(sph, mat, lig, Settings(True, False, bounces), Camera.looking_at((0, 5, 18), width=48, height=32),
                     name=f"{n_lights}-lights")


def _cases():
    base = scene_from_camera_overrides(plain_cpu_scene(), 64, 48)
    emis = base.materials.copy()
    emis["emission_color"][1] = (1.0, 0.5, 0.25)
    emis["emission_intensity"][1] = 3.0
    return {
        "plain-cpu": base,
        "32-lights": _lights_scene(32, 5),          # every light has a bit
        "33-lights": _lights_scene(33, 6),          # too many: shadow rays of bounce 0 traced
        "no-lights": FlatScene(base.spheres, base.materials, make_lights(0), Settings(True, False, 5), base.camera,
                               name="no-lights"),
        "sky-one-bounce": FlatScene(base.spheres, base.materials, base.lights, Settings(True, True, 1), base.camera,
                                    name="sky-one-bounce"),
        "zero-bounces": FlatScene(base.spheres, base.materials, base.lights, Settings(True, False, 0), base.camera,
                                  name="zero-bounces"),
        "emissive": FlatScene(base.spheres, emis, base.lights, Settings(True, True, 6), base.camera, name="emissive"),
        # no sphere: every primary ray misses, no path job at all, k_accumulate alone
        "all-sky": FlatScene(make_spheres(0), make_materials(0), base.lights, Settings(True, True, 5), base.camera,
                             name="all-sky"),
        # looking up: most pixels miss; sky light off, so their samples are black
        "mostly-miss": FlatScene(base.spheres, base.materials, base.lights, Settings(True, False, 5),
                                 Camera.looking_at((0.0, 1.0, 6.0), target=(0.0, 9.0, -4.0), width=64, height=48),
                                 name="mostly-miss"),
    }


CASES = _cases()


def _render(sc, kernel, frames, per_call):
    cam = sc.camera
    with Renderer() as r:
        r.set_kernel(kernel)
        r.on_resize(cam.width, cam.height)
        r.set_settings(sc.settings)
        r.set_scene(sc)
        rays = reused = 0
        for _ in range(frames // per_call):
            r.render_frames(cam, per_call)
            st = r.stats()
            rays += st["rays"]
            reused += st["reused_rays"]
        return r.get_accumulation().copy(), r.get_image().copy(), rays, reused


def _expected_reused(sc, frames):
    """Paths that take bounce 0 from the cache reuse their primary ray, and
    their shadow ray when the primary ray hits and the scene has 1..32 lights."""
    if sc.settings.max_bounces <= 0:
        return 0
    cam = O.camera_from(sc.camera)
    osc = O.Scene(sc.spheres, sc.materials, sc.lights)
    st = O.settings(sc.settings.max_bounces, sc.settings.sky_light, sc.settings.accumulation)
    _, hit, _, _ = O.render_frame_debug(cam, osc, st, 1, max(1, sc.settings.max_bounces))
    n = sc.camera.width * sc.camera.height
    shadow = 1 <= len(sc.lights) <= 32
    return frames * (n + (int((hit[:, 0] >= 0).sum()) if shadow else 0))


@pytest.mark.parametrize("name", list(CASES))
def test_cached_bounce0_equals_plain_and_oracle(name):
    sc = CASES[name]
    frames = 4
    print(f"case {name} {sc.camera.width}x{sc.camera.height} lights={len(sc.lights)} "
          f"bounces={sc.settings.max_bounces} frames={frames}", flush=True)
    acc_p, img_p, rays_p, reused_p = _render(sc, ATX_KERNEL_PLAIN, frames, frames)
    acc_w, img_w, rays_w, reused_w = _render(sc, ATX_KERNEL_WAVEFRONT, frames, frames)
    assert np.array_equal(acc_w, acc_p)
    assert np.array_equal(img_w, img_p)
    assert rays_w == rays_p and reused_p == 0
    assert reused_w == _expected_reused(sc, frames)
    oacc, oimg = oracle_accumulate(sc, frames)
    assert np.array_equal(acc_w.reshape(-1, 4), oacc)
    assert np.array_equal(img_w.reshape(-1), oimg)


def test_one_frame_calls_reuse_the_cache():
    """render_frames(cam, 1) four times (the reference's Render loop) equals
    four frames in one call; the first call of a camera traces every ray,
    the next ones (same camera) use the cache."""
    sc = CASES["plain-cpu"]
    acc1, img1, rays1, reused1 = _render(sc, ATX_KERNEL_WAVEFRONT, 4, 1)
    acc4, img4, rays4, reused4 = _render(sc, ATX_KERNEL_WAVEFRONT, 4, 4)
    assert np.array_equal(acc1, acc4) and np.array_equal(img1, img4)
    assert rays1 == rays4 and reused4 > 0 and reused1 * 4 == reused4 * 3


def test_moving_camera_never_builds_the_cache():
    """One frame per call with a different camera each call: no call reuses
    a ray, and the images equal the plain kernel's."""
    sc = CASES["plain-cpu"]
    cams = [Camera.looking_at((x, 4.0, 14.0), width=sc.camera.width, height=sc.camera.height)
            for x in (0.0, 0.5, 1.0, 0.5)]
    out = {}
    for kernel in (ATX_KERNEL_PLAIN, ATX_KERNEL_WAVEFRONT):
        with Renderer() as r:
            r.set_kernel(kernel)
            r.on_resize(sc.camera.width, sc.camera.height)
            r.set_settings(sc.settings)
            r.set_scene(sc)
            accs = []
            for cam in cams:
                r.reset_frame_index()
                r.render_frames(cam, 1)
                assert r.stats()["reused_rays"] == 0
                accs.append(r.get_accumulation().copy())
        out[kernel] = accs
    for x, y in zip(out[ATX_KERNEL_PLAIN], out[ATX_KERNEL_WAVEFRONT]):
        assert np.array_equal(x, y)


def test_cache_follows_camera_and_scene():
    """Camera B after camera A (same size, no new scene), then a new scene
    under camera B: the cache must be rebuilt each time."""
    sc = CASES["plain-cpu"]
    other = CASES["32-lights"]
    cam_a = sc.camera
    cam_b = Camera.looking_at((3.0, 4.0, 14.0), target=(0.5, 0.5, 0.0), width=cam_a.width, height=cam_a.height)
    out = {}
    for kernel in (ATX_KERNEL_PLAIN, ATX_KERNEL_WAVEFRONT):
        with Renderer() as r:
            r.set_kernel(kernel)
            r.on_resize(cam_a.width, cam_a.height)
            r.set_settings(sc.settings)
            r.set_scene(sc)
            r.render_frames(cam_a, 2)
            a = r.get_accumulation().copy()
            r.reset_frame_index()
            r.render_frames(cam_b, 2)
            b = r.get_accumulation().copy()
            r.set_scene(other)
            r.reset_frame_index()
            r.render_frames(cam_b, 1)
            r.render_frames(cam_b, 1)     # same camera as the previous call: uses a cache of the new scene
            assert (r.stats()["reused_rays"] > 0) == (kernel == ATX_KERNEL_WAVEFRONT)
            c = r.get_accumulation().copy()
        out[kernel] = (a, b, c)
    for x, y in zip(out[ATX_KERNEL_PLAIN], out[ATX_KERNEL_WAVEFRONT]):
        assert np.array_equal(x, y)
    assert not np.array_equal(out[ATX_KERNEL_WAVEFRONT][0], out[ATX_KERNEL_WAVEFRONT][1])


def test_cache_kept_across_accel_and_settings_changes():
    """The renderer keeps its cache while camera and scene stay; switching
    tree/flat loop, sky light and bounce depth between calls must give the
    plain kernel's accumulation after every call."""
    from ataraxia_amd import ATX_ACCEL_BVH, ATX_ACCEL_NONE
    sc = CASES["plain-cpu"]
    cam = sc.camera
    steps = [(ATX_ACCEL_BVH, Settings(True, False, 5), 2), (ATX_ACCEL_NONE, Settings(True, False, 5), 1),
             (ATX_ACCEL_BVH, Settings(True, True, 5), 2), (ATX_ACCEL_BVH, Settings(True, True, 0), 1),
             (ATX_ACCEL_NONE, Settings(True, False, 3), 3)]
    out = {}
    for kernel in (ATX_KERNEL_PLAIN, ATX_KERNEL_WAVEFRONT):
        accs = []
        with Renderer() as r:
            r.set_kernel(kernel)
            r.on_resize(cam.width, cam.height)
            r.set_settings(sc.settings)
            r.set_scene(sc)
            for accel, st, frames in steps:
                r.set_accel(accel)
                r.set_settings(st)
                r.render_frames(cam, frames)
                accs.append(r.get_accumulation().copy())
        out[kernel] = accs
    for k, (x, y) in enumerate(zip(out[ATX_KERNEL_PLAIN], out[ATX_KERNEL_WAVEFRONT])):
        assert np.array_equal(x, y), k


def test_reserve_keeps_results_and_redoes_dropped_caches():
    """Renderer.reserve: before the first call, between calls of one camera
    (the sample buffer grows) and after a resize, the accumulation equals a
    renderer that never reserved; growing the buffers drops the per-camera
    caches, which the next call rebuilds."""
    sc = CASES["plain-cpu"]
    cam = sc.camera
    ref, _, _, _ = _render(sc, ATX_KERNEL_WAVEFRONT, 6, 2)
    with Renderer() as r:
        r.on_resize(cam.width, cam.height)
        r.set_settings(sc.settings)
        r.reserve(4)                 # no scene yet: sample buffer and primary directions only
        r.set_scene(sc)
        r.render_frames(cam, 2)
        r.reserve(64)                # more frames per call than reserved before
        r.render_frames(cam, 2)
        r.render_frames(cam, 2)
        assert np.array_equal(r.get_accumulation(), ref)
        small = scene_from_camera_overrides(sc, cam.width // 2, cam.height // 2)
        r.on_resize(small.camera.width, small.camera.height)
        r.reserve(2)
        r.reset_frame_index()
        r.render_frames(small.camera, 2)
        acc_small = r.get_accumulation().copy()
    ref_small, _, _, _ = _render(small, ATX_KERNEL_WAVEFRONT, 2, 2)
    assert np.array_equal(acc_small, ref_small)
