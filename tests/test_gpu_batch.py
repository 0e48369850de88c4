"""GPU parity of the persistent wavefront kernel in its batched form.

* every scene of a mixed batch (different sizes, sphere counts, lights,
  bounce limits and spp) equals the oracle's frame-by-frame accumulation;
* the last frame's closest-hit indices equal the oracle's exactly;
* results do not depend on the tuning (chunk size, grid size, graph or not),
  on graph replays, or on which other scenes share the batch;
* on the full 1080p workload the wavefront kernel equals the plain kernel.
"""
import numpy as np
import pytest

import oracle as O
from ataraxia_amd import (ATX_ACCUM_KAHAN, ATX_KERNEL_PLAIN, ATX_KERNEL_WAVEFRONT, Batch, Camera, Renderer,
                          Settings, default_scene, plain_1080p_scene, plain_cpu_scene)
from ataraxia_amd.scene import (FlatScene, SplitMix64, batch_mix_scenes, make_lights, make_materials, make_spheres,
                                random_scene, scene_from_camera_overrides)

from _scenes import mixed_scenes, oracle_accumulate

pytestmark = pytest.mark.gpu


def render_batch(scenes, tuning=None, kahan=False, capture=()):
    with Batch() as b:
        ids = [b.add_scene(s) for s in scenes]
        if kahan:
            b.set_accumulator(ATX_ACCUM_KAHAN)
        if tuning:
            b.set_tuning(**tuning)
        for i in capture:
            b.capture(ids[i], max(1, scenes[i].settings.max_bounces))
        st = b.render()
        accs = [b.get_accumulation(i).reshape(-1, 4).copy() for i in ids]
        imgs = [b.get_image(i).reshape(-1).copy() for i in ids]
        caps = {i: b.get_capture(ids[i], max(1, scenes[i].settings.max_bounces)) for i in capture}
    return accs, imgs, caps, st


def test_batch_matches_oracle_per_scene():
    scenes = mixed_scenes()
    print("case batch mixed", [s.name for s in scenes], flush=True)
    accs, imgs, caps, st = render_batch(scenes, capture=(0, 2))
    assert st["samples"] == sum(s.camera.width * s.camera.height * s.spp for s in scenes)
    for i, s in enumerate(scenes):
        oacc, oimg = oracle_accumulate(s, s.spp)
        assert np.array_equal(accs[i], oacc), s.name
        assert np.array_equal(imgs[i], oimg), s.name
    for i in (0, 2):
        s = scenes[i]
        nb = max(1, s.settings.max_bounces)
        rad, hit, sh = caps[i]
        orad, ohit, osh, _ = O.render_frame_debug(O.camera_from(s.camera), O.Scene(s.spheres, s.materials, s.lights),
                                                  O.settings(s.settings.max_bounces, s.settings.sky_light), s.spp, nb)
        assert np.array_equal(hit, ohit) and np.array_equal(sh, osh)
        assert np.array_equal(rad, orad)


def test_batch_kahan_matches_oracle():
    # with the sky-lit multi-frame scenes, whose missing pixels take the
    # compensated sum of the constant sky colour
    scenes = [s for i, s in enumerate(mixed_scenes()) if i in (0, 1, 2, 3, 6)]
    assert sum(s.settings.sky_light and s.spp > 1 for s in scenes) == 2
    accs, _, _, _ = render_batch(scenes, kahan=True)
    for i, s in enumerate(scenes):
        oacc, _ = oracle_accumulate(s, s.spp, kahan=True)
        assert np.array_equal(accs[i], oacc), s.name


@pytest.mark.parametrize("tuning", [
    {"chunk_px": 64, "workgroups": 2, "use_graph": True, "frames_per_launch": 1},
    {"chunk_px": 64, "workgroups": 1, "use_graph": False, "frames_per_launch": 2},
    {"chunk_px": 256, "workgroups": 0, "use_graph": True, "frames_per_launch": 64},
    {"chunk_px": 192, "workgroups": 37, "use_graph": False, "frames_per_launch": 3},
    {"chunk_px": 256, "workgroups": 0, "use_graph": True, "frames_per_launch": 4, "waves_per_simd": 5},
    {"chunk_px": 128, "workgroups": 0, "use_graph": True, "frames_per_launch": 2, "waves_per_simd": 7},
    {"chunk_px": 512, "workgroups": 0, "use_graph": False, "frames_per_launch": 8, "waves_per_simd": 8},
], ids=["c64-g2-f1", "c64-g1-f2-eager", "c256-full-f64", "c192-g37-f3-eager", "w5", "w7", "w8-eager"])
def test_batch_independent_of_tuning(tuning):
    scenes = mixed_scenes()
    print("case tuning", tuning, flush=True)
    ref, _, _, _ = render_batch(scenes)
    got, _, _, _ = render_batch(scenes, tuning=tuning)
    for a, b_, s in zip(ref, got, scenes):
        assert np.array_equal(a, b_), s.name


def test_graph_replays_and_alone_vs_mix():
    scenes = mixed_scenes()
    with Batch() as b:
        ids = [b.add_scene(s) for s in scenes]
        b.render()
        first = [b.get_accumulation(i).copy() for i in ids]
        for _ in range(3):  # replays restart accumulation at frame 1 -> same bytes
            b.render()
            for i in ids:
                assert np.array_equal(b.get_accumulation(i), first[i])
    # the same scene alone in a batch
    for k in (1, 2):
        alone, _, _, _ = render_batch([scenes[k]])
        assert np.array_equal(alone[0], first[k].reshape(-1, 4))


def test_wavefront_equals_plain_on_full_1080p():
    sc = plain_1080p_scene()
    cam = sc.camera
    out = {}
    for kernel in (ATX_KERNEL_PLAIN, ATX_KERNEL_WAVEFRONT):
        with Renderer() as r:
            r.set_kernel(kernel)
            r.on_resize(cam.width, cam.height)
            r.set_settings(sc.settings)
            r.set_scene(sc)
            r.render_frames(cam, 3)
            out[kernel] = (r.get_accumulation().copy(), r.get_image().copy(), r.stats()["rays"])
    assert np.array_equal(out[ATX_KERNEL_PLAIN][0], out[ATX_KERNEL_WAVEFRONT][0])
    assert np.array_equal(out[ATX_KERNEL_PLAIN][1], out[ATX_KERNEL_WAVEFRONT][1])
    assert out[ATX_KERNEL_PLAIN][2] == out[ATX_KERNEL_WAVEFRONT][2]


@pytest.mark.parametrize("size,frames,launches", [((48, 30), 600, 3), ((4096, 4096), 90, 3)],
                         ids=["256-frame-launches", "8GiB-sample-cap"])
def test_wavefront_launch_splits_equal_plain(size, frames, launches):
    """600 frames split into 256 + 256 + 88 frames per launch, and a
    4096x4096 image whose 8 GiB sample buffer holds 42 frames of 12-byte
    samples (42 + 42 + 6): both must match the plain kernel's frame-by-frame
    accumulation bit for bit."""
    w, h = size
    sc = scene_from_camera_overrides(plain_cpu_scene(), w, h)
    sc.settings = Settings(True, False, 4)
    out = {}
    for kernel in (ATX_KERNEL_PLAIN, ATX_KERNEL_WAVEFRONT):
        with Renderer() as r:
            r.set_kernel(kernel)
            r.set_accumulator(ATX_ACCUM_KAHAN)
            r.on_resize(w, h)
            r.set_settings(sc.settings)
            r.set_scene(sc)
            r.render_frames(sc.camera, frames)
            assert r.frame_index == frames + 1
            out[kernel] = (r.get_accumulation().copy(), r.get_image().copy(), r.stats())
    acc_p, img_p, st_p = out[ATX_KERNEL_PLAIN]
    acc_w, img_w, st_w = out[ATX_KERNEL_WAVEFRONT]
    assert np.array_equal(acc_p, acc_w)
    assert np.array_equal(img_p, img_w)
    assert st_p["launches"] == frames and st_p["path_ms"] == st_p["gpu_ms"]
    assert st_w["launches"] == launches
    assert 0.0 < st_w["path_ms"] <= st_w["gpu_ms"]


def test_wavefront_accumulation_off_equals_plain():
    """Accumulation off: every frame is frame 1 (one frame per launch); after
    five frames both kernels hold that frame's samples, bit for bit."""
    sc = scene_from_camera_overrides(plain_1080p_scene(), 120, 70)
    sc.settings = Settings(False, True, 5)
    out = {}
    for kernel in (ATX_KERNEL_PLAIN, ATX_KERNEL_WAVEFRONT):
        with Renderer() as r:
            r.set_kernel(kernel)
            r.on_resize(120, 70)
            r.set_settings(sc.settings)
            r.set_scene(sc)
            r.render_frames(sc.camera, 5)
            assert r.frame_index == 1
            out[kernel] = (r.get_accumulation().copy(), r.get_image().copy(), r.stats())
    assert np.array_equal(out[ATX_KERNEL_PLAIN][0], out[ATX_KERNEL_WAVEFRONT][0])
    assert np.array_equal(out[ATX_KERNEL_PLAIN][1], out[ATX_KERNEL_WAVEFRONT][1])
    assert out[ATX_KERNEL_WAVEFRONT][2]["launches"] == 5


def test_mixed_workload_batch_equals_plain_kernel():
    """The 64 scenes of the mixed workload (16..4096 spheres, 1..16 lights,
    2..10 bounces) at 1/12 of their resolution and at most 40 spp (two
    launches for scenes past 32 frames): each batched image equals rendering
    that scene alone with the plain kernel, bit for bit."""
    scenes = []
    for s in batch_mix_scenes():
        r = scene_from_camera_overrides(s, max(8, s.camera.width // 12), max(8, s.camera.height // 12))
        r.spp = min(s.spp, 40)
        scenes.append(r)
    accs, imgs, _, st = render_batch(scenes, tuning=dict(chunk_px=128, frames_per_launch=32))
    assert st["samples"] == sum(s.camera.width * s.camera.height * s.spp for s in scenes)
    assert st["box_tests"] > 0
    for i, s in enumerate(scenes):
        with Renderer() as r:
            r.set_kernel(ATX_KERNEL_PLAIN)
            r.on_resize(s.camera.width, s.camera.height)
            r.set_settings(s.settings)
            r.set_scene(s)
            r.render_frames(s.camera, s.spp)
            acc = r.get_accumulation().reshape(-1, 4)
            img = r.get_image().reshape(-1)
        assert np.array_equal(accs[i], acc), s.name
        assert np.array_equal(imgs[i], img), s.name


def test_batch_argument_errors():
    from ataraxia_amd import AtxError
    with Batch() as b:
        with pytest.raises(AtxError):
            b.render()  # no scenes
        bad = scene_from_camera_overrides(plain_cpu_scene(), 8, 8)
        with pytest.raises(AtxError):
            b.add_scene(bad, spp=0)
        with pytest.raises(AtxError):
            b.set_tuning(chunk_px=100)
        from ataraxia_amd import _lib
        buf = np.zeros(64, np.uint32)
        assert _lib.lib.atx_batch_get_image(b._h, 5, buf.ctypes.data, buf.size) == _lib.ATX_ERR_NOTFOUND


def test_autotune_picks_valid_tuning_and_caches_it(tmp_path):
    """The autotuner returns one of its candidates, records it in the cache
    file keyed by device and workload, reuses it for the same workload, and
    the tuned batch still equals the oracle."""
    scenes = mixed_scenes()
    cache = tmp_path / "tune.txt"
    with Batch() as b:
        ids = [b.add_scene(s) for s in scenes]
        t1 = b.autotune(str(cache))
        assert t1["chunk_px"] in (256, 512) and t1["frames_per_launch"] in (32, 64, 128, 256)
        assert t1["workgroups"] > 0 and t1["waves_per_simd"] in (5, 6, 7, 8)
        b.render()
        accs = [b.get_accumulation(i).reshape(-1, 4).copy() for i in ids]
    lines = cache.read_text().splitlines()
    assert len(lines) == 1 and lines[0].split("\t")[1:] == [str(t1["chunk_px"]), str(t1["workgroups"]),
                                                           str(t1["frames_per_launch"]), str(t1["waves_per_simd"])]
    with Batch() as b:
        for s in scenes:
            b.add_scene(s)
        t2 = b.autotune(str(cache))      # cache hit: same choice, no new line
    assert t2 == t1
    assert len(cache.read_text().splitlines()) == 1
    for i, s in enumerate(scenes):
        oacc, _ = oracle_accumulate(s, s.spp)
        assert np.array_equal(accs[i], oacc), s.name
