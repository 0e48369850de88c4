"""GPU parity of the plain single-scene kernel against the CPU oracle.

Every case renders frames through the C ABI (libatx.so, HIP kernels on the
device) and through the oracle on the same inputs and seeds, then checks
* closest-hit sphere indices of every continuation and shadow ray: exact;
* per-sample radiance: relative error <= 1e-4 per channel (the restated
  arithmetic is identical, so this is expected to be 0);
* accumulation buffer and RGBA8 image after several frames: radiance bar on
  the accumulation, exact RGBA8 where the accumulations are bit-equal.
"""
import numpy as np
import pytest

import oracle as O
from ataraxia_amd import ATX_ACCUM_KAHAN, ATX_KERNEL_PLAIN, ATX_KERNEL_WAVEFRONT, Camera, Renderer, Settings
from ataraxia_amd import default_scene, plain_1080p_scene
from ataraxia_amd import plain_cpu_scene
from ataraxia_amd.scene import FlatScene, make_lights, make_materials, make_spheres, scene_from_camera_overrides

pytestmark = pytest.mark.gpu

KERNELS = pytest.mark.parametrize("kernel", [ATX_KERNEL_PLAIN, ATX_KERNEL_WAVEFRONT], ids=["plain", "wavefront"])

RAD_TOL = 1e-4


def _oracle_scene(sc: FlatScene) -> O.Scene:
    sph = sc.spheres.copy()
    # the oracle indexes materials with the raw id; clamp like the host does
    bad = (sph["material"] < 0) | (sph["material"] >= len(sc.materials))
    sph["material"][bad] = 0
    return O.Scene(sph, sc.materials, sc.lights)


def _rel_err(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-6)


def run_parity(sc: FlatScene, frames: int, kahan: bool = False, kernel: int = ATX_KERNEL_PLAIN):
    cam = sc.camera
    st = sc.settings
    nb = max(1, st.max_bounces)
    print(f"case {sc.name} {cam.width}x{cam.height} spheres={sc.num_spheres} lights={len(sc.lights)} "
          f"bounces={st.max_bounces} frames={frames} kahan={kahan} kernel={kernel}", flush=True)
    osc = _oracle_scene(sc)
    ocam = O.camera_from(cam)
    oset = O.settings(st.max_bounces, st.sky_light, st.accumulation)
    n = cam.width * cam.height
    o_acc = np.zeros((n, 4), np.float32)
    o_comp = np.zeros((n, 4), np.float32) if kahan else None
    with Renderer() as r:
        r.set_kernel(kernel)
        r.on_resize(cam.width, cam.height)
        r.set_settings(st)
        if kahan:
            r.set_accumulator(ATX_ACCUM_KAHAN)
        r.set_scene(sc)
        total_rays = 0
        for f in range(1, frames + 1):
            assert r.frame_index == f
            rad, hit, sh = r.render_debug(cam, nb)
            orad, ohit, osh, orays = O.render_frame_debug(ocam, osc, oset, f, nb)
            o_acc, o_comp, o_img, _, _ = O.render_frame(ocam, osc, oset, f, accum=o_acc, comp=o_comp,
                                                        want_radiance=False)
            mism = np.argwhere(hit != ohit)
            assert mism.size == 0, f"frame {f}: {len(mism)} continuation hit mismatches, first {mism[:5]}"
            mism = np.argwhere(sh != osh)
            assert mism.size == 0, f"frame {f}: {len(mism)} shadow hit mismatches, first {mism[:5]}"
            err = _rel_err(rad, orad).max() if n else 0.0
            assert err <= RAD_TOL, f"frame {f}: radiance rel err {err}"
            st_ = r.stats()
            assert st_["rays"] == int(orays.sum()), (st_["rays"], int(orays.sum()))
            total_rays += st_["rays"]
        acc = r.get_accumulation().reshape(n, 4)
        img = r.get_image().reshape(n)
    assert _rel_err(acc, o_acc).max() <= RAD_TOL
    same = np.all(acc == o_acc, axis=1)
    assert np.array_equal(img[same], o_img[same])
    return acc, o_acc, total_rays


@KERNELS
def test_default_scene_small(kernel):
    sc = scene_from_camera_overrides(default_scene(), 64, 36)
    sc.settings = Settings(True, False, 6)
    acc, oacc, _ = run_parity(sc, 4, kernel=kernel)
    assert np.array_equal(acc, oacc)


@KERNELS
def test_default_scene_sky_light(kernel):
    sc = scene_from_camera_overrides(default_scene(), 48, 40)
    sc.settings = Settings(True, True, 4)
    run_parity(sc, 3, kernel=kernel)


@KERNELS
def test_plain_cpu_config_reduced(kernel):
    sc = scene_from_camera_overrides(plain_cpu_scene(), 160, 90)
    acc, oacc, _ = run_parity(sc, 4, kernel=kernel)
    assert np.array_equal(acc, oacc)


@KERNELS
def test_plain_1080p_scene_reduced(kernel):
    sc = scene_from_camera_overrides(plain_1080p_scene(), 192, 108)
    run_parity(sc, 3, kernel=kernel)


@KERNELS
def test_kahan_accumulator(kernel):
    sc = scene_from_camera_overrides(plain_cpu_scene(), 96, 54)
    acc, oacc, _ = run_parity(sc, 5, kahan=True, kernel=kernel)
    assert np.array_equal(acc, oacc)


@KERNELS
def test_odd_sizes_and_tiny_images(kernel):
    for w, h in [(1, 1), (17, 5), (33, 31)]:
        sc = scene_from_camera_overrides(plain_cpu_scene(), w, h)
        run_parity(sc, 2, kernel=kernel)


@KERNELS
def test_no_lights_zero_bounces_empty_scene(kernel):
    base = scene_from_camera_overrides(plain_cpu_scene(), 40, 24)
    nl = FlatScene(base.spheres, base.materials, make_lights(0), Settings(True, False, 5), base.camera, name="nolights")
    run_parity(nl, 2, kernel=kernel)
    zb = FlatScene(base.spheres, base.materials, base.lights, Settings(True, False, 0), base.camera, name="zero")
    acc, _, _ = run_parity(zb, 2, kernel=kernel)
    assert np.all(acc[:, :3] == 0)
    empty = FlatScene(make_spheres(0), make_materials(0), base.lights, Settings(True, True, 3), base.camera,
                      name="empty-sky")
    acc, _, _ = run_parity(empty, 2, kernel=kernel)
    # a sky-only path returns exactly (0.6, 0.7, 0.9) every frame
    assert np.allclose(acc[:, :3] / 2.0, np.array([0.6, 0.7, 0.9], np.float32), rtol=0, atol=0)


@KERNELS
def test_emissive_metallic_and_clamped_ids(kernel):
    sc = scene_from_camera_overrides(plain_cpu_scene(), 64, 48)
    mat = sc.materials.copy()
    mat["emission_color"][1] = (1.0, 0.5, 0.25)
    mat["emission_intensity"][1] = 3.0
    mat["metallic"][2] = 1.0
    sph = sc.spheres.copy()
    sph["material"][3] = 99          # out of range: clamped to 0 on upload
    s2 = FlatScene(sph, mat, sc.lights, Settings(True, False, 5), sc.camera, name="emissive")
    run_parity(s2, 3, kernel=kernel)


@KERNELS
def test_large_scene_global_memory_path(kernel):
    """More spheres than the LDS budget (64 KiB = 4096 spheres): global loads."""
    from ataraxia_amd.scene import SplitMix64, random_scene
    rng = SplitMix64(77)
    sph, mat, lig = random_scene(rng, 4200, 3, 40.0, (0.0, 6.0), (10.0, 20.0))
    cam = Camera.looking_at((0.0, 16.0, 80.0), width=40, height=24)
    sc = FlatScene(sph, mat, lig, Settings(True, False, 4), cam, name="n4201")
    run_parity(sc, 2, kernel=kernel)


@KERNELS
def test_accumulation_off_resets_every_frame(kernel):
    sc = scene_from_camera_overrides(plain_cpu_scene(), 32, 18)
    sc.settings = Settings(False, False, 5)
    with Renderer() as r:
        r.set_kernel(kernel)
        r.on_resize(32, 18)
        r.set_settings(sc.settings)
        r.set_scene(sc)
        for _ in range(3):
            rad, _, _ = r.render_debug(sc.camera, 5)
            assert r.frame_index == 1
        acc = r.get_accumulation().reshape(-1, 4)
    # with accumulation off the frame index stays 1: accum == that frame's sample
    assert np.array_equal(acc, rad)


@KERNELS
def test_full_1080p_sampled_pixels_and_determinism(kernel):
    """Full-size workload (1920x1080, 257 spheres, 8 lights, 8 bounces): the
    oracle checks 4096 sampled pixels per frame; the whole image must be
    bit-identical across two renderers (determinism)."""
    sc = plain_1080p_scene()
    cam = sc.camera
    st = sc.settings
    nb = st.max_bounces
    osc = _oracle_scene(sc)
    ocam = O.camera_from(cam)
    oset = O.settings(st.max_bounces, st.sky_light, st.accumulation)
    dirs = O.ray_directions(ocam).reshape(-1, 3)
    rng = np.random.default_rng(5)
    pix = rng.choice(cam.width * cam.height, size=4096, replace=False)
    outs = []
    print(f"case plain_1080p kernel={kernel}", flush=True)
    for rep in range(2):
        with Renderer() as r:
            r.set_kernel(kernel)
            r.on_resize(cam.width, cam.height)
            r.set_settings(st)
            r.set_scene(sc)
            for f in (1, 2):
                rad, hit, sh = r.render_debug(cam, nb)
                if rep == 0:
                    for p in pix:
                        x, y = int(p % cam.width), int(p // cam.width)
                        orad, ohit, osh, _ = O.per_pixel(x, y, cam.width, dirs[p], cam.position, osc, f, oset)
                        assert np.array_equal(hit[p], ohit), (x, y, f)
                        assert np.array_equal(sh[p], osh), (x, y, f)
                        assert _rel_err(rad[p], orad).max() <= RAD_TOL
            outs.append(r.get_accumulation().copy())
    assert np.array_equal(outs[0], outs[1])


@KERNELS
def test_plain_cpu_config_full_size_all_frames(kernel):
    """The reference's own CPU-runnable configuration at full size: 640x360,
    4 spheres + ground, 1 light, 5 bounces, 64 frames accumulated.  Every
    pixel's accumulation and RGBA8 value after the 64 frames equal the
    oracle's frame-by-frame sums bit for bit (14.7 M paths)."""
    sc = plain_cpu_scene()
    cam = sc.camera
    frames = sc.spp
    print(f"case plain_cpu full 640x360x{frames} kernel={kernel}", flush=True)
    with Renderer() as r:
        r.set_kernel(kernel)
        r.on_resize(cam.width, cam.height)
        r.set_settings(sc.settings)
        r.set_scene(sc)
        r.render_frames(cam, frames)
        acc = r.get_accumulation().reshape(-1, 4)
        img = r.get_image().reshape(-1)
    st = sc.settings
    oacc, oimg = O.accumulate_pixels(O.camera_from(cam), _oracle_scene(sc),
                                     O.settings(st.max_bounces, st.sky_light, st.accumulation), 1, frames,
                                     np.arange(cam.width * cam.height), threads=16)
    assert np.array_equal(acc, oacc)
    assert np.array_equal(img, oimg)
