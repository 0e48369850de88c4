"""Precision of the accumulators at 4096 spp (the precision workload's scene,
1024x1024 reduced to 40x40 pixels; 512 spheres + ground, 4 lights, 6 bounces).

The same FP32 per-sample radiance stream (captured from the GPU every frame)
is summed three ways: by the GPU's naive accumulator (must equal the host's
float32 running sum bit for bit), by the GPU's Kahan accumulator, and in
float64 on the host (the reference).  The Kahan sum must stay within 1e-6
relative of the float64 sum for every pixel and channel."""
import numpy as np
import pytest

from ataraxia_amd import ATX_ACCUM_KAHAN, Renderer, precision_scene
from ataraxia_amd.scene import scene_from_camera_overrides

pytestmark = pytest.mark.gpu

SPP = 4096


def rel_err(x, ref):
    m = ref > 0
    return np.abs(x[m].astype(np.float64) - ref[m]) / np.maximum(np.abs(ref[m]), 1e-6)


def test_kahan_within_1e6_of_float64_at_4096_spp():
    sc = scene_from_camera_overrides(precision_scene(), 40, 40)
    cam = sc.camera
    n = cam.width * cam.height
    print("case precision 40x40 spp", SPP, flush=True)
    s64 = np.zeros((n, 4), np.float64)
    s32 = np.zeros((n, 4), np.float32)
    with Renderer() as naive, Renderer() as kahan:
        for r in (naive, kahan):
            r.on_resize(cam.width, cam.height)
            r.set_settings(sc.settings)
            r.set_scene(sc)
        kahan.set_accumulator(ATX_ACCUM_KAHAN)
        for f in range(SPP):
            rad, _, _ = kahan.render_debug(cam, 1)
            rad2, _, _ = naive.render_debug(cam, 1)
            assert np.array_equal(rad, rad2)
            s64 += rad
            s32 = s32 + rad
        acc_naive = naive.get_accumulation().reshape(n, 4)
        acc_kahan = kahan.get_accumulation().reshape(n, 4)
    assert np.array_equal(acc_naive, s32)           # GPU naive == float32 running sum
    assert np.all(acc_kahan[:, 3] == SPP)
    ek = rel_err(acc_kahan[:, :3], s64[:, :3])
    en = rel_err(acc_naive[:, :3], s64[:, :3])
    print(f"kahan max {ek.max():.3g} mean {ek.mean():.3g} | naive max {en.max():.3g} mean {en.mean():.3g}")
    assert ek.max() < 1e-6
    assert ek.mean() < en.mean()


def test_production_path_kahan_against_float64_oracle():
    """The production path (persistent kernel, first-bounce cache, 4096
    frames in 256-frame launches) at 64x64 of the precision scene: the naive
    accumulation equals the oracle's float32 frame-by-frame sum bit for bit,
    and the Kahan accumulation stays within 1e-6 relative of the oracle's
    float64 sum of the identical float32 samples."""
    import oracle as O
    sc = scene_from_camera_overrides(precision_scene(), 64, 64)
    cam = sc.camera
    n = cam.width * cam.height
    print("case precision production 64x64 spp", SPP, flush=True)
    accs = {}
    for kind in ("naive", "kahan"):
        with Renderer() as r:
            r.on_resize(cam.width, cam.height)
            r.set_settings(sc.settings)
            r.set_scene(sc)
            if kind == "kahan":
                r.set_accumulator(ATX_ACCUM_KAHAN)
            r.render_frames(cam, SPP)
            assert r.stats()["launches"] == SPP // 256
            accs[kind] = r.get_accumulation().reshape(n, 4).copy()
    st = sc.settings
    ocam, osc = O.camera_from(cam), O.Scene(sc.spheres, sc.materials, sc.lights)
    oset = O.settings(st.max_bounces, st.sky_light, st.accumulation)
    s64 = O.accumulate_pixels_f64(ocam, osc, oset, 1, SPP, np.arange(n), threads=16)
    s32, _ = O.accumulate_pixels(ocam, osc, oset, 1, SPP, np.arange(n), threads=16)
    assert np.array_equal(accs["naive"], s32)
    ek = rel_err(accs["kahan"][:, :3], s64[:, :3])
    en = rel_err(accs["naive"][:, :3], s64[:, :3])
    print(f"kahan max {ek.max():.3g} mean {ek.mean():.3g} | naive max {en.max():.3g} mean {en.mean():.3g}")
    assert ek.max() < 1e-6
    assert en.max() > ek.max()
