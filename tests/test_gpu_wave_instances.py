"""Every compiled instance of the persistent path kernel renders the oracle's
images.

The renderer picks the instance (resident waves per SIMD: 5 and 6 keep the
path state in VGPRs, 7 and 8 park it in LDS during each trace) by the
launch's job count; small test scenes therefore run only the default ones.
Here ATX_WF_WAVES forces each instance in turn, on a tree scene whose
bounce 0 traces its shadow rays (33 lights), a tree scene with the
first-bounce cache's shadow bits, sky light and the compensated sum, and a
scene whose primary rays mostly miss (short hit list).  Each case checks,
against the oracle, a multi-frame call bit for bit (accumulation, RGBA8 and
the count of cached rays) and then the per-frame debug capture (per-bounce
hit and shadow ids, radiance), which the path kernel writes at the pixel it
finds from the job.
"""
import numpy as np
import pytest

import oracle as O
from ataraxia_amd import ATX_ACCUM_KAHAN, ATX_KERNEL_WAVEFRONT, Camera, Renderer, Settings, plain_cpu_scene
from ataraxia_amd.scene import FlatScene, SplitMix64, random_scene, scene_from_camera_overrides
from _scenes import oracle_accumulate
from test_gpu_plain_parity import run_parity

pytestmark = pytest.mark.gpu

WAVES = [5, 6, 7, 8]


def _cases():
    rng = SplitMix64(41)
    sph, mat, lig = random_scene(rng, 40, 33, 8.0, (0.0, 3.0), (6.0, 12.0))
    many = FlatScene(sph, mat, lig, Settings(True, False, 5), Camera.looking_at((0, 5, 18), width=56, height=40),
                     name="tree-33-lights")
    rng = SplitMix64(42)
    sph, mat, lig = random_scene(rng, 40, 6, 8.0, (0.0, 3.0), (6.0, 12.0))
    sky = FlatScene(sph, mat, lig, Settings(True, True, 6), Camera.looking_at((0, 4, 16), width=48, height=40),
                    name="tree-sky-kahan")
    base = scene_from_camera_overrides(plain_cpu_scene(), 64, 48)
    miss = FlatScene(base.spheres, base.materials, base.lights, Settings(True, True, 5),
                     Camera.looking_at((0.0, 1.0, 6.0), target=(0.0, 9.0, -4.0), width=64, height=48),
                     name="mostly-miss")
    return {"tree-33-lights": (many, False), "tree-sky-kahan": (sky, True), "mostly-miss": (miss, False)}


CASES = _cases()


def _expected_reused(sc, frames):
    """Cached rays of a call: every pixel's primary ray, and the shadow ray of
    each hit pixel when the scene has 1..32 lights."""
    cam = O.camera_from(sc.camera)
    osc = O.Scene(sc.spheres, sc.materials, sc.lights)
    st = O.settings(sc.settings.max_bounces, sc.settings.sky_light, sc.settings.accumulation)
    _, hit, _, _ = O.render_frame_debug(cam, osc, st, 1, max(1, sc.settings.max_bounces))
    n = sc.camera.width * sc.camera.height
    shadow = 1 <= len(sc.lights) <= 32
    return frames * (n + (int((hit[:, 0] >= 0).sum()) if shadow else 0))


@pytest.mark.parametrize("waves", WAVES)
@pytest.mark.parametrize("name", list(CASES))
def test_instance_matches_oracle(name, waves, monkeypatch):
    sc, kahan = CASES[name]
    monkeypatch.setenv("ATX_WF_WAVES", str(waves))
    frames = 6
    cam = sc.camera
    print(f"case {name} waves={waves} {cam.width}x{cam.height} spheres={sc.num_spheres} lights={len(sc.lights)} "
          f"frames={frames} kahan={kahan}", flush=True)
    with Renderer() as r:
        r.set_kernel(ATX_KERNEL_WAVEFRONT)
        r.on_resize(cam.width, cam.height)
        r.set_settings(sc.settings)
        if kahan:
            r.set_accumulator(ATX_ACCUM_KAHAN)
        r.set_scene(sc)
        r.render_frames(cam, frames)
        st = r.stats()
        acc = r.get_accumulation().reshape(-1, 4).copy()
        img = r.get_image().reshape(-1).copy()
    oacc, oimg = oracle_accumulate(sc, frames, kahan=kahan)
    assert np.array_equal(acc, oacc)
    assert np.array_equal(img, oimg)
    assert st["reused_rays"] == _expected_reused(sc, frames)
    # per-frame calls with the debug capture (the cache and hit list are used
    # from the second call on: the camera stays)
    run_parity(sc, 3, kahan=kahan, kernel=ATX_KERNEL_WAVEFRONT)
