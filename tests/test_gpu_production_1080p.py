"""The 1080p workload (1920x1080, 256 spheres + ground, 8 lights, 8 bounces)
through the renderer's production path at full size: the persistent kernel
with the sphere tree, the first-bounce cache and its hit list, chunked job
hand-out and k_accumulate (the per-pixel function is Renderer.cu:287-387,
the frame-ordered sum Renderer.cu:164-168).

One render call of 20 frames (27 M path jobs: the 8-wave instance) and one
more call of 3 frames that continues the accumulation (the 6-wave instance)
are checked at 4096 sampled pixels against the oracle's frame-by-frame sums
after 20 and 23 frames, accumulation and RGBA8 bit for bit; the compensated
accumulator is checked the same way over one call of 12 frames.
"""
import numpy as np
import pytest

import oracle as O
from ataraxia_amd import ATX_ACCUM_KAHAN, ATX_KERNEL_WAVEFRONT, Renderer, plain_1080p_scene

pytestmark = pytest.mark.gpu

ORACLE_THREADS = 16


def _oracle(sc, frames, pix, kahan=False):
    st = sc.settings
    return O.accumulate_pixels(O.camera_from(sc.camera), O.Scene(sc.spheres, sc.materials, sc.lights),
                               O.settings(st.max_bounces, st.sky_light, st.accumulation), 1, frames, pix,
                               kahan=kahan, threads=ORACLE_THREADS)


def _pixels(sc, k, seed):
    n = sc.camera.width * sc.camera.height
    return np.sort(np.random.default_rng(seed).choice(n, size=k, replace=False))


def test_1080p_production_path_against_oracle():
    sc = plain_1080p_scene()
    cam = sc.camera
    pix = _pixels(sc, 4096, 11)
    print(f"case plain_1080p production path {cam.width}x{cam.height} calls of 20 + 3 frames", flush=True)
    with Renderer() as r:
        r.set_kernel(ATX_KERNEL_WAVEFRONT)
        r.on_resize(cam.width, cam.height)
        r.set_settings(sc.settings)
        r.set_scene(sc)
        r.render_frames(cam, 20)
        st = r.stats()
        acc20 = r.get_accumulation().reshape(-1, 4)[pix].copy()
        img20 = r.get_image().reshape(-1)[pix].copy()
        r.render_frames(cam, 3)
        acc23 = r.get_accumulation().reshape(-1, 4)[pix].copy()
        img23 = r.get_image().reshape(-1)[pix].copy()
    assert st["samples"] == cam.width * cam.height * 20 and st["launches"] == 1
    assert st["box_tests"] > 0 and st["reused_rays"] > 0
    oacc, oimg = _oracle(sc, 20, pix)
    assert np.array_equal(acc20, oacc)
    assert np.array_equal(img20, oimg)
    oacc, oimg = _oracle(sc, 23, pix)
    assert np.array_equal(acc23, oacc)
    assert np.array_equal(img23, oimg)


def test_1080p_production_path_kahan_against_oracle():
    sc = plain_1080p_scene()
    cam = sc.camera
    pix = _pixels(sc, 2048, 12)
    print(f"case plain_1080p production path kahan {cam.width}x{cam.height} 12 frames", flush=True)
    with Renderer() as r:
        r.set_kernel(ATX_KERNEL_WAVEFRONT)
        r.on_resize(cam.width, cam.height)
        r.set_settings(sc.settings)
        r.set_accumulator(ATX_ACCUM_KAHAN)
        r.set_scene(sc)
        r.render_frames(cam, 12)
        acc = r.get_accumulation().reshape(-1, 4)[pix].copy()
        img = r.get_image().reshape(-1)[pix].copy()
    oacc, oimg = _oracle(sc, 12, pix, kahan=True)
    assert np.array_equal(acc, oacc)
    assert np.array_equal(img, oimg)
