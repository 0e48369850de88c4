"""End to end through the C++ façade: the headless driver renders the
reference's sample scene.json and its PPM equals the oracle's image."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from ataraxia_amd import SceneGraph

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def read_ppm(path):
    data = open(path, "rb").read()
    parts = data.split(b"\n", 3)
    assert parts[0] == b"P6"
    w, h = map(int, parts[1].split())
    px = np.frombuffer(parts[3], np.uint8).reshape(h, w, 3)
    return px


def test_driver_sample_scene_matches_oracle(tmp_path):
    scene_path = os.path.join(ROOT, "tests", "golden", "sample_scene.json")
    out = tmp_path / "out.ppm"
    w, h, frames = 64, 36, 3
    print("case atx_render sample_scene 64x36x3", flush=True)
    r = subprocess.run([os.path.join(ROOT, "ataraxia_amd", "atx_render"), scene_path, str(out), str(w), str(h),
                        str(frames)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = read_ppm(out)
    g = SceneGraph.load_json(scene_path)
    sc = g.to_flat_scene(w, h)
    cam = O.camera_from(sc.camera)
    osc = O.Scene(sc.spheres, sc.materials, sc.lights)
    st = O.settings(sc.settings.max_bounces, sc.settings.sky_light, sc.settings.accumulation)
    acc = np.zeros((w * h, 4), np.float32)
    for f in range(1, frames + 1):
        acc, _, img, _, _ = O.render_frame(cam, osc, st, f, accum=acc, want_radiance=False)
    img = img.reshape(h, w)[::-1]          # the PPM is written top row first
    exp = np.stack([img & 0xFF, (img >> 8) & 0xFF, (img >> 16) & 0xFF], axis=-1).astype(np.uint8)
    assert np.array_equal(got, exp)


def test_driver_accumulation_off_uploads_once(tmp_path):
    """With accumulation off the frame index stays 1 and the reference
    re-uploads the scene every frame (Renderer.cu:175-179); the façade
    uploads it once, since the flattened scene does not change, and every
    frame is the frame-1 image."""
    scene_path = os.path.join(ROOT, "tests", "golden", "sample_scene.json")
    out = tmp_path / "off.ppm"
    w, h, frames = 48, 32, 4
    print("case atx_render sample_scene accumulation off", flush=True)
    r = subprocess.run([os.path.join(ROOT, "ataraxia_amd", "atx_render"), scene_path, str(out), str(w), str(h),
                        str(frames), "0"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "1 scene upload(s)" in r.stdout, r.stdout
    got = read_ppm(out)
    sc = SceneGraph.load_json(scene_path).to_flat_scene(w, h)
    cam = O.camera_from(sc.camera)
    osc = O.Scene(sc.spheres, sc.materials, sc.lights)
    st = O.settings(sc.settings.max_bounces, sc.settings.sky_light, False)
    _, _, img, _, _ = O.render_frame(cam, osc, st, 1, want_radiance=False)
    img = img.reshape(h, w)[::-1]
    exp = np.stack([img & 0xFF, (img >> 8) & 0xFF, (img >> 16) & 0xFF], axis=-1).astype(np.uint8)
    assert np.array_equal(got, exp)
