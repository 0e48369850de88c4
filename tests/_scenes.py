"""Small mixed scenes shared by the GPU tests of the batched and served forms."""
import numpy as np

import oracle as O
from ataraxia_amd import Camera, Settings, default_scene, plain_1080p_scene, plain_cpu_scene
from ataraxia_amd.scene import (FlatScene, SplitMix64, make_lights, make_materials, make_spheres, random_scene,
                                scene_from_camera_overrides)


def mixed_scenes():
    out = []
    s = scene_from_camera_overrides(default_scene(), 37, 23)
    s.settings = Settings(True, False, 6)
    s.spp = 3
    out.append(s)
    s = scene_from_camera_overrides(plain_cpu_scene(), 64, 36)
    s.spp = 4
    out.append(s)
    s = scene_from_camera_overrides(plain_1080p_scene(), 96, 54)
    s.spp = 2
    out.append(s)
    rng = SplitMix64(31)
    sph, mat, lig = random_scene(rng, 40, 0, 8.0, (0.0, 3.0), (8.0, 12.0))
    out.append(FlatScene(sph, mat, lig, Settings(True, True, 3), Camera.looking_at((0, 4, 16), width=30, height=20),
                         spp=5, name="nolight-sky"))
    out.append(FlatScene(make_spheres(0), make_materials(0), make_lights(0), Settings(True, True, 2),
                         Camera(width=1, height=1), spp=1, name="empty-1x1"))
    sph, mat, lig = random_scene(rng, 12, 3, 6.0, (0.0, 3.0), (8.0, 12.0))
    out.append(FlatScene(sph, mat, lig, Settings(True, False, 0), Camera.looking_at((0, 3, 12), width=20, height=10),
                         spp=2, name="zero-bounce"))
    # no sphere, several frames: a first-bounce cache whose hit list is empty,
    # so the scene has no path job at all inside a launch shared with others
    out.append(FlatScene(make_spheres(0), make_materials(0), make_lights(2), Settings(True, True, 4),
                         Camera(width=24, height=16), spp=3, name="sky-only"))
    return out


def oracle_accumulate(sc, frames, kahan=False):
    cam = O.camera_from(sc.camera)
    osc = O.Scene(sc.spheres, sc.materials, sc.lights)
    st = O.settings(sc.settings.max_bounces, sc.settings.sky_light, sc.settings.accumulation)
    n = sc.camera.width * sc.camera.height
    acc = np.zeros((n, 4), np.float32)
    comp = np.zeros((n, 4), np.float32) if kahan else None
    img = None
    for f in range(1, frames + 1):
        acc, comp, img, _, _ = O.render_frame(cam, osc, st, f, accum=acc, comp=comp, want_radiance=False)
    return acc, img
