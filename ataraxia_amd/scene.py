"""Host-side scene model and the seeded synthetic scene generator.

Mirrors the reference's scene types (Engine/include/Scene.h:11-80,
Engine/include/SceneNode.h:11-21, Engine/include/Camera.h) as flat numpy
arrays with the reference's packed layouts, plus the generator for the
benchmark workloads:

* ``plain_cpu``  640x360, 4 spheres + ground, 1 light, 5 bounces, 64 frames
* ``plain_1080p`` 1920x1080, 256 spheres + ground, 8 lights, 8 bounces, 256 frames
* ``batch_mix``  64 scenes of mixed size/sphere count/lights/bounces/spp
* ``precision``  1024x1024, 512 spheres + ground, 4 lights, 6 bounces, 4096 spp
* serving requests drawn per request (resolution, spheres, spp)

Randomness is splitmix64 (Steele, Lea, Flood 2014) with our own draws:
uniform floats from the top 24 bits, log-uniform as exp(U[ln a, ln b]),
Bernoulli(0.5) from the top bit, integers U{a..b} as a + floor(u*(b-a+1)).
No library distribution is used, so the scenes are identical everywhere.
Draw order: per sphere (cx, cy, cz, r, albedo rgb, roughness, metallic),
then the ground material (albedo rgb, roughness, metallic), then per light
(px, py, pz, colour rgb, intensity).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from ._lib import LIGHT_DTYPE, MATERIAL_DTYPE, SPHERE_DTYPE, CameraDesc, Settings as _CSettings

_MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & _MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """U[0,1) with 24 random bits."""
        return (self.next_u64() >> 40) * (1.0 / (1 << 24))

    def range(self, a: float, b: float) -> float:
        return a + (b - a) * self.uniform()

    def log_uniform(self, a: float, b: float) -> float:
        la, lb = math.log(a), math.log(b)
        return math.exp(la + (lb - la) * self.uniform())

    def bernoulli(self) -> bool:
        return bool(self.next_u64() >> 63)

    def integer(self, a: int, b: int) -> int:
        return a + min(b - a, int(self.uniform() * (b - a + 1)))


@dataclass
class Settings:
    """Settings (Engine/include/Scene.h:49-56): defaults accumulation on,
    sky light off, 15 bounces."""
    accumulation: bool = True
    sky_light: bool = False
    max_bounces: int = 15

    def to_c(self) -> _CSettings:
        s = _CSettings()
        s.accumulation = 1 if self.accumulation else 0
        s.sky_light = 1 if self.sky_light else 0
        s.max_bounces = int(self.max_bounces)
        return s


@dataclass
class Camera:
    """Camera state that defines the primary rays (Engine/include/Camera.h).

    Defaults follow Camera(fov, near, far): position (0,0,3), direction
    (0,0,-1), viewport 1600x900 (Camera.cpp:14-21)."""
    fov: float = 45.0
    near_clip: float = 0.1
    far_clip: float = 100.0
    position: tuple = (0.0, 0.0, 3.0)
    direction: tuple = (0.0, 0.0, -1.0)
    width: int = 1600
    height: int = 900

    def resize(self, width: int, height: int) -> None:
        """Camera::Resize (Camera.cpp:110-127); zero sizes are rejected."""
        if width == 0 or height == 0:
            raise ValueError("Width or height cannot be zero")
        self.width, self.height = int(width), int(height)

    def to_c(self) -> CameraDesc:
        c = CameraDesc()
        c.position[:] = [np.float32(v) for v in self.position]
        c.direction[:] = [np.float32(v) for v in self.direction]
        c.fov_deg = self.fov
        c.near_clip = self.near_clip
        c.far_clip = self.far_clip
        c.width = self.width
        c.height = self.height
        return c

    @staticmethod
    def looking_at(position, target=(0.0, 0.0, 0.0), fov=45.0, width=1600, height=900) -> "Camera":
        p = np.asarray(position, dtype=np.float64)
        d = np.asarray(target, dtype=np.float64) - p
        d = d / np.linalg.norm(d)
        return Camera(fov=fov, position=tuple(float(np.float32(v)) for v in p),
                      direction=tuple(float(np.float32(v)) for v in d), width=width, height=height)


@dataclass
class FlatScene:
    """A flattened scene: the sphere list the renderer consumes
    (Renderer::traverseSceneGraph output), materials and lights."""
    spheres: np.ndarray                 # SPHERE_DTYPE
    materials: np.ndarray               # MATERIAL_DTYPE
    lights: np.ndarray                  # LIGHT_DTYPE
    settings: Settings = field(default_factory=Settings)
    camera: Camera = field(default_factory=Camera)
    spp: int = 1
    name: str = ""

    @property
    def num_spheres(self) -> int:
        return int(self.spheres.shape[0])


def make_spheres(n: int) -> np.ndarray:
    return np.zeros(n, dtype=SPHERE_DTYPE)


def make_materials(n: int) -> np.ndarray:
    m = np.zeros(n, dtype=MATERIAL_DTYPE)
    m["albedo"] = 1.0
    m["F0"] = 0.04
    m["id"] = np.arange(n)
    return m


def make_lights(n: int) -> np.ndarray:
    return np.zeros(n, dtype=LIGHT_DTYPE)


def random_scene(rng: SplitMix64, n_spheres: int, n_lights: int, extent: float, sphere_y: tuple,
                 light_y: tuple, ground: bool = True) -> tuple:
    """Random spheres (centres x,z in U[-extent, extent], y in U[sphere_y],
    radius U[0.3, 1.5]), one material per sphere (albedo U[0,1]^3, roughness
    U[0.05, 1], metallic Bernoulli(0.5), F0 0.04, no emission), an optional
    ground sphere (centre (0,-1000,0), r 1000) at index 0, and lights (x,z in
    U[-extent, extent], y in U[light_y], colour U[0.5,1]^3, intensity U[5,20])."""
    total = n_spheres + (1 if ground else 0)
    sph = make_spheres(total)
    mat = make_materials(total)
    base = 1 if ground else 0
    for i in range(n_spheres):
        k = base + i
        sph["center"][k] = (rng.range(-extent, extent), rng.range(*sphere_y), rng.range(-extent, extent))
        sph["radius"][k] = rng.range(0.3, 1.5)
        sph["material"][k] = k
        mat["albedo"][k] = (rng.uniform(), rng.uniform(), rng.uniform())
        mat["roughness"][k] = rng.range(0.05, 1.0)
        mat["metallic"][k] = 1.0 if rng.bernoulli() else 0.0
    if ground:
        sph["center"][0] = (0.0, -1000.0, 0.0)
        sph["radius"][0] = 1000.0
        sph["material"][0] = 0
        mat["albedo"][0] = (rng.uniform(), rng.uniform(), rng.uniform())
        mat["roughness"][0] = rng.range(0.05, 1.0)
        mat["metallic"][0] = 1.0 if rng.bernoulli() else 0.0
    lig = make_lights(n_lights)
    for j in range(n_lights):
        lig["position"][j] = (rng.range(-extent, extent), rng.range(*light_y), rng.range(-extent, extent))
        lig["color"][j] = (rng.range(0.5, 1.0), rng.range(0.5, 1.0), rng.range(0.5, 1.0))
        lig["intensity"][j] = rng.range(5.0, 20.0)
    return sph, mat, lig


def plain_cpu_scene(seed: int = 0) -> FlatScene:
    """640x360, 4 spheres + ground, 1 light, 5 bounces, 64 frames."""
    rng = SplitMix64(seed)
    sph, mat, lig = random_scene(rng, 4, 1, 5.0, (0.0, 3.0), (8.0, 12.0))
    cam = Camera.looking_at((0.0, 3.0, 12.0), width=640, height=360)
    return FlatScene(sph, mat, lig, Settings(True, False, 5), cam, spp=64, name="plain_cpu_640x360")


def plain_1080p_scene(seed: int = 1) -> FlatScene:
    """1920x1080, 256 spheres + ground, 8 lights, 8 bounces, 256 frames."""
    rng = SplitMix64(seed)
    sph, mat, lig = random_scene(rng, 256, 8, 20.0, (0.0, 6.0), (10.0, 20.0))
    cam = Camera.looking_at((0.0, 8.0, 40.0), width=1920, height=1080)
    return FlatScene(sph, mat, lig, Settings(True, False, 8), cam, spp=256, name="plain_1080p")


def precision_scene(seed: int = 3) -> FlatScene:
    """1024x1024, 512 spheres + ground, 4 lights, 6 bounces, 4096 spp."""
    rng = SplitMix64(seed)
    sph, mat, lig = random_scene(rng, 512, 4, 20.0, (0.0, 6.0), (10.0, 20.0))
    cam = Camera.looking_at((0.0, 8.0, 40.0), width=1024, height=1024)
    return FlatScene(sph, mat, lig, Settings(True, False, 6), cam, spp=4096, name="precision_1024")


def batch_mix_scenes(seed: int = 2, count: int = 64) -> List[FlatScene]:
    """64 scenes: W ~ U{256..1920}, H ~ U{256..1080}, N ~ logU[16, 4096]
    spheres (+ ground) with extent 20*sqrt(N/256), L ~ U{1..16}, bounces
    U{2..10}, spp U{16..256}; camera at (0, 8, 40)*extent/20."""
    rng = SplitMix64(seed)
    out = []
    for i in range(count):
        w = rng.integer(256, 1920)
        h = rng.integer(256, 1080)
        n = int(round(rng.log_uniform(16, 4096)))
        lights = rng.integer(1, 16)
        bounces = rng.integer(2, 10)
        spp = rng.integer(16, 256)
        ext = 20.0 * math.sqrt(n / 256.0)
        sph, mat, lig = random_scene(rng, n, lights, ext, (0.0, 6.0), (10.0, 20.0))
        cam = Camera.looking_at((0.0, 8.0 * ext / 20.0, 40.0 * ext / 20.0), width=w, height=h)
        out.append(FlatScene(sph, mat, lig, Settings(True, False, bounces), cam, spp=spp,
                             name=f"mix{i:02d}_{w}x{h}_n{n}"))
    return out


def serve_request_scene(rng: SplitMix64, index: int) -> FlatScene:
    """One serving request: resolution from {512x512, 1280x720, 1920x1080},
    N ~ logU[4, 1024] spheres (+ ground), spp ~ logU[16, 1024], 8 bounces,
    L ~ U{1..8}."""
    res = [(512, 512), (1280, 720), (1920, 1080)][rng.integer(0, 2)]
    n = int(round(rng.log_uniform(4, 1024)))
    spp = int(round(rng.log_uniform(16, 1024)))
    lights = rng.integer(1, 8)
    ext = max(5.0, 20.0 * math.sqrt(n / 256.0))
    sph, mat, lig = random_scene(rng, n, lights, ext, (0.0, 6.0), (10.0, 20.0))
    cam = Camera.looking_at((0.0, 8.0 * ext / 20.0, 40.0 * ext / 20.0), width=res[0], height=res[1])
    return FlatScene(sph, mat, lig, Settings(True, False, 8), cam, spp=spp,
                     name=f"req{index}_{res[0]}x{res[1]}_n{n}")


def default_scene() -> FlatScene:
    """The editor's start-up scene (Engine/src/main.cpp:235-265) flattened:
    spheres at (0,0,0), (2,0,0) and (2,2,0), radius 1, three materials and
    one white light at (10,10,0)."""
    sph = make_spheres(3)
    sph["center"] = [(0, 0, 0), (2, 0, 0), (2, 2, 0)]
    sph["radius"] = 1.0
    sph["material"] = [0, 1, 2]
    mat = make_materials(3)
    mat["albedo"] = [(1.022, 0.782, 0.344), (1.0, 0.0, 0.0), (0.972, 0.960, 0.915)]
    mat["roughness"] = [1.0, 0.3, 0.25]
    mat["metallic"] = [0.0, 0.0, 1.0]
    lig = make_lights(1)
    lig["position"][0] = (10, 10, 0)
    lig["color"][0] = (1, 1, 1)
    lig["intensity"][0] = 1.0
    cam = Camera(width=320, height=180)
    return FlatScene(sph, mat, lig, Settings(), cam, name="default")


def scene_from_camera_overrides(scene: FlatScene, width: Optional[int] = None,
                                height: Optional[int] = None) -> FlatScene:
    cam = Camera(scene.camera.fov, scene.camera.near_clip, scene.camera.far_clip,
                 scene.camera.position, scene.camera.direction,
                 width or scene.camera.width, height or scene.camera.height)
    return FlatScene(scene.spheres, scene.materials, scene.lights, scene.settings, cam,
                     scene.spp, scene.name)
