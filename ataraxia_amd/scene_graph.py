"""SceneGraph: Python view of the C ABI's scene graph and scene.json I/O
(Scene.h:58-80, SceneNode.h:23-66, Utils.cpp:3-187, Renderer.cu:67-96)."""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import LIGHT_DTYPE, MATERIAL_DTYPE, SPHERE_DTYPE, check, lib, ptr
from .scene import Camera, FlatScene, Settings


def _f3(v):
    return (C.c_float * 3)(*[float(x) for x in v]) if v is not None else None


class SceneGraph:
    ROOT = 0

    def __init__(self, handle: Optional[C.c_void_p] = None):
        if handle is None:
            handle = C.c_void_p()
            check(lib.atx_scene_create(C.byref(handle)))
        self._h = handle

    def close(self) -> None:
        if self._h:
            lib.atx_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def load_json(path: str) -> "SceneGraph":
        h = C.c_void_p()
        check(lib.atx_scene_load_json(path.encode(), C.byref(h)))
        return SceneGraph(h)

    @staticmethod
    def parse_json(text: str) -> "SceneGraph":
        b = text.encode()
        h = C.c_void_p()
        check(lib.atx_scene_parse_json(b, len(b), C.byref(h)))
        return SceneGraph(h)

    def to_json(self) -> str:
        n = C.c_uint64()
        check(lib.atx_scene_to_json(self._h, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value + 1)
        check(lib.atx_scene_to_json(self._h, buf, n.value + 1, C.byref(n)))
        return buf.value.decode()

    def save_json(self, path: str) -> None:
        check(lib.atx_scene_save_json(self._h, path.encode()))

    def add_node(self, parent: int, name: str = "Untitled", position=(0, 0, 0), rotation_xyzw=(0, 0, 0, 0),
                 scale=(1, 1, 1)) -> int:
        k = C.c_uint32()
        rot = (C.c_float * 4)(*[float(x) for x in rotation_xyzw])
        check(lib.atx_scene_add_node(self._h, parent, name.encode(), _f3(position), rot, _f3(scale), C.byref(k)))
        return k.value

    def add_sphere(self, node: int, center, radius: float, material: int) -> None:
        s = _lib.Sphere()
        s.center[:] = [float(x) for x in center]
        s.radius = radius
        s.material = material
        check(lib.atx_scene_add_sphere(self._h, node, C.byref(s)))

    def add_material(self, albedo=(1, 1, 1), roughness=0.0, metallic=0.0, F0=(0.04, 0.04, 0.04),
                     emission_color=(0, 0, 0), emission_intensity=0.0) -> None:
        m = _lib.Material()
        m.albedo[:] = [float(x) for x in albedo]
        m.roughness, m.metallic = roughness, metallic
        m.F0[:] = [float(x) for x in F0]
        m.emission_color[:] = [float(x) for x in emission_color]
        m.emission_intensity = emission_intensity
        check(lib.atx_scene_add_material(self._h, C.byref(m)))

    def add_light(self, position, color=(1, 1, 1), intensity=1.0) -> None:
        li = _lib.Light()
        li.position[:] = [float(x) for x in position]
        li.color[:] = [float(x) for x in color]
        li.intensity = intensity
        check(lib.atx_scene_add_light(self._h, C.byref(li)))

    @property
    def settings(self) -> Settings:
        s = _lib.Settings()
        check(lib.atx_scene_get_settings(self._h, C.byref(s)))
        return Settings(bool(s.accumulation), bool(s.sky_light), int(s.max_bounces))

    @settings.setter
    def settings(self, st: Settings) -> None:
        check(lib.atx_scene_set_settings(self._h, C.byref(st.to_c())))

    def set_camera(self, position, direction, fov: float = 45.0) -> None:
        check(lib.atx_scene_set_camera(self._h, _f3(position), _f3(direction), fov))

    def get_camera(self):
        p, d, f = (C.c_float * 3)(), (C.c_float * 3)(), C.c_float()
        check(lib.atx_scene_get_camera(self._h, p, d, C.byref(f)))
        return tuple(p), tuple(d), f.value

    def counts(self) -> dict:
        v = [C.c_uint64() for _ in range(4)]
        check(lib.atx_scene_counts(self._h, *[C.byref(x) for x in v]))
        return dict(zip(["nodes", "spheres", "materials", "lights"], [x.value for x in v]))

    def flatten(self) -> np.ndarray:
        n = C.c_uint64()
        check(lib.atx_scene_flatten(self._h, None, 0, C.byref(n)))
        out = np.zeros(n.value, SPHERE_DTYPE)
        if n.value:
            check(lib.atx_scene_flatten(self._h, ptr(out), n.value, C.byref(n)))
        return out

    def materials(self) -> np.ndarray:
        n = C.c_uint64()
        check(lib.atx_scene_get_materials(self._h, None, 0, C.byref(n)))
        out = np.zeros(n.value, MATERIAL_DTYPE)
        if n.value:
            check(lib.atx_scene_get_materials(self._h, ptr(out), n.value, C.byref(n)))
        return out

    def lights(self) -> np.ndarray:
        n = C.c_uint64()
        check(lib.atx_scene_get_lights(self._h, None, 0, C.byref(n)))
        out = np.zeros(n.value, LIGHT_DTYPE)
        if n.value:
            check(lib.atx_scene_get_lights(self._h, ptr(out), n.value, C.byref(n)))
        return out

    def to_flat_scene(self, width: int, height: int, spp: int = 1) -> FlatScene:
        pos, d, fov = self.get_camera()
        cam = Camera(fov=fov, position=pos, direction=d, width=width, height=height)
        return FlatScene(self.flatten(), self.materials(), self.lights(), self.settings, cam, spp=spp,
                         name="scene.json")


def write_ppm(path: str, rgba: np.ndarray, flip_vertical: bool = True) -> None:
    a = np.ascontiguousarray(rgba, dtype=np.uint32)
    h, w = a.shape
    check(lib.atx_write_ppm(path.encode(), ptr(a), w, h, 1 if flip_vertical else 0))
