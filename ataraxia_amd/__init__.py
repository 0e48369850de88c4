"""ataraxia_amd -- MI355X-native (gfx950) Ataraxia per-pixel path tracer.

The compute path is libatx.so (hand-written HIP kernels behind the C ABI in
include/atx.h).  This package is the Python view of that ABI used by the
tests and the benchmark; it has no CPU rendering path of its own.
"""
from ._lib import (ATX_ACCEL_BVH, ATX_ACCEL_NONE, ATX_ACCUM_KAHAN, ATX_ACCUM_NAIVE, ATX_KERNEL_PLAIN,
                   ATX_KERNEL_WAVEFRONT, AtxError, LIB_PATH, device_count, last_error)
from .batch import Batch
from .renderer import Renderer, camera_matrices, camera_ray_directions
from .scene_graph import SceneGraph, write_ppm
from .server import Server
from .scene import (Camera, FlatScene, Settings, SplitMix64, batch_mix_scenes, default_scene,
                    plain_1080p_scene, plain_cpu_scene, precision_scene, serve_request_scene)

__all__ = [
    "ATX_ACCEL_BVH", "ATX_ACCEL_NONE", "ATX_ACCUM_KAHAN", "ATX_ACCUM_NAIVE", "ATX_KERNEL_PLAIN",
    "ATX_KERNEL_WAVEFRONT", "AtxError", "Batch", "SceneGraph", "Server", "write_ppm", "LIB_PATH", "device_count",
    "last_error",
    "Renderer", "camera_matrices", "camera_ray_directions", "Camera", "FlatScene", "Settings",
    "SplitMix64", "batch_mix_scenes", "default_scene", "plain_1080p_scene", "plain_cpu_scene",
    "precision_scene", "serve_request_scene",
]
