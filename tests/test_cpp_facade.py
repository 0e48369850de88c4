"""The C++ façade (include/ataraxia.hpp) compiles on its own and the headless
driver built on it reports errors instead of exiting abnormally (CPU)."""
import os
import subprocess

import pytest

from ataraxia_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "ataraxia_amd", "atx_render")


def test_facade_header_compiles(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include "ataraxia.hpp"\n'
                   'int main() { ataraxia::Scene s; s.rootNode->addChild(std::make_shared<ataraxia::SceneNode>("c"));\n'
                   '  ataraxia::Camera c(45.0f, 0.1f, 100.0f); c.Resize(4, 4); return (int)s.materials.size(); }\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        f"-I{os.path.join(ROOT, 'include')}", str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_driver_usage_and_missing_device():
    assert os.path.exists(DRIVER)
    r = subprocess.run([DRIVER], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
    if _lib.device_count() == 0:
        r = subprocess.run([DRIVER, os.path.join(ROOT, "tests", "golden", "sample_scene.json"), "/dev/null"],
                           capture_output=True, text=True, timeout=60)
        assert r.returncode == 1 and "device" in r.stderr
