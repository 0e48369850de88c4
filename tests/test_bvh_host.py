"""Host-side checks of the sphere BVH (no GPU): tests/bvh_check.cpp builds
trees with ataraxia_amd/csrc/bvh_build.cpp, checks their structure and runs
a float restatement of the device traversal against the oracle's flat
closest-hit loop (ora_trace_ray) on rays chosen to stress the culling.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLAGS = ["-O2", "-ffp-contract=off", "-fno-fast-math"]


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not shutil.which("g++") or not shutil.which("gcc"):
        pytest.skip("no host C/C++ compiler")
    d = tmp_path_factory.mktemp("bvh")
    ora = d / "oracle.o"
    bvh = d / "bvh_build.o"
    exe = d / "bvh_check"
    subprocess.run(["gcc", "-std=c99", *FLAGS, "-c", os.path.join(ROOT, "oracle", "atx_oracle.c"), "-o", str(ora)],
                   check=True)
    subprocess.run(["g++", "-std=c++17", *FLAGS, "-c", os.path.join(ROOT, "ataraxia_amd", "csrc", "bvh_build.cpp"),
                    "-o", str(bvh)], check=True)
    subprocess.run(["g++", "-std=c++17", *FLAGS, os.path.join(ROOT, "tests", "bvh_check.cpp"), str(bvh), str(ora),
                    "-lm", "-o", str(exe)], check=True)
    return str(exe)


@pytest.mark.parametrize("layouts", [8, 1])
def test_bvh_structure_and_culling_match_flat_loop(checker, layouts):
    out = subprocess.run([checker, "64000", str(layouts)], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stderr
    lines = [l for l in out.stdout.splitlines() if "spheres=" in l]
    assert len(lines) == 5
    # the 257-sphere scene with a ground sphere: the ground stays out of the tree
    assert "random_257_ground" in lines[0] and "big= 1" in lines[0]
