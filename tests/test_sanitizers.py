"""The host-side code that parses untrusted files or builds data structures
(the scene.json reader/writer and flattening, the sphere-tree builder, the
CPU oracle) built with g++/gcc under AddressSanitizer and
UndefinedBehaviorSanitizer (with float-to-int/float-overflow checks) and
driven by tests/sanitize_host.cpp: the sample scene, every truncation of it,
3000 seeded mutations, pathological documents, random trees and oracle
frames must run clean (CPU only; no device code is built)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "ataraxia_amd", "csrc")
SAN = ["-fsanitize=address,undefined,float-cast-overflow", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
       "-g", "-O1"]
HIP_HOST = ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]


def test_host_code_under_asan_and_ubsan(tmp_path):
    if not shutil.which("g++") or not shutil.which("gcc") or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("needs gcc, g++ and the ROCm headers")
    objs = []
    o = tmp_path / "oracle.o"
    subprocess.run(["gcc", "-std=c99", *SAN, "-ffp-contract=off", "-D_GNU_SOURCE", "-c",
                    os.path.join(ROOT, "oracle", "atx_oracle.c"), "-o", str(o)], check=True)
    objs.append(str(o))
    for name in ("scene_graph", "host_common", "bvh_build"):
        o = tmp_path / f"{name}.o"
        subprocess.run(["g++", "-std=c++17", *SAN, "-ffp-contract=off", *HIP_HOST, "-c",
                        os.path.join(CSRC, f"{name}.cpp"), "-o", str(o)], check=True)
        objs.append(str(o))
    o = tmp_path / "driver.o"
    subprocess.run(["g++", "-std=c++17", *SAN, *HIP_HOST, "-c", os.path.join(ROOT, "tests", "sanitize_host.cpp"),
                    "-o", str(o)], check=True)
    exe = tmp_path / "sanitize_host"
    subprocess.run(["g++", *SAN, str(o), *objs, "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib",
                    "-lpthread", "-o", str(exe)], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe), os.path.join(ROOT, "tests", "golden", "sample_scene.json")], capture_output=True,
                       text=True, timeout=600, env=env)
    print(r.stdout)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "sanitize_host: ok" in r.stdout
