"""Host-side checks of the persistent kernel's launch layout (no GPU): chunk
ids, pixel and sample offsets and the 32-bit guard of
ataraxia_amd/csrc/wavefront_host.h, the job order, the first-bounce cache
layout, and the device-block pool's hand-out rules, compiled host-only by
hipcc."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_launch_layout(tmp_path):
    if not (os.path.exists(HIPCC) or shutil.which("hipcc")):
        pytest.skip("hipcc not available")
    exe = tmp_path / "wavefront_layout_check"
    r = subprocess.run([HIPCC if os.path.exists(HIPCC) else "hipcc", "-std=c++17", "--offload-host-only", "-O1",
                        f"-I{os.path.join(ROOT, 'ataraxia_amd', 'csrc')}", f"-I{os.path.join(ROOT, 'include')}",
                        os.path.join(ROOT, "tests", "wavefront_layout_check.cpp"), "-o", str(exe)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert "all checks passed" in out.stdout
