"""The XCD-grouped tile order (kernels.hpp xcd_tile), compiled for the host from the device
source: a bijection on the tile range for every tile count and group size, runs of consecutive
tiles per XCD inside whole groups, and the launch order past the last whole group."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_xcd_tile_is_a_grouped_bijection(tmp_path):
    if not os.path.exists(HIPCC) and not shutil.which("hipcc"):
        pytest.skip("hipcc not available")
    exe = tmp_path / "xcd_tile_check"
    cmd = [HIPCC, "--offload-host-only", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROOT, "chasqui_amd", "csrc"), os.path.join(HERE, "native", "xcd_tile_check.cpp"),
           "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0" in r.stdout
