"""The reconstructions' exact division by 6 (numerics.hpp div6: an FMA remainder correction,
Markstein's theorem) against IEEE x / 6, compiled for the host from the device source."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_div6_bitwise_equals_division(tmp_path):
    if not os.path.exists(HIPCC) and not shutil.which("hipcc"):
        pytest.skip("hipcc not available")
    exe = tmp_path / "div6_check"
    cmd = [HIPCC, "--offload-host-only", "-O2", "-std=c++17", "-ffp-contract=off",
           "-I" + os.path.join(ROOT, "chasqui_amd", "csrc"), os.path.join(HERE, "native", "div6_check.cpp"),
           "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True)
    r = subprocess.run([str(exe), "20000000"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0" in r.stdout
