"""The multiply-and-shift division by launch-time divisors (kernels.hpp fast_div / fdiv) that
the face-flux kernels use for their cell and thread indices, compiled for the host from the
device source and compared with integer division over divisors 1..5000 and large ones, at
numerators around their multiples up to 2^31 - 1 (the Granlund-Montgomery bound it relies on
covers every numerator below 2^31)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_fast_div_matches_integer_division(tmp_path):
    if not os.path.exists(HIPCC) and not shutil.which("hipcc"):
        pytest.skip("hipcc not available")
    exe = tmp_path / "fast_div_check"
    cmd = [HIPCC, "--offload-host-only", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROOT, "chasqui_amd", "csrc"), os.path.join(HERE, "native", "fast_div_check.cpp"),
           "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0" in r.stdout
