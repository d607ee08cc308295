"""Host build of the device sine/cosine (csrc/gpe_math.hpp) checked against
long-double sin/cos, and its E-argument form against the scalar form."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_fast_sincos_accuracy_and_vector_form(tmp_path):
    exe = tmp_path / "math_check"
    cmd = ["g++", "-O2", "-std=c++17", "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", "-x", "c++",
           os.path.join(ROOT, "tests", "native", "math_check.cpp"), "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, timeout=300)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=300).stdout
    vals = dict(line.split(None, 1) for line in out.strip().splitlines())
    assert float(vals["fp64_max_abs_err_eps"]) <= 1.0
    assert float(vals["fp32_max_abs_err_eps"]) <= 1.0
    assert all(float(v) == 0.0 for v in vals["wide"].split())
    assert int(vals["vector_mismatch_fp64"]) == 0
    assert int(vals["vector_mismatch_fp32"]) == 0
