"""The transport fits' logarithm and exponential (corefl-cpc_amd/csrc/fastmath.h,
shared by host and device) against the C library: at most 4 units in the
last place over 1e-300 .. 1e300 (log) and -700 .. 700 (exp)."""
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
REPO = HERE.parent


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_fast_log_exp_ulp(tmp_path):
    exe = tmp_path / "check_fastmath"
    subprocess.run(["g++", "-O2", "-std=c++17", f"-I{REPO / 'corefl-cpc_amd' / 'csrc'}",
                    str(HERE / "fastmath" / "check_fastmath.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    worst_log, worst_exp = float(out[0]), float(out[2])
    assert worst_log <= 4.0, out
    assert worst_exp <= 4.0, out
