"""Integer helpers of the fused pair kernel, checked on the CPU with the same
header the kernels include (built with g++): the folded low-half prefilter
never misses a candidate whose value lies in an exact unary-chain window, and
the invariant-divisor division equals n / d."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_pair_math(tmp_path):
    exe = tmp_path / "wpc"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "simba_amd", "csrc"),
                    os.path.join(ROOT, "tests", "pair_math_check.cpp"), "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "misses 0" in r.stdout and "wrong 0" in r.stdout
