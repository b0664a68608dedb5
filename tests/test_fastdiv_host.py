"""The kernels' division of a sample index by the samples per ray (vxn_fastdiv.h: a
multiply-high, an add and a shift from host-computed constants) against C integer division:
the header itself is compiled with gcc and checked for every divisor up to 4096 and 2000
random larger ones, at the multiples of the divisor and either side of them, near 2^31 and
at random numerators below 2^31."""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "fastdiv_check.c")
INC = os.path.join(os.path.dirname(HERE), "voxneus_amd", "csrc")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_fastdiv_equals_integer_division(tmp_path):
    exe = str(tmp_path / "fastdiv_check")
    subprocess.run(["gcc", "-O2", f"-I{INC}", "-o", exe, SRC], check=True)
    r = subprocess.run([exe, "20000"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    assert "mismatches 0" in r.stdout
