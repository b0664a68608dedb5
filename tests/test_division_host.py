"""The grid kernels' stencil division (vxn_device.h div_rn): a / b from y = RN(1 / b) and
two fma residual corrections (Markstein), on the CPU against the correctly rounded float
division that the reference's (f[i+1] - f[i-1]) / dv performs (gradients.py:84-97).
The C restatement in tests/csrc/div_rn_check.c is compiled with gcc and run over 2^22
random quotients of general positive divisors and 2^22 of the stencil's eps / 2 eps
divisors, plus signed zeros, infinities and NaN: every quotient with a normal result must
be bit-identical.  The uncorrected quotient RN(a y) must fail the same check, so the check
can tell them apart.  (A single correction also passes these samples: RN(a y) is almost
always within one ulp, where one correction suffices; the second one covers the cases
where it is up to 1.5 ulp off.)"""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "div_rn_check.c")


def _run(tmp_path, defines=()):
    exe = str(tmp_path / "div_rn_check")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", *defines, "-o", exe, SRC, "-lm"], check=True)
    return subprocess.run([exe, str(1 << 22)], capture_output=True, text=True)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_stencil_division_is_correctly_rounded(tmp_path):
    r = _run(tmp_path)
    assert r.returncode == 0, r.stdout
    assert "mismatches 0" in r.stdout


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_uncorrected_quotient_fails(tmp_path):
    r = _run(tmp_path, ("-DNO_CORRECTION",))
    assert r.returncode == 1 and "mismatch a=" in r.stdout
