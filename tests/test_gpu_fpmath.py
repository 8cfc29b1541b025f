"""The step kernels' correctly rounded reciprocal, quotient and square root
(cosmic-collisions_amd/csrc/cc_fpmath.h) against the correctly rounded
results, on the GPU (tests/recip_check.hip, built by build()): the reciprocal
for every significand at exponents -100..100, the quotient on 2^34 random
signed pairs, the root for every x in [2^-100, FLT_MAX].  The check decides rounding
by exact midpoint comparisons in double, and shows it has teeth on the
hardware reciprocal alone, which it must reject for some significands."""
import os
import subprocess

import pytest

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "recip_check")


@pytest.mark.gpu
def test_helpers_are_correctly_rounded():
    assert os.path.exists(EXE), "tests/recip_check is not built: run __graft_entry__.build()"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all correctly rounded" in r.stdout
