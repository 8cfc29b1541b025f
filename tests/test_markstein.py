"""The pairwise pass's divisions by its two fixed divisors, and by a divisor
whose reciprocal four quotients share, use a single residual correction
(qdiv_by, cosmic-collisions_amd/csrc/cc_kernels.hip).
Its dividends are a distance (0, or above 2^-75 as the root of a sum of FP32
squares) and the cut minus a distance (0, or above 2^-24).  This test runs an
exhaustive check of the correction against the correctly rounded quotient
over every FP32 dividend in [2^-80, 1] and 0 (tests/markstein_check.c), and
shows that the bound matters: below it, where the residual underflows, the
one-step quotient differs."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GCC = shutil.which("gcc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if GCC is None:
        pytest.skip("gcc is not installed")
    exe = str(tmp_path_factory.mktemp("mk") / "markstein_check")
    subprocess.run([GCC, "-O2", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "markstein_check.c"), "-lm"],
                   check=True)
    return exe


def test_one_step_division_exact_over_the_dividends(checker):
    r = subprocess.run([checker], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert r.stdout.count(": 0 differences") == 2


def test_one_step_division_fails_where_the_residual_underflows(checker):
    r = subprocess.run([checker, "0x1p-126", "0x1p-120"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 1, r.stdout


def test_one_step_division_exact_for_random_divisors(checker):
    """The same correction from a shared divisor's correctly rounded
    reciprocal (the pairwise pass's direction and table quotients), on 2e8
    random dividend and divisor pairs with exponents in [-60, 60]."""
    r = subprocess.run([checker, "random", "200000000"], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
