"""The correctly rounded cos of the delta function (host build of the same
header the kernels use): the table-driven fast path with its rounding test
must give the full double-double series' result, which must be cos rounded
from quad precision (2M arguments over the delta's range |x| <= 3 pi / 2, a
quarter of them within 2^-50 .. 1 of a multiple of pi/2); and the fast delta
argument must equal the exact integer rounding RN53(RN64(pi_64 r)) n / 2
(2M values of r over |n r| < 3, n = 2 .. 2^13, with tiny ones and ones at
the support's ends)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "afmibins_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_fast_cos_equals_full_series_and_quad(tmp_path):
    exe = str(tmp_path / "cos_paths")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-I", CSRC,
                           os.path.join(HERE, "native", "cos_paths.cpp"), "-o", exe, "-lquadmath"])
    for seed in (1, 2):
        out = subprocess.run([exe, "1000000", str(seed)], check=True, capture_output=True, text=True).stdout
        bad_full, bad_quad, bad_arg = map(int, out.split())
        assert bad_full == 0 and bad_quad == 0 and bad_arg == 0, out
