"""The C++ drop-in headers: a program in the style of the reference's test/
programs, compiled against include/afm/afm_dropin.hpp by build(), runs on
the GPU through libafm."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "dropin_main")


def test_dropin_programs():
    assert os.path.exists(BIN), "tests/cpp/dropin_main is built by __graft_entry__.build()"
    out = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    v = dict(line.split(None, 1) for line in out.stdout.strip().splitlines())
    assert v["done"] == "1"
    # the v-grid window of a marker at (0.5, 0.5) has one row at r = 2h,
    # where the cosine delta is exactly zero: 12 of 16 taps are non-zero
    assert int(v["ib_nonzero"]) == 12
    assert abs(float(v["ib_value"]) - 1.0) < 1e-14
    assert float(v["spread_momentum_error"]) < 1e-13
    assert 1 <= int(v["mg_cycles"]) <= 20
    assert float(v["mg_rel_error"]) < 1e-9
    assert float(v["ns_zero_max"]) == 0.0
    assert float(v["ns_uniform_error"]) < 1e-12
    # the reference's 3-D geometry through the same headers
    assert 1 <= int(v["mg3_cycles"]) <= 30
    assert float(v["mg3_rms"]) <= 1.05e-12    # the solve's own test is <= 1e-12 in its summation order
    assert float(v["ns3_vw_max"]) < 1e-10
    assert float(v["ns3_xz_spread"]) < 1e-12
    # u = (1 - y) y has u'' = -2: over one full step (dt = 5e-5, mu = rho = 1)
    # the mid-channel value falls by 2 mu dt (the discrete laplacian of a
    # quadratic is exact away from the walls)
    assert abs(float(v["ns3_u_decay"]) - 2 * 5e-5) < 1e-9
    # 3-D markers through ib::lagrangian_grid / mac_grid / eulerian_grid: the
    # v-face window of a marker at the centre has one zero row (r = 2h in y)
    assert int(v["ib3_nonzero"]) == 48
    assert abs(float(v["ib3_value"]) - 1.0) < 1e-14
    assert float(v["spread3_node_error"]) < 1e-13
    assert float(v["spread3_w_error"]) < 1e-13
    assert float(v["interp3_unity_error"]) < 1e-14
    # the protected operators of mg::solver (test/mg_test2.cc, restr_test.cc)
    assert float(v["L8_error"]) < 1e-13
    assert float(v["L8_ones"]) == 0.0
    # the 64 weights of a fine cell sum to 1/8 and every product is exact
    # for these dyadic inputs: the mean 1.625 / 8, then added once more
    assert float(v["restr_first"]) == 0.203125
    assert float(v["restr_second"]) == 0.40625
    assert float(v["prolong_error"]) == 0.0
    # a matrix the program fills: y = A x, CG, one lexicographic sweep
    assert float(v["csr_apply_error"]) == 0.0
    assert 1 <= int(v["csr_cg_iters"]) <= 400
    assert float(v["csr_cg_error"]) < 1e-7
    assert float(v["csr_smooth_error"]) == 0.0
