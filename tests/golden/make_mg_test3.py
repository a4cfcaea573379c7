"""Writes tests/golden/mg_test3.npz from the reference's test/mg_test3.cc:
its 16^3 right-hand side b (the fixture data the test holds) and the solution
of mg::solver<geometry<4,1,1,1>>(tol, 2, 2).solve(b) computed by the
reference's own code (oracle/_ref, built from the reference headers).

The test asks for tol = 1e-16, which the solve never reaches (its RMS
residual stalls near 6e-16 in either smoothing order, and the reference's
loop has no bound), so the expected output is taken at tol = 1e-15.

Run in a container that has the reference tree:
    python tests/golden/make_mg_test3.py
"""
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = "/root/reference/test/mg_test3.cc"
TOL = 1e-15


def read_b(path=SRC):
    text = open(path).read()
    start = text.index("dt::vector<4096> b {") + len("dt::vector<4096> b {")
    body = text[start:text.index("};", start)]
    b = np.array([float(t) for t in re.findall(r"[-+]?[0-9]*\.?[0-9]+(?:[eE][-+]?[0-9]+)?", body)])
    assert b.size == 4096
    return b


def main():
    sys.path.insert(0, ROOT)
    from oracle import oracle
    oracle.build()
    R = oracle.ref()
    b = read_b()
    x = b.copy()
    cycles = R.ref_mg_solve(0, TOL, 2, 2, x.ctypes.data_as(oracle._P))
    np.savez(os.path.join(HERE, "mg_test3.npz"), b=b, x=x, cycles=np.int64(cycles), tol=np.float64(TOL))
    print(f"mg_test3: {cycles} V-cycles at tol {TOL}")


if __name__ == "__main__":
    main()
