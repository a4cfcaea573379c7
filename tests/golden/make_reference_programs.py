"""Writes tests/golden/reference_programs.npz: what the reference's own test
programs compute, from the inputs they hold, by the reference's own code
(oracle/_ref, instantiated from the reference headers by oracle/ref_shim.cc).

  ib_test     test/ib_test.cc:73-93 -- geometry<5,1,1,1>, one marker of
              volume 1/n^3 at (0.5, 0.5, 0.5) holding 1, spread to the v grid
              (mac_grid<y>), non-zero entries set to 1, interpolated back.
              Kept: the spread field's non-zero entries (index, value) and the
              marker's final value.
  ib_ns_test  test/ib_ns_test.cc:29-63 -- ns::solver<geometry<4,1,1,1>>,
              mu = rho = 1, dt = 1e-5, tol = 1e-15, u0 = (H - y) y at the cell
              centres, fx = 1, one step (backward Euler): u, v, w and get_p().
  ns_test     test/ns_test.cc:29-84 -- the same solver over zero fields,
              dt = 5e-5, two steps: u, v, w, p.
  mg_test     test/mg_test.cc:31-37 -- mg::solver<geometry<4,1,1,1>>(1e-16,
              2, 2).solve(b) on the program's 16^3 right-hand side (entries of
              order 1e-16): b, the solution and the cycle count.
  mg_test2    test/mg_test2.cc:39-47 -- mg::solver<geometry<1,8,8,8>>(1e-15,
              2, 2) on ones with b[0] = 2; the program prints L * x: kept
              with the cycle count.
  restr_test  test/restr_test.cc:38-50 -- restriction of the program's 8
              values on geometry<1,1,1,1>.

Arrays are flat in the reference's order (x fastest, then y, then z).  Run in
a container that has the reference tree:
    python tests/golden/make_reference_programs.py
"""
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
TEST = "/root/reference/test"
NUM = r"[-+]?[0-9]*\.?[0-9]+(?:[eE][-+]?[0-9]+)?"


def read_values(name, opener, count):
    """The numbers of the brace initialiser that follows `opener` in test/name."""
    text = open(os.path.join(TEST, name)).read()
    start = text.index(opener) + len(opener)
    vals = np.array([float(t) for t in re.findall(NUM, text[start:text.index("};", start)])])
    assert vals.size == count, (name, vals.size)
    return vals


def ns_program(R, P, mu, rho, dt, tol, u0, fx, steps):
    n = 16
    u = np.ascontiguousarray(u0, dtype=np.float64).copy()
    v, w = np.zeros(n * (n - 1) * n), np.zeros(n ** 3)
    walls = [np.zeros(n * n) for _ in range(4)]
    f = (np.ascontiguousarray(fx, dtype=np.float64), np.zeros(n * (n - 1) * n), np.zeros(n ** 3))
    p = np.empty(n ** 3)
    rc = R.ref_ns3_run(0, mu, rho, dt, tol, *(a.ctypes.data_as(P) for a in (u, v, w, *walls, *f)), steps,
                       p.ctypes.data_as(P))
    assert rc == 0
    return u, v, w, p


def main():
    sys.path.insert(0, ROOT)
    from oracle import oracle
    oracle.build()
    R, P = oracle.ref(), oracle._P
    out = {}

    # test/ib_test.cc
    n = 32
    field = np.empty(n * (n - 1) * n)
    val = np.zeros(1)
    size = R.ref_ib_test(0.5, 0.5, 0.5, field.ctypes.data_as(P), val.ctypes.data_as(P))
    assert size == field.size
    nz = np.flatnonzero(field)
    out.update(ib_test_X=np.array([0.5, 0.5, 0.5]), ib_test_idx=nz.astype(np.int64), ib_test_val=field[nz],
               ib_test_out=val)

    # test/ib_ns_test.cc: u0 = (H - y) y with y = (iy + 0.5)/n, fx = 1
    n = 16
    iy = (np.arange(n ** 3) % (n * n)) // n
    y = (iy + 0.5) / n
    u0 = (1.0 - y) * y
    u, v, w, p = ns_program(R, P, 1.0, 1.0, 1e-5, 1e-15, u0, np.ones(n ** 3), 1)
    out.update(ib_ns_test_u=u, ib_ns_test_v=v, ib_ns_test_w=w, ib_ns_test_p=p,
               ib_ns_test_params=np.array([1.0, 1.0, 1e-5, 1e-15]))

    # test/ns_test.cc: zero fields, two steps
    u, v, w, p = ns_program(R, P, 1.0, 1.0, 5e-5, 1e-15, np.zeros(n ** 3), np.zeros(n ** 3), 2)
    out.update(ns_test_u=u, ns_test_v=v, ns_test_w=w, ns_test_p=p, ns_test_params=np.array([1.0, 1.0, 5e-5, 1e-15]))

    # test/mg_test.cc
    b = read_values("mg_test.cc", "dt::vector<4096> b = {", 4096)
    x = b.copy()
    cyc = R.ref_mg_solve(0, 1e-16, 2, 2, x.ctypes.data_as(P))
    out.update(mg_test_b=b, mg_test_x=x, mg_test_cycles=np.int64(cyc), mg_test_tol=np.float64(1e-16))

    # test/mg_test2.cc: geometry<1,8,8,8> (16^3 cells), ones with a spike
    b = np.ones(4096)
    b[0] = 2
    x = b.copy()
    cyc2 = R.ref_mg_solve(2, 1e-15, 2, 2, x.ctypes.data_as(P))
    Lx = np.empty(4096)
    assert R.ref_lap_apply(2, x.ctypes.data_as(P), Lx.ctypes.data_as(P)) == 0
    out.update(mg_test2_b=b, mg_test2_x=x, mg_test2_Lx=Lx, mg_test2_cycles=np.int64(cyc2),
               mg_test2_tol=np.float64(1e-15))

    # test/restr_test.cc
    fine = read_values("restr_test.cc", "dt::vector<8> fine = {", 8)
    coarse = np.zeros(1)
    assert R.ref_restrict_g1(fine.ctypes.data_as(P), coarse.ctypes.data_as(P)) == 0
    out.update(restr_test_fine=fine, restr_test_coarse=coarse)

    np.savez_compressed(os.path.join(HERE, "reference_programs.npz"), **out)
    print(f"ib_test: {nz.size} non-zero v faces, value {val[0]!r}; mg_test {cyc} cycles; mg_test2 {cyc2} cycles; "
          f"restr_test {coarse[0]!r}")


if __name__ == "__main__":
    main()
