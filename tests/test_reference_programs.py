"""The oracle against what the reference's own test programs compute
(tests/golden/reference_programs.npz, written by make_reference_programs.py
with the reference's code): test/ib_test.cc, ib_ns_test.cc, ns_test.cc,
mg_test.cc, mg_test2.cc and restr_test.cc.  The oracle's reference order
(glibc cos, lexicographic Gauss-Seidel) must give the same bits; the order
the GPU follows (correctly rounded cos, red-black) must agree to rounding.
CPU only; the GPU is held to the same fixtures in
test_gpu_reference_programs.py."""
import os

import numpy as np
import pytest

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_programs.npz"))


def ib_ns_inputs(n=16):
    iy = (np.arange(n ** 3) % (n * n)) // n
    y = (iy + 0.5) / n
    return ((1.0 - y) * y).reshape(n, n, n)


def test_ib_test_program(oracle):
    """test/ib_test.cc: spread 1 from a marker of volume 1/n^3 at the centre
    of geometry<5,1,1,1> to the v grid, binarise, interpolate back."""
    n = 32
    X = G["ib_test_X"].reshape(1, 3)
    for mode in (0, 1):
        f = oracle.spread3(1, np.ones(1), np.full(1, 1.0 / n ** 3), X, n, mode=mode)
        nz = np.flatnonzero(f)
        assert np.array_equal(nz, G["ib_test_idx"])
        if mode == 0:   # glibc cos: the reference's bits
            assert np.array_equal(f[nz], G["ib_test_val"])
        else:           # correctly rounded cos: within an ulp or two of glibc's
            assert np.max(np.abs(f[nz] / G["ib_test_val"] - 1)) <= 4e-16
        b = (f != 0).astype(np.float64).reshape(n, n - 1, n)
        out = oracle.interp3(1, b, X, n, mode=mode)
        assert out[0] == G["ib_test_out"][0] == 1.0


@pytest.mark.parametrize("name,steps,fx", [("ib_ns_test", 1, 1.0), ("ns_test", 2, 0.0)])
def test_ns_programs(oracle, name, steps, fx):
    """test/ib_ns_test.cc (Poiseuille-like u0, fx = 1, one backward-Euler
    step) and test/ns_test.cc (zero fields, two steps) on geometry<4,1,1,1>
    at the programs' tolerance 1e-15."""
    n = 16
    mu, rho, dt, tol = G[name + "_params"]
    u0 = ib_ns_inputs(n) if fx else np.zeros((n, n, n))
    walls = [np.zeros((n, n)) for _ in range(4)]
    f = (np.full((n, n, n), fx), np.zeros((n, n - 1, n)), np.zeros((n, n, n)))
    want = [G[f"{name}_{c}"] for c in "uvwp"]
    u, v, w, _, p, _ = oracle.ns3_run(4, u0, np.zeros((n, n - 1, n)), np.zeros((n, n, n)), walls, f, mu, rho, dt,
                                      tol, steps, smoother=0)
    for a, b in zip((u, v, w, p), want):
        assert np.array_equal(a.ravel(), b)
    # the red-black order (the GPU's) to the tolerance's rounding
    u, v, w, _, p, _ = oracle.ns3_run(4, u0, np.zeros((n, n - 1, n)), np.zeros((n, n, n)), walls, f, mu, rho, dt,
                                      tol, steps, smoother=1)
    scale = max(np.max(np.abs(want[0])), 1.0)
    for a, b in zip((u, v, w), want[:3]):
        assert np.max(np.abs(a.ravel() - b)) <= 1e-13 * scale
    if not fx:
        for a in (u, v, w, p):
            assert not np.any(a)


def test_mg_test_program(oracle):
    """test/mg_test.cc: its 16^3 right-hand side of order 1e-16, solved at
    the absolute tolerance 1e-16 (met after the nested start)."""
    b = G["mg_test_b"]
    x, cyc, _ = oracle.mg_solve(b.reshape(16, 16, 16), 4, dim=3, tol=float(G["mg_test_tol"]), tol_mode=0,
                                smoother=0)
    assert cyc == int(G["mg_test_cycles"])
    assert np.array_equal(x.ravel(), G["mg_test_x"])


def test_mg_test2_program(oracle):
    """test/mg_test2.cc: geometry<1,8,8,8> (16^3 cells, coarsest level 8^3
    by CG), ones with b[0] = 2 at 1e-15; the program prints L x."""
    x, cyc, _ = oracle.mg_solve(G["mg_test2_b"].reshape(16, 16, 16), 1, dim=3, W=8, H=8, D=8,
                                tol=float(G["mg_test2_tol"]), tol_mode=0, smoother=0)
    assert cyc == int(G["mg_test2_cycles"])
    assert np.array_equal(x.ravel(), G["mg_test2_x"])
    Lx = oracle.stencil_apply(x, 16, 16, 16, 2)
    assert np.array_equal(Lx.ravel(), G["mg_test2_Lx"])
    # L x is the right-hand side less its mean (the solve removes it)
    b = G["mg_test2_b"]
    assert np.max(np.abs(Lx.ravel() - (b - b.mean()))) < 1e-12


def test_restr_test_program(oracle):
    """test/restr_test.cc: restriction of its 8 values on geometry<1,1,1,1>."""
    out = oracle.restrict3(G["restr_test_fine"], 2, 2, 2)
    assert np.array_equal(out, G["restr_test_coarse"])
