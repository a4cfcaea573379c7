"""The reference's native 3-D geometry for the multigrid solve (n^3 cells,
periodic x and z, walls in y) on the GPU, against the oracle: the laplacian
and one V-cycle bit-exact (the oracle's red-black order, the same row
arithmetic), solves to rounding of the oracle's red-black solve and within
the tolerance of its lexicographic one -- which is bit-exact to the reference
build (test_oracle_pins.py::test_mg_solve_matches_reference)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def T(a, dev):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)


def H(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("L", [1, 2, 3, 4, 5])
def test_laplacian3_bit_exact(ctx_of, oracle, dev, L):
    n = 1 << L
    ctx = ctx_of(64)
    x = np.random.default_rng(L).standard_normal((n, n, n))
    got = H(ctx.laplacian3(T(x, dev)))
    assert np.array_equal(got.ravel(), oracle.stencil_apply(x, n, n, n, n))


@pytest.mark.parametrize("L,nu1,nu2", [(1, 2, 2), (2, 2, 2), (3, 2, 2), (4, 2, 2), (4, 1, 1), (4, 2, 1), (5, 2, 2)])
def test_vcycle3_bit_exact(ctx_of, oracle, dev, L, nu1, nu2):
    n = 1 << L
    ctx = ctx_of(64)
    r = np.random.default_rng(10 * L + nu1 + 3 * nu2).standard_normal((n, n, n))
    got = H(ctx.mg3_vcycle(T(r, dev), nu1=nu1, nu2=nu2))
    want = oracle.mg_vcycle(r, L, dim=3, W=1, H=1, D=1, nu1=nu1, nu2=nu2, smoother=1)
    assert np.array_equal(got.ravel(), want.ravel())


@pytest.mark.parametrize("L", [3, 4, 5])
def test_mg3_solve_matches_oracle(ctx_of, oracle, dev, L):
    n = 1 << L
    ctx = ctx_of(64)
    # test/mg_test3.cc's kind of input: U(0, 1) right-hand side, mean kept
    rhs = np.random.default_rng(300 + L).uniform(0, 1, (n, n, n))
    x = T(rhs, dev)
    cyc, res = ctx.mg3_solve(x, tol=1e-10, tol_mode=1)
    got = H(x).ravel()
    want, cyc_o, _ = oracle.mg_solve(rhs, L, dim=3, tol=1e-10, tol_mode=1, smoother=1)
    assert abs(cyc - cyc_o) <= 1
    d = (got - got.mean()) - (want.ravel() - want.mean())
    assert np.max(np.abs(d)) <= 1e-12 * np.max(np.abs(want - want.mean()))
    # the true residual of the returned solution meets the test
    b = rhs.ravel() - rhs.mean()
    r = b - oracle.stencil_apply(got, n, n, n, n)
    assert np.sqrt(np.mean(r * r)) <= 1.0001e-10 * np.sqrt(np.mean(b * b))
    # the reference's own smoother (lexicographic; bit-exact to the reference
    # in the oracle) reaches the same solution within the tolerance
    lex, cyc_l, _ = oracle.mg_solve(rhs, L, dim=3, tol=1e-10, tol_mode=1, smoother=0)
    dl = (got - got.mean()) - (lex.ravel() - lex.mean())
    assert np.max(np.abs(dl)) <= 1e-8 * np.max(np.abs(lex - lex.mean()))


def test_mg3_solves_the_mg_test3_fixture(ctx_of, dev):
    """test/mg_test3.cc's right-hand side (tests/golden/mg_test3.npz, with
    the reference's own solution at the absolute tolerance 1e-15): the GPU
    solve (red-black order) meets the tolerance and agrees to rounding."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "mg_test3.npz"))
    ctx = ctx_of(64)
    x = T(g["b"].reshape(16, 16, 16), dev)
    cyc, res = ctx.mg3_solve(x, tol=float(g["tol"]), tol_mode=0)
    got, want = H(x).ravel(), g["x"]
    assert res <= float(g["tol"]) and abs(cyc - int(g["cycles"])) <= 3
    d = (got - got.mean()) - (want - want.mean())
    assert np.max(np.abs(d)) <= 1e-12 * np.max(np.abs(want - want.mean()))


def test_mg3_rejects_bad_input(ctx_of, dev):
    from afmibins_amd._lib import AfmError
    ctx = ctx_of(64)
    with pytest.raises(TypeError):
        ctx.mg3_solve(torch.zeros((4, 4, 8), dtype=torch.float64, device=dev))
    with pytest.raises(ValueError):
        ctx.mg3_solve(torch.zeros((6, 6, 6), dtype=torch.float64, device=dev))
    # no cycle allowed and a tolerance the nested start cannot meet
    rhs = torch.rand((8, 8, 8), dtype=torch.float64, device=dev)
    with pytest.raises(AfmError):
        ctx.mg3_solve(rhs, tol=0.0, tol_mode=0, max_cycles=0)
