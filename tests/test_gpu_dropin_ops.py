"""The GPU operations behind the drop-in headers, against the oracle:

* the 3-D multigrid on cubes of c 2^L cells per axis (geometry<L,c,c,c>:
  <1,8,8,8> is test/mg_test2.cc's), its laplacian, restriction,
  prolongation and red-black sweeps bit-exact, V-cycles and solves to the
  oracle's rounding (its coarsest c^3 level is a CG whose dot products sum
  in another order) and within the tolerance of the reference-pinned
  lexicographic solve;
* alg::conjugate_gradient on the laplacian and on a dt::matrix held in CSR,
  its SpMV and lexicographic sweep bit-exact;
* helmholtz_*::boundary and projection::boundary bit-exact;
* interpolation and spreading on the node grid eulerian_grid (comp 3)
  bit-exact."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def T(a, dev, dtype=torch.float64):
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device=dev)


def H(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


CUBES = [(1, 8), (2, 2), (2, 4), (3, 1), (1, 1)]


@pytest.mark.parametrize("L,W", CUBES)
def test_laplacian_restrict_prolong_smooth3_bit_exact(ctx_of, oracle, dev, L, W):
    s, n = W << L, 1 << L
    ctx = ctx_of(64)
    rng = np.random.default_rng(10 * L + W)
    x, b = rng.standard_normal((s, s, s)), rng.standard_normal((s, s, s))
    assert np.array_equal(H(ctx.laplacian3(T(x, dev), width=W)).ravel(), oracle.stencil_apply(x, s, s, s, n))
    assert np.array_equal(H(ctx.mg3_restrict(T(x, dev), width=W)).ravel(), oracle.restrict3(x, s, s, s))
    c = rng.standard_normal((s // 2, s // 2, s // 2))
    want = oracle.prolong3(c, np.zeros(s ** 3), s, s, s)
    assert np.array_equal(H(ctx.mg3_prolong(T(c, dev), width=W)).ravel(), want)
    got = H(ctx.mg3_smooth(T(x, dev), T(b, dev), sweeps=2, width=W)).ravel()
    assert np.array_equal(got, oracle.smooth(x, b, s, s, s, n, smoother=1, sweeps=2))


@pytest.mark.parametrize("L,W", CUBES)
def test_vcycle3_on_cube_multiples(ctx_of, oracle, dev, L, W):
    s = W << L
    ctx = ctx_of(64)
    r = np.random.default_rng(20 * L + W).standard_normal((s, s, s))
    got = H(ctx.mg3_vcycle(T(r, dev), width=W, tol=1e-13)).ravel()
    want = oracle.mg_vcycle(r, L, dim=3, W=W, H=W, D=W, tol=1e-13, smoother=1).ravel()
    if W == 1:
        assert np.array_equal(got, want)
    else:
        assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))


def test_mg3_solve_of_mg_test2(ctx_of, oracle, dev):
    """test/mg_test2.cc: geometry<1,8,8,8> (16^3 cells, n = 2, coarsest 8^3
    by CG), ones with b[0] = 2.  Against the oracle's red-black solve (to its
    rounding) and its lexicographic solve (bit-exact to the reference build,
    test_oracle_pins.py), and L x = b - mean(b) (the program's intent)."""
    L, W, s, tol = 1, 8, 16, 1e-10
    b = np.ones(s ** 3); b[0] = 2.0
    ctx = ctx_of(64)
    x = T(b.reshape(s, s, s), dev)
    cyc, rms = ctx.mg3_solve(x, tol=tol, tol_mode=0, width=W)
    got = H(x).ravel()
    rb, cyc_rb, _ = oracle.mg_solve(b, L, dim=3, W=W, H=W, D=W, tol=tol, smoother=1)
    lex, cyc_lex, _ = oracle.mg_solve(b, L, dim=3, W=W, H=W, D=W, tol=tol, smoother=0)
    print(f"geometry<1,8,8,8>: cycles GPU {cyc}, red-black oracle {cyc_rb}, lexicographic {cyc_lex}", flush=True)
    assert abs(cyc - cyc_rb) <= 1
    sc = np.max(np.abs(rb - rb.mean()))
    assert np.max(np.abs((got - got.mean()) - (rb - rb.mean()))) <= 1e-12 * sc
    assert np.max(np.abs((got - got.mean()) - (lex - lex.mean()))) <= 1e-8 * sc
    r = (b - b.mean()) - oracle.stencil_apply(got, s, s, s, 2)
    assert np.sqrt(np.mean(r * r)) <= tol * 1.0001


@pytest.mark.parametrize("L,W", [(3, 1), (1, 8)])
def test_laplacian3_cg_matches_oracle(ctx_of, oracle, dev, L, W):
    s, n = W << L, 1 << L
    ctx = ctx_of(64)
    b = np.random.default_rng(5 + L).standard_normal(s ** 3)
    b -= b.mean()
    x = T(np.zeros((s, s, s)), dev)
    it, nrm = ctx.laplacian3_cg(T(b.reshape(s, s, s), dev), x, tol=1e-10, width=W)
    xo, ito, _ = oracle.cg(np.zeros(s ** 3), b, s, s, s, n, kind=0, tol=1e-10)
    assert abs(it - ito) <= 1
    assert np.max(np.abs(H(x).ravel() - xo)) <= 1e-9 * np.max(np.abs(xo))
    assert nrm <= 1e-10


@pytest.mark.parametrize("kind", [1, 3])
def test_helmholtz3_boundary_bit_exact(ctx_of, oracle, dev, kind):
    n, scale = 16, -2.5e-4
    ctx = ctx_of(64)
    rng = np.random.default_rng(kind)
    top, bot = rng.standard_normal((n, n)), rng.standard_normal((n, n))
    got = H(ctx.helmholtz3_boundary(kind, scale, T(top, dev), T(bot, dev)))
    assert np.array_equal(got.ravel(), oracle.hbound3(kind, n, scale, top, bot).ravel())


def test_projection3_boundary_bit_exact(ctx_of, oracle, dev):
    n = 16
    ctx = ctx_of(64)
    rng = np.random.default_rng(4)
    phx, phz = rng.standard_normal((n, n, n)), rng.standard_normal((n, n, n))
    walls = [rng.standard_normal((n, n)) for _ in range(4)]
    wt = [T(w, dev) for w in walls]
    ctx.projection3_boundary(T(phx, dev), T(phz, dev), *wt)
    want = oracle.proj_boundary3(phx, phz, *walls)
    for g, w in zip(wt, want):
        assert np.array_equal(H(g), w)


def test_csr_spmv_cg_and_lexicographic_sweep(ctx_of, oracle, dev):
    # helmholtz_periodic's CSR on 16^3 as a matrix a program fills itself
    n, scale = 16, -3e-4
    ctx = ctx_of(64)
    rs, col, val = oracle.stencil_csr(n, n, n, n, 1, scale)
    rsd, cold, vald = T(rs, dev, torch.int64), T(col, dev, torch.int64), T(val, dev)
    rng = np.random.default_rng(6)
    x, b = rng.standard_normal(n ** 3), rng.standard_normal(n ** 3)
    assert np.array_equal(H(ctx.csr_spmv(rsd, cold, vald, T(x, dev))), oracle.stencil_apply(x, n, n, n, n, 1, scale))
    xg = T(np.zeros(n ** 3), dev)
    it, nrm = ctx.csr_cg(rsd, cold, vald, T(b, dev), xg, tol=1e-12)
    xo, ito, _ = oracle.cg(np.zeros(n ** 3), b, n, n, n, n, kind=1, scale=scale, tol=1e-12)
    assert abs(it - ito) <= 1 and nrm <= 1e-12
    assert np.max(np.abs(H(xg) - xo)) <= 1e-12 * np.max(np.abs(xo))
    y = T(b, dev)
    ctx.axpy(0.5, T(x, dev), y)
    assert np.array_equal(H(y), b + 0.5 * x)


def test_csr_lexicographic_sweep_is_the_reference_smoother(ctx_of, oracle, dev):
    # the laplacian's CSR: alg::smooth on it is the oracle's lexicographic
    # sweep, which is bit-exact to the reference build
    n = 16
    ctx = ctx_of(64)
    rs, col, val = oracle.stencil_csr(n, n, n, n, 0, 0.0)
    rng = np.random.default_rng(7)
    x, b = rng.standard_normal(n ** 3), rng.standard_normal(n ** 3)
    got = H(ctx.csr_smooth(T(rs, dev, torch.int64), T(col, dev, torch.int64), T(val, dev), T(x, dev), T(b, dev),
                           sweeps=3))
    assert np.array_equal(got, oracle.smooth(x, b, n, n, n, n, smoother=0, sweeps=3))


@pytest.mark.parametrize("n", [16, 32])
def test_node_grid_transfer3_bit_exact(ctx_of, oracle, dev, n):
    ctx = ctx_of(64)
    rng = np.random.default_rng(n + 9)
    X = rng.uniform(3.5 / n, 1 - 3.5 / n, (200, 3))
    f = rng.standard_normal((n, n, n))
    got = H(ctx.interp3(3, T(f, dev), T(X, dev)))
    assert np.array_equal(got, oracle.interp3(3, f, X, n, mode=1))
    F, ds = rng.standard_normal(200), rng.uniform(0.5, 1.5, 200) / n
    got = H(ctx.spread3(3, T(F, dev), T(ds, dev), T(X, dev), n))
    assert np.array_equal(got.ravel(), oracle.spread3(3, F, ds, X, n, mode=1).ravel())
