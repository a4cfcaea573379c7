"""Pins the oracle's 3-D wall-bounded Navier-Stokes step (the reference's
native geometry, n^3 cells, periodic x and z, walls in y) against the
reference's own ns::solver, helmholtz_dirichlet, div and grad (oracle/_ref),
bit for bit, and against properties the mathematics fixes."""
import numpy as np
import pytest

GEOS = {0: 4, 1: 5}   # ref_shim geometry id -> levels (16^3, 32^3)


def _p(a):
    from oracle.oracle import _ptr
    return _ptr(a)


def fields(n, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    u = scale * rng.standard_normal((n, n, n))
    v = scale * rng.standard_normal((n, n - 1, n))
    w = scale * rng.standard_normal((n, n, n))
    return u, v, w


@pytest.mark.parametrize("geo", [0, 1])
def test_dirichlet_helmholtz_csr_matches_reference(oracle, ref, geo):
    n = 1 << GEOS[geo]
    s = -3e-4
    nnz = ref.ref_csr(geo, 3, s, None, None, None)
    rs = np.empty(n * (n - 1) * n + 1, dtype=np.int64)
    col = np.empty(nnz, dtype=np.int64)
    val = np.empty(nnz)
    ref.ref_csr(geo, 3, s, _p(rs), _p(col), _p(val))
    a, b, c = oracle.stencil_csr(n, n - 1, n, n, 3, s)
    assert np.array_equal(rs, a) and np.array_equal(col, b) and np.array_equal(val, c)


def test_div_grad_3d_match_reference(oracle, ref):
    n = 16
    u, v, w = fields(n, 11)
    p = np.random.default_rng(12).standard_normal((n, n, n))
    d, px, pz = np.empty(n ** 3), np.empty(n ** 3), np.empty(n ** 3)
    py = np.empty(n * (n - 1) * n)
    ref.ref_div_grad(0, _p(u), _p(v), _p(w), _p(d), _p(p), _p(px), _p(py), _p(pz))
    assert np.array_equal(d, oracle.div3(u, v, w).ravel())
    a, b, c = oracle.grad3(p)
    assert np.array_equal(px, a.ravel()) and np.array_equal(py, b.ravel()) and np.array_equal(pz, c.ravel())


@pytest.mark.parametrize("geo,steps", [(0, 4), (1, 2)])
def test_ns3_steps_match_reference(oracle, ref, geo, steps):
    """ns::solver<geometry<L,1,1,1>>: constructor projection, alternating
    BE-half / CN-full steps with walls that move (wall vectors updated in
    place by projection::boundary), forces in all three components."""
    L = GEOS[geo]
    n = 1 << L
    u, v, w = fields(n, 20 + geo, 0.5)
    rng = np.random.default_rng(30 + geo)
    walls = [rng.uniform(-1, 1, (n, n)) for _ in range(4)]
    f = (rng.standard_normal((n, n, n)), rng.standard_normal((n, n - 1, n)), rng.standard_normal((n, n, n)))
    mu, rho, dt, tol = 0.05, 1.0, 1e-3, 1e-9
    uo, vo, wo, wo4, po, _ = oracle.ns3_run(L, u, v, w, walls, f, mu, rho, dt, tol, steps, smoother=0)
    ur, vr, wr = (np.ascontiguousarray(a).copy() for a in (u, v, w))
    wl = [np.ascontiguousarray(a).copy() for a in walls]
    fr = [np.ascontiguousarray(a) for a in f]
    pr = np.empty((n, n, n))
    assert ref.ref_ns3_run(geo, mu, rho, dt, tol, _p(ur), _p(vr), _p(wr), *(_p(a) for a in wl),
                           *(_p(a) for a in fr), steps, _p(pr)) == 0
    assert np.array_equal(uo, ur) and np.array_equal(vo, vr) and np.array_equal(wo, wr)
    for a, b in zip(wo4, wl):
        assert np.array_equal(a, b)
    assert np.array_equal(po, pr)


def test_mg_test3_fixture(oracle):
    """test/mg_test3.cc's right-hand side, solved by the reference's
    mg::solver<geometry<4,1,1,1>> (tests/golden/make_mg_test3.py): the
    oracle's lexicographic solve gives the same cycles and bits."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "mg_test3.npz"))
    x, c, _ = oracle.mg_solve(g["b"].reshape(16, 16, 16), 4, dim=3, tol=float(g["tol"]), tol_mode=0, smoother=0)
    assert c == int(g["cycles"])
    assert np.array_equal(x.ravel(), g["x"])


# ------------------------------------------------------------ properties
def test_spread3_is_adjoint_of_interp3_up_to_the_walls(oracle):
    """<S F, u>_h = <F, J u>_ds for all three MAC components with markers
    anywhere, including within 2h of the walls (where the v window's wall
    faces are skipped both ways) and of the periodic seams; momentum
    sum f h^3 = sum F ds for markers at least 3h from walls and seams."""
    n = 16
    rng = np.random.default_rng(8)
    X = rng.uniform(0, 1, (60, 3))
    X[:10, 1] = rng.uniform(0, 2.0 / n, 10)
    X[10:20, 1] = rng.uniform(1 - 2.0 / n, 1, 10)
    X[30:, :] = rng.uniform(3.0 / n, 1 - 3.0 / n, (30, 3))
    F, ds = rng.standard_normal(60), rng.uniform(0.5, 1.5, 60) / n
    for comp in (0, 1, 2):
        shape = (n, n - 1, n) if comp == 1 else (n, n, n)
        u = rng.standard_normal(shape)
        f = oracle.spread3(comp, F, ds, X, n, mode=1)
        J = oracle.interp3(comp, u, X, n, mode=1)
        lhs = np.sum(f * u.ravel()) / n ** 3
        rhs = np.sum(F * ds * J)
        assert abs(lhs - rhs) <= 1e-13 * (abs(lhs) + abs(rhs))
        fi = oracle.spread3(comp, F[30:], ds[30:], X[30:], n, mode=1)
        assert abs(fi.sum() / n ** 3 - np.sum(F[30:] * ds[30:])) < 1e-13
    # a v window next to a wall loses exactly its wall-face taps: constant v
    # interpolates to the stored-face share of the unit partition
    Xw = np.array([[0.43, 0.2 / n, 0.61], [0.43, 0.5, 0.61], [0.43, 1 - 0.7 / n, 0.61]])
    J = oracle.interp3(1, np.ones((n, n - 1, n)), Xw, n, mode=1)
    assert abs(J[1] - 1.0) < 1e-12 and J[0] < 0.9 and J[2] < 0.9


# ------------------------------------------------------- properties (NS)
def test_dirichlet_helmholtz_is_symmetric_and_boundary_is_its_wall_column(oracle):
    """The Dirichlet rows are I + s*L7 on the interior faces with the wall
    faces eliminated: symmetric, and boundary() is the eliminated column
    s*n^2 times the wall value."""
    n, s = 8, -2e-3
    rs, col, val = oracle.stencil_csr(n, n - 1, n, n, 3, s)
    N = n * (n - 1) * n
    A = np.zeros((N, N))
    for i in range(N):
        for k in range(rs[i], rs[i + 1]):
            A[i, col[k]] += val[k]
    assert np.array_equal(A, A.T)
    # a field that is constant in x and z and linear in y across the full
    # face column (walls at y = 0 and y = n): the interior rows of
    # (I + s*Lap) applied to it, walls moved to the right-hand side, are the
    # field itself (the Laplacian of a linear function is 0)
    yb, yt = 0.3, 1.1
    yf = np.arange(1, n) / n
    vfield = np.broadcast_to((yb + (yt - yb) * yf)[None, :, None], (n, n - 1, n)).copy()
    top, bot = np.full((n, n), yt), np.full((n, n), yb)
    b = oracle.hbound3(3, n, s, top, bot)
    Av = A @ vfield.ravel() + b.ravel()
    assert np.allclose(Av, vfield.ravel(), rtol=0, atol=1e-13)


def test_periodic_helmholtz_boundary_closure_is_second_order(oracle):
    """helmholtz_periodic's wall rows (ghost value 8/3 g - 2 u0 + u1/3, a
    quadratic through the wall value and the first two cells): a quadratic
    in y with wall values g is reproduced, so (I + s*L) q + boundary = q + s*q''."""
    n, s = 8, -1e-3
    rs, col, val = oracle.stencil_csr(n, n, n, n, 1, s)
    y = (np.arange(n) + 0.5) / n
    qy = 0.2 + 0.7 * y - 0.9 * y ** 2
    q = np.broadcast_to(qy[None, :, None], (n, n, n)).ravel().copy()
    top, bot = np.full((n, n), 0.2 + 0.7 - 0.9), np.full((n, n), 0.2)
    Aq = np.array([sum(val[k] * q[col[k]] for k in range(rs[i], rs[i + 1])) for i in range(n ** 3)])
    b = oracle.hbound3(1, n, s, top, bot).ravel()
    assert np.allclose(Aq + b, q + s * (-1.8), rtol=0, atol=1e-12)


def test_advection_of_uniform_flow_is_zero_and_projection_is_div_free(oracle):
    n = 8
    u = np.full((n, n, n), 0.7)
    v = np.zeros((n, n - 1, n))
    w = np.full((n, n, n), -0.4)
    a, b, c = oracle.advect3(u, v, w)
    assert np.max(np.abs(a)) == 0 and np.max(np.abs(b)) == 0 and np.max(np.abs(c)) == 0
    # the constructor's projection leaves a discretely divergence-free field
    u, v, w = fields(n, 40)
    z = [np.zeros((n, n))] * 4
    f = (np.zeros((n, n, n)), np.zeros((n, n - 1, n)), np.zeros((n, n, n)))
    uo, vo, wo, _, _, _ = oracle.ns3_run(3, u, v, w, z, f, 0.01, 1.0, 1e-3, 1e-12, 0)
    d = oracle.div3(uo, vo, wo)
    assert np.sqrt(np.mean(d ** 2)) < 1e-9
