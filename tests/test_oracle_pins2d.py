"""Pins of the oracle's 2-D doubly periodic restatements, which the reference
cannot run on that plane (its geometry has walls in y):

* advection: the reference's own ns::solver<geometry<4,1,1,1>>::advection
  (ns/solver.h:142-236), called through oracle/_ref, is bit-equal to the
  oracle's 3-D advection on random data, and on y-independent u, w with
  v = 0 every one of its y-planes is bit-equal to the oracle's 2-D advection;
* interpolation and spreading on the plane: a brute-force sum of the kernel
  over every face of the grid at the documented staggering (u at
  (i, j+1/2)/n, v at (i+1/2, j)/n), with no windows and no index arithmetic;
* the 2-D Navier-Stokes step (ns/solver.h:90-140 on the plane): a shear mode,
  for which advection and projection vanish exactly and each step is a
  closed form in the Helmholtz eigenvalue, and a manufactured unsteady
  solution whose error falls as h^2 + dt^2 (advection, viscosity, forcing and
  the backward-Euler / Crank-Nicolson pairing all enter it).
"""
import numpy as np
import pytest


def _p(a):
    from oracle.oracle import _ptr
    return _ptr(a)


# ---------------------------------------------------------------- advection
def _ref_advect(ref, u, v, w):
    n = 16
    u, v, w = (np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, w))
    ua, wa = np.empty((n, n, n)), np.empty((n, n, n))
    va = np.empty((n, n - 1, n))
    assert ref.ref_advect(_p(u), _p(v), _p(w), _p(ua), _p(va), _p(wa)) == 0
    return ua, va, wa


def test_advect3_matches_reference(oracle, ref):
    n = 16
    rng = np.random.default_rng(61)
    u, w = rng.standard_normal((n, n, n)), rng.standard_normal((n, n, n))
    v = rng.standard_normal((n, n - 1, n))
    ra, rb, rc = _ref_advect(ref, u, v, w)
    oa, ob, oc = oracle.advect3(u, v, w)
    assert np.array_equal(ra, oa) and np.array_equal(rb, ob) and np.array_equal(rc, oc)


def test_advect2_matches_reference_on_y_independent_flow(oracle, ref):
    # The reference's plane (x, z) is the build's (x, y).  With u, w constant
    # in the reference's y and v = 0, the y-fluxes u v and v w are zero and
    # every y-plane of its advection is the doubly periodic one.
    n = 16
    rng = np.random.default_rng(62)
    u2, w2 = rng.standard_normal((n, n)), rng.standard_normal((n, n))    # [z, x]
    u3 = np.repeat(u2[:, None, :], n, axis=1)
    w3 = np.repeat(w2[:, None, :], n, axis=1)
    ra, _, rc = _ref_advect(ref, u3, np.zeros((n, n - 1, n)), w3)
    a2, b2 = oracle.advect2(u2, w2, n)
    for y in range(n):
        assert np.array_equal(ra[:, y, :], a2), y
        assert np.array_equal(rc[:, y, :], b2), y


# ------------------------------------------------- interpolation, spreading
def _phi(r, n):
    """The 4-point cosine kernel (ib/grid.h:51-55) on its window |n r| < 2."""
    t = n * r
    return np.where(np.abs(t) < 2, 0.25 * n * (1 + np.cos(np.pi * t / 2)), 0.0)


def _faces(comp, n):
    x = np.arange(n) / n + (0.5 / n if comp == 1 else 0.0)
    y = np.arange(n) / n + (0.5 / n if comp == 0 else 0.0)
    return x, y


@pytest.mark.parametrize("n", [16, 32])
def test_interp2_equals_brute_force_sum_over_all_faces(oracle, n):
    rng = np.random.default_rng(63 + n)
    X = rng.uniform(3.0 / n, 1 - 3.0 / n, (60, 2))      # away from the seams
    X[:4] = [[0.5, 0.5], [0.5 + 0.5 / n, 0.5], [0.5, 0.5 + 0.5 / n], [0.25 + 1e-9, 0.75 - 1e-9]]
    for comp in (0, 1):
        f = rng.standard_normal((n, n))                  # [y, x]
        xs, ys = _faces(comp, n)
        want = np.array([_phi(ys - Y, n) @ f @ _phi(xs - Xx, n) / n ** 2 for Xx, Y in X])
        got = oracle.interp2(comp, f, X, n, mode=1)
        assert np.allclose(got, want, rtol=0, atol=1e-13 * np.abs(f).max()), comp
        # the other staggering is a different function: the check is sharp
        xo, yo = _faces(1 - comp, n)
        other = np.array([_phi(yo - Y, n) @ f @ _phi(xo - Xx, n) / n ** 2 for Xx, Y in X])
        assert np.max(np.abs(other - got)) > 1e-3 * np.abs(f).max()


@pytest.mark.parametrize("n", [16, 32])
def test_spread2_equals_brute_force_sum_over_all_markers(oracle, n):
    rng = np.random.default_rng(64 + n)
    m = 50
    X = rng.uniform(3.0 / n, 1 - 3.0 / n, (m, 2))
    F, ds = rng.standard_normal(m), rng.uniform(0.5, 1.5, m) / n
    for comp in (0, 1):
        xs, ys = _faces(comp, n)
        want = np.zeros((n, n))
        for (Xx, Y), Fk, dk in zip(X, F, ds):
            want += np.outer(_phi(ys - Y, n), _phi(xs - Xx, n)) * Fk * dk
        got = oracle.spread2(comp, F, ds, X, n, n, n, mode=1)
        assert np.allclose(got, want, rtol=0, atol=1e-13 * np.abs(want).max()), comp


# ------------------------------------------------------ 2-D NS step (plane)
@pytest.mark.parametrize("comp", [0, 1])
def test_ns2_step_on_a_shear_mode_is_the_closed_form(oracle, comp):
    """u = a sin(2 pi k y) (or v = a sin(2 pi k x)) with a force of the same
    mode: advection and divergence vanish exactly, so the backward-Euler half
    step and the Crank-Nicolson full step are, with lam = -4 n^2 sin^2(pi k/n)
    and s = mu dt / (2 rho),
        u1 = (u0 + dt/2 F / rho) / (1 - s lam),
        u2 = ((1 + s lam) u0 + dt F / rho) / (1 - s lam)."""
    n, kk, amp, F0 = 32, 3, 0.8, -2.5
    mu, rho, dt = 0.05, 2.0, 1e-3
    c = (np.arange(n) + 0.5) / n
    mode = np.sin(2 * np.pi * kk * c)
    shape = mode[:, None] * np.ones((1, n)) if comp == 0 else np.ones((n, 1)) * mode[None, :]
    z = np.zeros((n, n))
    u0, v0 = (amp * shape, z) if comp == 0 else (z, amp * shape)
    fx, fy = (F0 * shape, z) if comp == 0 else (z, F0 * shape)
    lam = -4 * n * n * np.sin(np.pi * kk / n) ** 2
    s = mu * dt / (2 * rho)
    u1, v1, _, _ = oracle.ns_run(u0, v0, fx, fy, n, mu, rho, dt, 1e-14, calls=1)
    u2, v2, _, _ = oracle.ns_run(u0, v0, fx, fy, n, mu, rho, dt, 1e-14, calls=2)
    w1 = (amp + dt / 2 * F0 / rho) / (1 - s * lam) * shape
    w2 = ((1 + s * lam) * amp + dt * F0 / rho) / (1 - s * lam) * shape
    g1, g2 = (u1, u2) if comp == 0 else (v1, v2)
    o1, o2 = (v1, v2) if comp == 0 else (u1, u2)
    assert np.max(np.abs(g1 - w1)) <= 1e-13 and np.max(np.abs(g2 - w2)) <= 1e-13
    # the other component stays zero up to rounding (the CSR rows at the
    # seam sum in another order, so u is x-constant only to rounding)
    assert np.abs(o1).max() <= 1e-15 and np.abs(o2).max() <= 1e-15


_K = 2 * np.pi
_A = 0.5


def _g(t):
    return 1.0 + 0.8 * np.sin(6.0 * t)


def _dg(t):
    return 4.8 * np.cos(6.0 * t)


def _mms_velocity(x, y, t):
    """u = d psi / dy, v = -d psi / dx for
    psi = g(t)/(2 pi) [sin(2 pi x) sin(2 pi y) + a cos(4 pi x + 1) sin(2 pi y + 0.3)]."""
    g = _g(t)
    u = g * (np.sin(_K * x) * np.cos(_K * y) + _A * np.cos(2 * _K * x + 1) * np.cos(_K * y + 0.3))
    v = g * (-np.cos(_K * x) * np.sin(_K * y) + 2 * _A * np.sin(2 * _K * x + 1) * np.sin(_K * y + 0.3))
    return u, v


def _mms_force(x, y, t, nu, rho, comp):
    """rho (du/dt + (u.grad) u - nu lap u) at (x, y): the force for which the
    velocity above solves the Navier-Stokes equations with zero pressure."""
    g, k, a = _g(t), _K, _A
    u, v = _mms_velocity(x, y, t)
    sx, cx, sy, cy = np.sin(k * x), np.cos(k * x), np.sin(k * y), np.cos(k * y)
    s2, c2 = np.sin(2 * k * x + 1), np.cos(2 * k * x + 1)
    s3, c3 = np.sin(k * y + 0.3), np.cos(k * y + 0.3)
    if comp == 0:
        ux = g * (k * cx * cy - 2 * a * k * s2 * c3)
        uy = g * (-k * sx * sy - a * k * c2 * s3)
        lu = g * (-2 * k * k * sx * cy - 5 * a * k * k * c2 * c3)
        return rho * (_dg(t) / g * u + u * ux + v * uy - nu * lu)
    vx = g * (k * sx * sy + 4 * a * k * c2 * s3)
    vy = g * (-k * cx * cy + 2 * a * k * s2 * c3)
    lv = g * (2 * k * k * cx * sy - 10 * a * k * k * s2 * s3)
    return rho * (_dg(t) / g * v + u * vx + v * vy - nu * lv)


def _mms_error(oracle, n, T, mu, rho):
    dt = T / (n * 25 / 16)           # dt proportional to h
    steps = int(round(T / dt))
    nu = mu / rho
    i = np.arange(n) / n
    ih = (np.arange(n) + 0.5) / n
    xu, yu = np.meshgrid(i, ih)      # x-faces [y, x]
    xv, yv = np.meshgrid(ih, i)      # y-faces
    # discrete curl of psi sampled at the cell corners: exactly divergence-free
    xc, yc = np.meshgrid(i, i)
    psi = _g(0.0) / _K * (np.sin(_K * xc) * np.sin(_K * yc) + _A * np.cos(2 * _K * xc + 1) * np.sin(_K * yc + 0.3))
    u = n * (np.roll(psi, -1, 0) - psi)
    v = -n * (np.roll(psi, -1, 1) - psi)
    for st in range(steps):
        tm = (st + 0.5) * dt         # forcing at the step's midpoint
        fx = _mms_force(xu, yu, tm, nu, rho, 0)
        fy = _mms_force(xv, yv, tm, nu, rho, 1)
        u, v, _, _ = oracle.ns_run(u, v, fx, fy, n, mu, rho, dt, 1e-13, calls=2)
    ue, _ = _mms_velocity(xu, yu, steps * dt)
    _, ve = _mms_velocity(xv, yv, steps * dt)
    return max(np.abs(u - ue).max(), np.abs(v - ve).max()) / max(np.abs(ue).max(), np.abs(ve).max())


def test_ns2_steps_converge_to_a_manufactured_solution(oracle):
    """Second order in h and dt together (dt ~ h): the error ratio between
    successive refinements is near 4.  A wrong sign or factor of the
    advection, viscosity or force terms leaves an O(1) error that does not
    fall, and a first-order step pairing gives a ratio near 2."""
    mu, rho, T = 0.04, 2.0, 0.16
    errs = [_mms_error(oracle, n, T, mu, rho) for n in (16, 32, 64)]
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert errs[2] < 5e-3, errs
    assert 3.4 < r1 < 4.6 and 3.4 < r2 < 4.6, (errs, r1, r2)


@pytest.mark.parametrize("n", [16, 64])
def test_kd_tree_gather_spreading_equals_the_scatter(oracle, n):
    """The reference spreads by gathering, for every face, the markers in a
    kd-tree box of half-width 2/n (dt/kdtree.h:39-146); the oracle's scatter
    over each marker's window sums the same terms in another order.  Markers
    at least 3h from the seams, as the reference's own tests place them (near
    a seam the reference's gather uses periodic images while its window uses
    the unwrapped coordinate)."""
    rng = np.random.default_rng(70 + n)
    m = 400
    X = rng.uniform(3.0 / n, 1 - 3.0 / n, (m, 2))
    X[:50] = 0.5 + rng.uniform(-0.5, 0.5, (50, 2)) / n          # a crowded cell
    F, ds = rng.standard_normal(m), rng.uniform(0.5, 1.5, m) / n
    for comp in (0, 1):
        a = oracle.spread2(comp, F, ds, X, n, n, n, mode=1)
        b = oracle.spread2_gather(comp, F, ds, X, n, n, n, mode=1)
        assert np.array_equal(a != 0, b != 0)
        assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(a)), comp
