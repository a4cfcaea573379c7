"""The IB transfer on the reference's 3-D MAC grid (n^3 cells, periodic x and
z, walls in y) on the GPU, against the oracle's windows (which are bit-exact
to the reference's ib::grid::interpolate inside the reference's defined
domain, test_oracle_pins.py): interpolation and spreading bit for bit for u,
v (the mac_grid<y> window) and w, with markers anywhere -- next to the walls,
on the seams, on grid lines, crowded into one cell, repeated."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def T(a, dev):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)


def H(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def markers(n, seed, m=600):
    rng = np.random.default_rng(seed)
    X = rng.uniform(0, 1, (m, 3))
    X[:60, 1] = rng.uniform(0, 2.5 / n, 60)                 # next to the bottom wall
    X[60:120, 1] = rng.uniform(1 - 2.5 / n, 1, 60)          # next to the top wall
    X[120:160, 0] = rng.uniform(-1e-12, 1e-12, 40) % 1.0    # on the x seam
    X[160:200, 2] = 1 - rng.uniform(0, 1e-9, 40)            # just below the z seam
    g = rng.integers(0, 2 * n + 1, (100, 3)) / (2 * n)       # on cell faces and centres
    X[200:300] = np.minimum(g, 1.0)
    X[300:310, 1] = 0.0
    X[310:320, 1] = 1.0
    X[320:420] = np.array([0.51, 0.49, 0.52]) + rng.uniform(0, 0.4 / n, (100, 3))   # one crowded cell
    X[420:440] = X[421]                                      # one position, twenty markers
    return X


@pytest.mark.parametrize("n", [4, 16, 32])
@pytest.mark.parametrize("comp", [0, 1, 2])
def test_interp3_bit_exact(ctx_of, oracle, dev, n, comp):
    ctx = ctx_of(64)
    X = markers(n, 10 * n + comp)
    shape = (n, n - 1, n) if comp == 1 else (n, n, n)
    f = np.random.default_rng(n + comp).standard_normal(shape)
    U = H(ctx.interp3(comp, T(f, dev), T(X, dev)))
    assert np.array_equal(U, oracle.interp3(comp, f, X, n, mode=1))


@pytest.mark.parametrize("n", [4, 16, 32])
@pytest.mark.parametrize("comp", [0, 1, 2])
def test_spread3_bit_exact_and_reproducible(ctx_of, oracle, dev, n, comp):
    ctx = ctx_of(64)
    X = markers(n, 20 * n + comp)
    rng = np.random.default_rng(3 * n + comp)
    F, ds = rng.standard_normal(len(X)), rng.uniform(0.5, 1.5, len(X)) / n
    Xd, Fd, dsd = T(X, dev), T(F, dev), T(ds, dev)
    f1 = H(ctx.spread3(comp, Fd, dsd, Xd, n))
    assert np.array_equal(f1.ravel(), oracle.spread3(comp, F, ds, X, n, mode=1))
    f2 = H(ctx.spread3(comp, Fd, dsd, Xd, n))
    assert np.array_equal(f1, f2)


def test_ib3_empty_marker_sets_and_bad_input(ctx_of, dev):
    from afmibins_amd._lib import AfmError
    ctx = ctx_of(64)
    n = 8
    X = torch.empty((0, 3), dtype=torch.float64, device=dev)
    e = torch.empty(0, dtype=torch.float64, device=dev)
    for comp in (0, 1, 2, 3):
        f = H(ctx.spread3(comp, e, e, X, n))
        assert f.shape == ((n, n - 1, n) if comp == 1 else (n, n, n)) and not f.any()
        shape = (n, n - 1, n) if comp == 1 else (n, n, n)
        assert ctx.interp3(comp, torch.zeros(shape, dtype=torch.float64, device=dev), X).numel() == 0
    with pytest.raises(ValueError):
        ctx.interp3(1, torch.zeros((n, n, n), dtype=torch.float64, device=dev), X)
    with pytest.raises(AfmError):
        ctx.spread3(4, e, e, X, n)
