"""Differentiation form: adjoint gradients dJ/dk and dJ/dR on the GPU,
checked against central finite differences (h = 1e-6 along random
directions) and against the CPU oracle's FP64 adjoint on the same inputs."""
import numpy as np
import pytest
import torch

from afmibins_amd import workloads as W

pytestmark = pytest.mark.gpu


def T(a, dev):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)


def setup_case(afm, dev, n, count, msize, steps, every, tol=1e-13, seed=3, precision=0):
    from afmibins_amd.autograd import IBSetup
    ctx = afm.Context(n, 0)
    u0, v0 = W.fourier_velocity(n, 3, seed=seed, rms=0.3)
    X, R, k, ds = W.membranes(count, msize, seed=seed + 1, center=(0.3, 0.7), axes=(0.05, 0.12))
    k = k * (1.0 + 0.3 * np.random.default_rng(seed).standard_normal(k.shape))
    st = IBSetup(ctx, T(u0, dev), T(v0, dev), T(X, dev), T(ds, dev), 1e-2, 1.0, 1e-3, tol, steps, msize, every,
                 precision)
    return st, (u0, v0, X, R, k, ds)


def test_adjoint_matches_oracle_and_fd(afm, oracle, dev):
    from afmibins_amd.autograd import IBObjective
    n, steps, msize = 32, 3, 32
    st, (u0, v0, X, R, k, ds) = setup_case(afm, dev, n, 3, msize, steps, every=2)
    kt = T(k, dev).requires_grad_(True)
    Rt = T(R, dev).requires_grad_(True)
    J = IBObjective.apply(kt, Rt, st)
    J.backward()
    gk, gR = kt.grad.cpu().numpy(), Rt.grad.cpu().numpy()
    Jo, gko, gRo = oracle.ib_grad(u0, v0, X, R, k, ds, n, 1e-2, 1.0, 1e-3, 1e-13, steps, msize)
    assert abs(J.item() - Jo) <= 1e-11 * abs(Jo)
    assert np.max(np.abs(gk - gko)) <= 1e-9 * np.max(np.abs(gko))
    assert np.max(np.abs(gR - gRo)) <= 1e-9 * np.max(np.abs(gRo))
    rng = np.random.default_rng(11)
    h = 1e-6
    for _ in range(4):
        dk, dR = rng.standard_normal(k.shape), rng.standard_normal(R.shape)
        with torch.no_grad():
            Jp = IBObjective.apply(T(k + h * dk, dev), T(R + h * dR, dev), st).item()
            Jm = IBObjective.apply(T(k - h * dk, dev), T(R - h * dR, dev), st).item()
        fd = (Jp - Jm) / (2 * h)
        ad = float(gk @ dk + np.sum(gR * dR))
        assert abs(fd - ad) <= 1e-5 * abs(ad), (fd, ad)


@pytest.mark.parametrize("n,precision", [(64, 0), (64, 1), (128, 1)])
def test_adjoint_checkpointing_is_exact(afm, dev, n, precision):
    # checkpoint interval does not change the gradient (recomputation is
    # deterministic; the mixed CG form's checkpoints carry its projections'
    # start potentials, so a recomputed segment repeats the forward run)
    steps, msize = 6, 32
    res = []
    for every in (1, 4, 6):
        st, (u0, v0, X, R, k, ds) = setup_case(afm, dev, n, 2, msize, steps, every, tol=1e-12, precision=precision)
        from afmibins_amd.autograd import run
        J, gk, gR = run(st, T(k, dev), T(R, dev), True)
        res.append((J, gk.cpu().numpy(), gR.cpu().numpy()))
        st.ctx.close()
    for J, gk, gR in res[1:]:
        assert J == res[0][0]
        assert np.array_equal(gk, res[0][1]) and np.array_equal(gR, res[0][2])
