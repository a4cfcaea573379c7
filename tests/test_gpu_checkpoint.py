"""An IB run saved to a checkpoint file and restored continues bit for bit
as the uninterrupted run (the reference's FP64 iteration, and the mixed CG
form, whose projections start from potentials the checkpoint carries: per
step kind the next start and the last potential it is extrapolated from)."""
import numpy as np
import pytest
import torch

from afmibins_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [0, 1])
def test_checkpoint_restart_continues_identically(tmp_path, precision):
    import afmibins_amd.api as afm
    from afmibins_amd import formats as FM
    dev = "cuda:0"
    T = lambda a: torch.tensor(a, dtype=torch.float64, device=dev)
    n, mu, rho, dt, tol = 128, 1e-2, 1.0, 1e-4, 1e-10
    u0, v0 = W.fourier_velocity(n, 4, seed=6, rms=0.2)
    X, R, k, ds = W.membranes(4, 64, seed=4, center=(0.3, 0.6), axes=(0.05, 0.1))
    ctx = afm.Context(n, 0)
    u, v = T(u0), T(v0)
    ns = afm.NSSolver(ctx, mu, rho, dt, tol, u, v, tol_mode=1, precision=precision)
    ib = afm.IBSolver(ns, T(X), T(R), T(k), T(ds))
    try:
        ib.step(3)
        path = str(tmp_path / "run.npz")
        FM.save_checkpoint(path, u, v, ib.markers(), R, k, ds, n=n, mu=mu, rho=rho, dt=dt, tol=tol, tol_mode=1,
                           precision=precision, step=3, starts=[ns.get_start(j) for j in range(4)])
        ib.step(2)
        ctx.sync()
        want = (u.cpu().numpy(), v.cpu().numpy(), ib.markers().cpu().numpy())
    finally:
        ib.close(); ns.close(); ctx.close()

    ctx2, ns2, ib2, u2, v2, meta = FM.restore(path, dev)
    try:
        assert meta["step"] == 3
        ib2.step(2)
        ctx2.sync()
        got = (u2.cpu().numpy(), v2.cpu().numpy(), ib2.markers().cpu().numpy())
    finally:
        ib2.close(); ns2.close(); ctx2.close()
    for g, w, name in zip(got, want, ("u", "v", "X")):
        assert np.array_equal(g, w), name
