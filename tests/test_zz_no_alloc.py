"""The allocation-free Helmholtz CG (afm_helmholtz_cg_noalloc), called through
the C ABI, against afm_helmholtz_cg: the same iterations, residual and
solution bits, on two calls in a row (the second on a smaller grid, so the
shared per-GPU buffer holds the first call's larger vectors), and a clear
refusal of a grid beyond its limit."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _noalloc(ctx, scale, b, x, tol, tol_mode=1, max_iters=10000):
    from afmibins_amd import _lib
    it, res = C.c_int(-1), C.c_double(-1.0)
    ctx._bind_stream()
    rc = _lib.load().afm_helmholtz_cg_noalloc(ctx._h, scale, C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()),
                                              tol, tol_mode, max_iters, C.byref(it), C.byref(res))
    torch.cuda.synchronize()
    return rc, it.value, res.value


def test_noalloc_cg_matches_the_allocating_cg(afm, dev):
    for n, seed in ((1024, 1), (128, 2)):
        print(f"noalloc CG: n={n}", flush=True)
        ctx = afm.Context(n, 0)
        rng = np.random.default_rng(seed)
        b = torch.tensor(rng.standard_normal((n, n)), dtype=torch.float64, device=dev)
        x0 = 0.25 * b
        scale = -0.05 / (n * n)          # s n^2 = 0.05: a few iterations
        xa = x0.clone()
        it_a, res_a = ctx.helmholtz_cg(scale, b, xa, tol=1e-13, tol_mode=1)
        xb = x0.clone()
        rc, it_b, res_b = _noalloc(ctx, scale, b, xb, 1e-13)
        assert rc == 0
        assert it_b == it_a and res_b == res_a, (it_a, it_b, res_a, res_b)
        assert torch.equal(xa, xb)
        assert it_a >= 2
        ctx.close()


def test_noalloc_cg_refuses_a_grid_beyond_its_buffer(afm, dev):
    from afmibins_amd import _lib
    n = 4096
    ctx = afm.Context(n, 0)
    b = torch.ones((n, n), dtype=torch.float64, device=dev)
    x = torch.zeros_like(b)
    rc, _, _ = _noalloc(ctx, -1e-8, b, x, 1e-10)
    assert rc == -1
    assert b"2048" in _lib.load().afm_last_error()
    assert not x.any()
    ctx.close()
