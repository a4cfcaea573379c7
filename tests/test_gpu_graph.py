"""HIP graph capture of the entry points that do not synchronise the host
(a V-cycle, spreading, interpolation): captured on the stream the context is
bound to, replayed several times with the program's own eager work and
synchronisations between the replays, every replay bit-identical to the
eager result.  Spreading with no markers clears its bins with a kernel, so
the graph holds no memset node.  Solves synchronise the host for their stop
tests and are not capturable."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_vcycle_spread_interp_replay_bit_identical(afm, dev):
    n, m = 256, 3000
    ctx = afm.Context(n, 0)     # its own context: it stays bound to the capture stream
    rng = np.random.default_rng(5)
    T = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    rhs0 = rng.standard_normal((n, n))
    rhs0 = T(rhs0 - rhs0.mean())
    X = T(rng.uniform(0.2, 0.8, (m, 2)))
    F, ds = T(rng.standard_normal(m)), T(rng.uniform(0.5, 1.5, m) / m)
    field = T(rng.standard_normal((n, n)))
    other = T(rng.standard_normal((n, n)))
    s = torch.cuda.Stream(dev)
    s.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(s):
        # eager results; this also binds the context to s before the capture
        r = rhs0.clone()
        ctx.mg_vcycle(r)
        want_v = r.clone()
        want_s = ctx.spread(0, F, ds, X).clone()
        want_i = ctx.interp(1, field, X).clone()
        torch.cuda.synchronize(dev)
        rg = rhs0.clone()
        out = torch.empty((n, n), dtype=torch.float64, device=dev)
        out0 = torch.empty((n, n), dtype=torch.float64, device=dev)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            ctx.mg_vcycle(rg)
            ctx.spread(0, F, ds, X, out=out)
            ctx.spread(0, F[:0], ds[:0], X[:0], out=out0)
            Ug = ctx.interp(1, field, X)
        for rep in range(4):
            rg.copy_(rhs0)
            out.fill_(7.0)
            out0.fill_(7.0)
            Ug.fill_(7.0)
            o = other.clone()
            ctx.mg_vcycle(o)                  # eager work on the same context between replays
            ctx.spread(1, F, ds, X)
            torch.cuda.synchronize(dev)
            g.replay()
            torch.cuda.synchronize(dev)
            assert torch.equal(rg, want_v), rep
            assert torch.equal(out, want_s), rep
            assert torch.equal(out0, torch.zeros_like(out0)), rep
            assert torch.equal(Ug, want_i), rep
        torch.cuda.synchronize(dev)
        del g
    del ctx
