"""Python mirror of the reference's ib / mg / ns interfaces over libafm.

Device memory is plain torch tensors (float64, contiguous, on the context's
GPU); every call goes through the C ABI of include/afm.h and is ordered on
torch's current stream.  Names follow the reference:

  Context.interp / Context.spread   ib::grid::interpolate (ib/grid.h:68-84)
  MGSolver(ctx, tol, nu1, nu2).solve(rhs)
                                    mg::solver<G>::solve (mg/solver.h:218-237)
  NSSolver(ctx, mu, rho, dt, tol, u0, v0).step(fx, fy)
                                    ns::solver<G>::step (ns/solver.h:90-140)
  IBSolver(ns, X, R, k, ds).step(n) the build's IB time step
"""
import ctypes as C

import torch

from . import _lib
from ._lib import check


def _ptr(t):
    return C.c_void_p(t.data_ptr())


class Context:
    """A 2-D doubly periodic unit-square MAC grid with n x n cells."""

    def __init__(self, n, device=0):
        L = _lib.load()
        self.device = torch.device("cuda", device)
        g = _lib.Geom(2, n, 1, 1)
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            check(L.afm_ctx_create(C.byref(g), device, C.byref(h)), "afm_ctx_create")
        self._h = h
        self.n = n
        nx, ny, lv, lc = C.c_uint(), C.c_uint(), C.c_uint(), C.c_uint()
        check(L.afm_ctx_info(h, C.byref(nx), C.byref(ny), C.byref(lv), C.byref(lc)), "afm_ctx_info")
        self.nx, self.ny, self.levels, self.coarse_level = nx.value, ny.value, lv.value, lc.value
        self._bind_stream()

    # -- plumbing ---------------------------------------------------------
    def _bind_stream(self):
        s = torch.cuda.current_stream(self.device)
        check(_lib.load().afm_set_stream(self._h, C.c_void_p(s.cuda_stream)), "afm_set_stream")
        self._stream = s.cuda_stream

    def _call(self, name, *args):
        s = torch.cuda.current_stream(self.device).cuda_stream
        if s != self._stream:
            self._bind_stream()
        return check(getattr(_lib.load(), name)(self._h, *args), name)

    def _f(self, t, shape=None, name="tensor"):
        if not isinstance(t, torch.Tensor) or t.dtype != torch.float64 or t.device != self.device \
                or not t.is_contiguous():
            raise TypeError(f"{name} must be a contiguous float64 tensor on {self.device}")
        if shape is not None and tuple(t.shape) != tuple(shape):
            raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")
        return t

    def field(self, level=None):
        n = self.n if level is None else (1 << level)
        return torch.zeros((n, n), dtype=torch.float64, device=self.device)

    def sync(self):
        torch.cuda.synchronize(self.device)
        check(_lib.load().afm_sync(self._h), "afm_sync")

    def update_timing(self, enable=True):
        """HIP-event timing of the finest-level post-smoothing + update launch
        of every multigrid solve on this context (enable resets)."""
        self._call("afm_update_timing", 1 if enable else 0)

    def update_timing_read(self):
        """(average duration in microseconds, number of launches timed)."""
        avg, n = C.c_double(), C.c_long()
        check(_lib.load().afm_update_timing_read(self._h, C.byref(avg), C.byref(n)), "afm_update_timing_read")
        return avg.value, n.value

    def kernel_timing_read(self, kind):
        """Timing of kind 0 (the outer update), 1 (the CG forms' last
        post-smoothing) or 2 (the finest pre-smoothing + residual +
        restriction): (average microseconds, launches timed)."""
        avg, n = C.c_double(), C.c_long()
        check(_lib.load().afm_kernel_timing_read(self._h, int(kind), C.byref(avg), C.byref(n)),
              "afm_kernel_timing_read")
        return avg.value, n.value

    def close(self):
        if self._h:
            torch.cuda.synchronize(self.device)
            _lib.load().afm_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- ib ---------------------------------------------------------------
    def delta(self, r):
        r = self._f(r, name="r")
        out = torch.empty_like(r)
        self._call("afm_delta", _ptr(r), r.numel(), _ptr(out))
        return out

    def interp(self, comp, field, X):
        """U_k = J(X) field for MAC component comp (0 = u, 1 = v)."""
        self._f(field, (self.ny, self.nx), "field")
        X = self._f(X, name="X")
        m = X.shape[0]
        U = torch.empty(m, dtype=torch.float64, device=self.device)
        self._call("afm_interp", comp, _ptr(field), _ptr(X), m, _ptr(U))
        return U

    def spread(self, comp, F, ds, X, out=None):
        """f = S(X) (F ds) on MAC component comp, ascending marker order per cell."""
        X = self._f(X, name="X")
        m = X.shape[0]
        self._f(F, (m,), "F")
        self._f(ds, (m,), "ds")
        if out is None:
            out = torch.empty((self.ny, self.nx), dtype=torch.float64, device=self.device)
        self._f(out, (self.ny, self.nx), "out")
        self._call("afm_spread", comp, _ptr(F), _ptr(ds), _ptr(X), m, _ptr(out))
        return out

    # -- mg ---------------------------------------------------------------
    def laplacian(self, x, level=None):
        level = self.levels if level is None else level
        n = 1 << level
        self._f(x, (n, n), "x")
        y = torch.empty_like(x)
        self._call("afm_laplacian_apply", level, _ptr(x), _ptr(y))
        return y

    def mg_smooth(self, b, x=None, sweeps=2, level=None):
        level = self.levels if level is None else level
        out = torch.empty_like(b)
        self._call("afm_mg_smooth", level, _ptr(x) if x is not None else None, _ptr(b), _ptr(out), sweeps,
                   1 if x is None else 0)
        return out

    def mg_restrict(self, fine, level=None):
        level = self.levels if level is None else level
        n = 1 << level
        out = torch.empty((n // 2, n // 2), dtype=torch.float64, device=self.device)
        self._call("afm_mg_restrict", level, _ptr(fine), _ptr(out))
        return out

    def mg_prolong(self, coarse, level=None):
        level = self.levels if level is None else level
        n = 1 << level
        out = torch.empty((n, n), dtype=torch.float64, device=self.device)
        self._call("afm_mg_prolong", level, _ptr(coarse), _ptr(out))
        return out

    def mg_vcycle(self, rhs, nu1=2, nu2=2):
        self._f(rhs, (self.ny, self.nx), "rhs")
        self._call("afm_mg_vcycle", _ptr(rhs), nu1, nu2)
        return rhs

    def mg_solve(self, rhs, tol=1e-10, tol_mode=1, nu1=2, nu2=2, precision=0, max_cycles=200):
        """In place: rhs <- solution.  Returns (cycles, final RMS residual)."""
        self._f(rhs, (self.ny, self.nx), "rhs")
        cyc, rms = C.c_int(), C.c_double()
        self._call("afm_mg_solve", _ptr(rhs), tol, tol_mode, nu1, nu2, precision, max_cycles, C.byref(cyc),
                   C.byref(rms))
        return cyc.value, rms.value

    # -- the 3-D mode: cubes of s = width * 2^levels cells per axis, walls in
    # y (shape (s, s, s) as (z, y, x)); width 1 is geometry<L,1,1,1> --
    def _f3(self, t, name, width=1):
        if not isinstance(t, torch.Tensor) or t.dtype != torch.float64 or t.device != self.device \
                or not t.is_contiguous() or t.dim() != 3 or len(set(t.shape)) != 1:
            raise TypeError(f"{name} must be a contiguous float64 (s, s, s) tensor on {self.device}")
        s = t.shape[0]
        if width < 1 or width & (width - 1) or s % width or (s // width) & (s // width - 1):
            raise ValueError(f"{name}: {s} cells per axis is not width * 2^levels for width {width}")
        return (s // width).bit_length() - 1

    def laplacian3(self, x, width=1):
        L = self._f3(x, "x", width)
        y = torch.empty_like(x)
        self._call("afm_laplacian3_apply", L, width, _ptr(x), _ptr(y))
        return y

    def mg3_vcycle(self, rhs, nu1=2, nu2=2, width=1, tol=0.0):
        """One V-cycle recurse(levels, rhs) of the 3-D solver, in place (tol:
        the coarsest level's CG tolerance when width > 1)."""
        L = self._f3(rhs, "rhs", width)
        self._call("afm_mg3_vcycle", L, width, _ptr(rhs), nu1, nu2, tol)
        return rhs

    def mg3_solve(self, rhs, tol=1e-10, tol_mode=1, nu1=2, nu2=2, max_cycles=200, width=1):
        """In place: rhs <- solution of the 3-D walled problem.  Returns
        (cycles, final RMS residual)."""
        L = self._f3(rhs, "rhs", width)
        cyc, rms = C.c_int(), C.c_double()
        self._call("afm_mg3_solve", L, width, _ptr(rhs), tol, tol_mode, nu1, nu2, max_cycles, C.byref(cyc),
                   C.byref(rms))
        return cyc.value, rms.value

    def mg3_restrict(self, fine, width=1):
        L = self._f3(fine, "fine", width)
        s = fine.shape[0] // 2
        coarse = torch.empty((s, s, s), dtype=torch.float64, device=self.device)
        self._call("afm_mg3_restrict", L, width, _ptr(fine), _ptr(coarse))
        return coarse

    def mg3_prolong(self, coarse, width=1):
        s = coarse.shape[0] * 2
        fine = torch.empty((s, s, s), dtype=torch.float64, device=self.device)
        L = self._f3(fine, "fine", width)
        self._call("afm_mg3_prolong", L, width, _ptr(self._f(coarse, (s // 2,) * 3, "coarse")), _ptr(fine))
        return fine

    def mg3_smooth(self, x, b, sweeps=1, width=1):
        L = self._f3(x, "x", width)
        self._f(b, tuple(x.shape), "b")
        self._call("afm_mg3_smooth", L, width, _ptr(x), _ptr(b), sweeps)
        return x

    def laplacian3_cg(self, b, x, tol=1e-10, max_iters=100000, width=1):
        L = self._f3(x, "x", width)
        self._f(b, tuple(x.shape), "b")
        it, nrm = C.c_int(), C.c_double()
        self._call("afm_laplacian3_cg", L, width, _ptr(b), _ptr(x), tol, max_iters, C.byref(it), C.byref(nrm))
        return it.value, nrm.value

    def helmholtz3_boundary(self, kind, scale, top, bottom):
        """helmholtz_periodic (kind 1) / helmholtz_dirichlet (kind 3)
        boundary(top, bottom): the wall terms of the right-hand side."""
        n = top.shape[0]
        self._f(top, (n, n), "top")
        self._f(bottom, (n, n), "bottom")
        out = torch.empty((n, n if kind == 1 else n - 1, n), dtype=torch.float64, device=self.device)
        self._call("afm_helmholtz3_boundary", n.bit_length() - 1, kind, scale, _ptr(top), _ptr(bottom), _ptr(out))
        return out

    def projection3_boundary(self, phx, phz, ut, ub, wt, wb):
        """projection::boundary: the wall vectors corrected in place."""
        n = phx.shape[0]
        self._f3(phx, "phx")
        self._f3(phz, "phz")
        for t, nm in ((ut, "ut"), (ub, "ub"), (wt, "wt"), (wb, "wb")):
            self._f(t, (n, n), nm)
        self._call("afm_projection3_boundary", n.bit_length() - 1, _ptr(phx), _ptr(phz), _ptr(ut), _ptr(ub),
                   _ptr(wt), _ptr(wb))

    # -- dt::matrix in device CSR form (int64 row starts and columns) --
    def _csr(self, rs, col, val):
        for t, nm, dt in ((rs, "row_start", torch.int64), (col, "column", torch.int64), (val, "values", torch.float64)):
            if not isinstance(t, torch.Tensor) or t.dtype != dt or t.device != self.device or not t.is_contiguous():
                raise TypeError(f"{nm} must be a contiguous {dt} tensor on {self.device}")
        return rs.numel() - 1

    def csr_spmv(self, rs, col, val, x):
        rows = self._csr(rs, col, val)
        y = torch.empty(rows, dtype=torch.float64, device=self.device)
        self._call("afm_csr_spmv", rows, _ptr(rs), _ptr(col), _ptr(val), _ptr(x), _ptr(y))
        return y

    def csr_cg(self, rs, col, val, b, x, tol=1e-10, max_iters=100000):
        rows = self._csr(rs, col, val)
        it, nrm = C.c_int(), C.c_double()
        self._call("afm_csr_cg", rows, _ptr(rs), _ptr(col), _ptr(val), _ptr(b), _ptr(x), tol, max_iters,
                   C.byref(it), C.byref(nrm))
        return it.value, nrm.value

    def csr_smooth(self, rs, col, val, x, b, sweeps=1):
        rows = self._csr(rs, col, val)
        self._call("afm_csr_smooth", rows, _ptr(rs), _ptr(col), _ptr(val), _ptr(x), _ptr(b), sweeps)
        return x

    def axpy(self, a, x, y):
        self._call("afm_axpy", x.numel(), a, _ptr(x), _ptr(y))
        return y

    # -- 3-D wall-bounded Navier-Stokes operators: u, w, p (n, n, n), v
    # (n, n-1, n) on the interior y-faces, wall vectors (n, n), all [z, y, x]
    def _v3(self, t, n, name):
        return self._f(t, (n, n - 1, n), name)

    def div3(self, u, v, w, top=None, bottom=None):
        L = self._f3(u, "u")
        n = 1 << L
        self._v3(v, n, "v")
        self._f(w, (n, n, n), "w")
        z = None
        if top is None or bottom is None:
            z = torch.zeros((n, n), dtype=torch.float64, device=self.device)
        top = z if top is None else self._f(top, (n, n), "top")
        bottom = z if bottom is None else self._f(bottom, (n, n), "bottom")
        d = torch.empty_like(u)
        self._call("afm_div3", L, _ptr(u), _ptr(v), _ptr(w), _ptr(top), _ptr(bottom), _ptr(d))
        return d

    def grad3(self, p):
        L = self._f3(p, "p")
        n = 1 << L
        px, pz = torch.empty_like(p), torch.empty_like(p)
        py = torch.empty((n, n - 1, n), dtype=torch.float64, device=self.device)
        self._call("afm_grad3", L, _ptr(p), _ptr(px), _ptr(py), _ptr(pz))
        return px, py, pz

    def advect3(self, u, v, w):
        L = self._f3(u, "u")
        n = 1 << L
        self._v3(v, n, "v")
        self._f(w, (n, n, n), "w")
        a, c = torch.empty_like(u), torch.empty_like(u)
        b = torch.empty_like(v)
        self._call("afm_advect3", L, _ptr(u), _ptr(v), _ptr(w), _ptr(a), _ptr(b), _ptr(c))
        return a, b, c

    def _h3(self, kind, x, name):
        if kind == 1:
            return self._f3(x, name)
        if kind != 3:
            raise ValueError("kind must be 1 (helmholtz_periodic) or 3 (helmholtz_dirichlet)")
        self._f(x, None, name)
        if x.dim() != 3 or x.shape[0] != x.shape[2] or x.shape[1] != x.shape[0] - 1:
            raise TypeError(f"{name} must have shape (n, n-1, n)")
        n = x.shape[0]
        if n < 2 or n & (n - 1):
            raise ValueError(f"{name}: n = {n} is not a power of two >= 2")
        return n.bit_length() - 1

    def helmholtz3(self, kind, scale, x):
        L = self._h3(kind, x, "x")
        y = torch.empty_like(x)
        self._call("afm_helmholtz3_apply", L, kind, scale, _ptr(x), _ptr(y))
        return y

    def helmholtz3_cg(self, kind, scale, b, x, tol=1e-10, max_iters=100000):
        """In place on x (the initial guess); returns (iterations, |r|_2)."""
        L = self._h3(kind, b, "b")
        self._f(x, tuple(b.shape), "x")
        it, res = C.c_int(), C.c_double()
        self._call("afm_helmholtz3_cg", L, kind, scale, _ptr(b), _ptr(x), tol, max_iters, C.byref(it), C.byref(res))
        return it.value, res.value

    def _markers3(self, X):
        self._f(X, None, "X")
        if X.dim() != 2 or X.shape[1] != 3:
            raise ValueError("X must have shape (m, 3)")
        return X.shape[0]

    def interp3(self, comp, field, X):
        """Velocity component comp (0 u, 1 v, 2 w) of the 3-D MAC grid at the
        markers X (m, 3); field is (n, n, n), or (n, n-1, n) for v."""
        n = field.shape[0] if isinstance(field, torch.Tensor) and field.dim() == 3 else 0
        shape = (n, n - 1, n) if comp == 1 else (n, n, n)
        L = self._h3(3 if comp == 1 else 1, self._f(field, shape, "field"), "field")
        m = self._markers3(X)
        U = torch.empty(m, dtype=torch.float64, device=self.device)
        self._call("afm_interp3", L, comp, _ptr(field), _ptr(X), m, _ptr(U))
        return U

    def spread3(self, comp, F, ds, X, n):
        """Forces F (m,) with source volumes ds (m,) at markers X (m, 3) onto
        component comp of the n^3 MAC grid: (n, n, n), or (n, n-1, n) for v."""
        m = self._markers3(X)
        self._f(F, (m,), "F")
        self._f(ds, (m,), "ds")
        if n < 2 or n & (n - 1):
            raise ValueError(f"n = {n} is not a power of two >= 2")
        shape = (n, n - 1, n) if comp == 1 else (n, n, n)
        f = torch.empty(shape, dtype=torch.float64, device=self.device)
        self._call("afm_spread3", n.bit_length() - 1, comp, _ptr(F), _ptr(ds), _ptr(X), m, _ptr(f))
        return f

    # -- ns operators -----------------------------------------------------
    def helmholtz(self, scale, x):
        self._f(x, (self.ny, self.nx), "x")
        y = torch.empty_like(x)
        self._call("afm_helmholtz_apply", scale, _ptr(x), _ptr(y))
        return y

    def helmholtz_cg(self, scale, b, x, tol=1e-10, tol_mode=1, max_iters=10000):
        self._f(b, (self.ny, self.nx), "b")
        self._f(x, (self.ny, self.nx), "x")
        it, res = C.c_int(), C.c_double()
        self._call("afm_helmholtz_cg", scale, _ptr(b), _ptr(x), tol, tol_mode, max_iters, C.byref(it), C.byref(res))
        return it.value, res.value

    def div(self, u, v):
        d = torch.empty_like(u)
        self._call("afm_div", _ptr(self._f(u, (self.ny, self.nx), "u")), _ptr(self._f(v, (self.ny, self.nx), "v")),
                   _ptr(d))
        return d

    def grad_sub(self, p, u, v):
        self._call("afm_grad_sub", _ptr(p), _ptr(u), _ptr(v))
        return u, v

    def advect(self, u, v):
        ua, va = torch.empty_like(u), torch.empty_like(v)
        self._call("afm_advect", _ptr(u), _ptr(v), _ptr(ua), _ptr(va))
        return ua, va


class MGSolver:
    """mg::solver<G>(tol, nu1, nu2) (mg/solver.h:60-62); tol is the
    reference's absolute RMS threshold unless tol_mode=1 (relative)."""

    def __init__(self, ctx, tol, nu1=2, nu2=2, tol_mode=0, precision=0, max_cycles=1000):
        self.ctx, self.tol, self.nu1, self.nu2 = ctx, tol, nu1, nu2
        self.tol_mode, self.precision, self.max_cycles = tol_mode, precision, max_cycles
        self.last_rms = None

    def solve(self, rhs):
        cyc, rms = self.ctx.mg_solve(rhs, self.tol, self.tol_mode, self.nu1, self.nu2, self.precision,
                                     self.max_cycles)
        self.last_rms = rms
        return cyc


class NSSolver:
    """ns::solver<G>(mu, rho, dt, tol, u0, v0, w0) (ns/solver.h:61-72) on the
    periodic plane: u0 and v0 are aliased and updated in place by step()."""

    def __init__(self, ctx, mu, rho, dt, tol, u0, v0, tol_mode=1, precision=0):
        self.ctx = ctx
        ctx._f(u0, (ctx.ny, ctx.nx), "u0")
        ctx._f(v0, (ctx.ny, ctx.nx), "v0")
        self.u, self.v = u0, v0
        h = C.c_void_p()
        ctx._call("afm_ns_create", mu, rho, dt, tol, tol_mode, precision, _ptr(u0), _ptr(v0), C.byref(h))
        self._h = h
        self.mu, self.rho, self.dt, self.tol = mu, rho, dt, tol

    def _lib(self, name, *args):
        self.ctx._call  # keep the stream binding fresh
        s = torch.cuda.current_stream(self.ctx.device).cuda_stream
        if s != self.ctx._stream:
            self.ctx._bind_stream()
        return check(getattr(_lib.load(), name)(self._h, *args), name)

    def step(self, fx, fy):
        self.ctx._f(fx, (self.ctx.ny, self.ctx.nx), "fx")
        self.ctx._f(fy, (self.ctx.ny, self.ctx.nx), "fy")
        self._lib("afm_ns_step", _ptr(fx), _ptr(fy))

    def project(self, u, v):
        self._lib("afm_ns_project", _ptr(u), _ptr(v))

    def get_p(self):
        p = torch.empty_like(self.u)
        self._lib("afm_ns_get_p", _ptr(p))
        return p

    def get_start(self, kind):
        """The CG forms' potentials: kind 0 / 1 the next start of the
        backward-Euler half step / Crank-Nicolson full step (extrapolated
        from the kind's last two potentials), kind 2 / 3 the kind's last
        potential; None while unset."""
        phi = torch.empty_like(self.u)
        ok = C.c_int()
        self._lib("afm_ns_get_start", kind, _ptr(phi), C.byref(ok))
        return phi if ok.value else None

    def set_start(self, kind, phi):
        """Set a potential of get_start (None: unset).  Setting kind 0 / 1
        unsets kind 2 / 3: restore 0, 1 before 2, 3."""
        if phi is not None:
            self.ctx._f(phi, (self.ctx.ny, self.ctx.nx), "phi")
        self._lib("afm_ns_set_start", kind, _ptr(phi) if phi is not None else None)

    def stats(self):
        a, b, c, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self._lib("afm_ns_stats", C.byref(a), C.byref(b), C.byref(c), C.byref(d))
        return {"cycles": a.value, "cg_u": b.value, "cg_v": c.value, "have_half_stepped": bool(d.value)}

    def last_residual(self):
        """(final RMS residual, its bound) of the last projection's solve."""
        a, b = C.c_double(), C.c_double()
        self._lib("afm_ns_last_residual", C.byref(a), C.byref(b))
        return a.value, b.value

    def close(self):
        if self._h:
            torch.cuda.synchronize(self.ctx.device)
            _lib.load().afm_ns_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class NS3Solver:
    """ns::solver<geometry<L,1,1,1>>(mu, rho, dt, tol, u0, v0, w0)
    (ns/solver.h:37-236): the reference's 3-D wall-bounded solver on n^3
    cells, n = 2^L.  u0, w0 are (n, n, n), v0 (n, n-1, n), indexed [z, y, x];
    they are aliased, projected at creation and advanced in place by step().
    tol is the reference's absolute bound for the multigrid and CG solves.
    Runs on any Context's device and stream (the 2-D grid is not used)."""

    def __init__(self, ctx, mu, rho, dt, tol, u0, v0, w0):
        self.ctx = ctx
        L = ctx._f3(u0, "u0")
        self.n = n = 1 << L
        ctx._v3(v0, n, "v0")
        ctx._f(w0, (n, n, n), "w0")
        self.u, self.v, self.w = u0, v0, w0
        h = C.c_void_p()
        ctx._call("afm_ns3_create", L, mu, rho, dt, tol, _ptr(u0), _ptr(v0), _ptr(w0), C.byref(h))
        self._h = h
        self.mu, self.rho, self.dt, self.tol = mu, rho, dt, tol

    def _lib(self, name, *args):
        s = torch.cuda.current_stream(self.ctx.device).cuda_stream
        if s != self.ctx._stream:
            self.ctx._bind_stream()
        return check(getattr(_lib.load(), name)(self._h, *args), name)

    def step(self, ut, ub, wt, wb, fx, fy, fz):
        """ns/solver.h:90-140.  ut, ub, wt, wb (n, n) receive projection::
        boundary's correction in place, as the reference's step does."""
        n, f = self.n, self.ctx._f
        for t, name in ((ut, "ut"), (ub, "ub"), (wt, "wt"), (wb, "wb")):
            f(t, (n, n), name)
        f(fx, (n, n, n), "fx")
        f(fy, (n, n - 1, n), "fy")
        f(fz, (n, n, n), "fz")
        self._lib("afm_ns3_step", _ptr(ut), _ptr(ub), _ptr(wt), _ptr(wb), _ptr(fx), _ptr(fy), _ptr(fz))

    def get_u(self):
        return self.u

    def get_v(self):
        return self.v

    def get_w(self):
        return self.w

    def get_p(self):
        p = torch.empty_like(self.u)
        self._lib("afm_ns3_get_p", _ptr(p))
        return p

    def stats(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        self._lib("afm_ns3_stats", C.byref(a), C.byref(b), C.byref(c))
        return {"cycles": a.value, "cg": b.value, "have_half_stepped": bool(c.value)}

    def close(self):
        if self._h:
            torch.cuda.synchronize(self.ctx.device)
            _lib.load().afm_ns3_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class IBSolver:
    """Tethered markers coupled to an NSSolver (see include/afm.h)."""

    def __init__(self, ns, X, R, k, ds):
        ctx = ns.ctx
        X = ctx._f(X, name="X")
        m = X.shape[0]
        ctx._f(R, (m, 2), "R")
        ctx._f(k, (m,), "k")
        ctx._f(ds, (m,), "ds")
        self.ns, self.m = ns, m
        h = C.c_void_p()
        check(_lib.load().afm_ib_create(ns._h, m, _ptr(X), _ptr(R), _ptr(k), _ptr(ds), C.byref(h)), "afm_ib_create")
        self._h = h

    def step(self, steps=1):
        ctx = self.ns.ctx
        if torch.cuda.current_stream(ctx.device).cuda_stream != ctx._stream:
            ctx._bind_stream()
        check(_lib.load().afm_ib_step(self._h, steps), "afm_ib_step")

    def grad(self, steps, msize, checkpoint_every=20, need_grad=True):
        """Run `steps` steps; return (J, dJ/dk, dJ/dR) by the discrete adjoint
        (J = kinetic energy + centroid displacements of membranes of msize
        markers).  Gradients are None with need_grad=False."""
        ctx = self.ns.ctx
        if torch.cuda.current_stream(ctx.device).cuda_stream != ctx._stream:
            ctx._bind_stream()
        J = C.c_double()
        gk = torch.empty(self.m, dtype=torch.float64, device=ctx.device) if need_grad else None
        gR = torch.empty((self.m, 2), dtype=torch.float64, device=ctx.device) if need_grad else None
        check(_lib.load().afm_ib_grad(self._h, steps, checkpoint_every, msize, C.byref(J),
                                      _ptr(gk) if need_grad else None, _ptr(gR) if need_grad else None),
              "afm_ib_grad")
        return J.value, gk, gR

    def markers(self):
        X = torch.empty((self.m, 2), dtype=torch.float64, device=self.ns.ctx.device)
        check(_lib.load().afm_ib_markers(self._h, _ptr(X)), "afm_ib_markers")
        return X

    def close(self):
        if self._h:
            torch.cuda.synchronize(self.ns.ctx.device)
            _lib.load().afm_ib_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def delta_host(r, n):
    """The build's delta evaluated by libafm on the CPU (numpy float64 in/out)."""
    import numpy as np
    r = np.ascontiguousarray(r, dtype=np.float64)
    out = np.empty_like(r)
    check(_lib.load().afm_delta_host(r.ctypes.data_as(C.c_void_p), r.size, n, out.ctypes.data_as(C.c_void_p)),
          "afm_delta_host")
    return out
