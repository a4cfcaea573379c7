"""ctypes bindings for the CPU oracle (liboracle.so) and the reference build
(_ref/libafmref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the afmibins_amd package.

Array conventions follow afm_oracle.cpp: 2-D fields are (ny, nx) row-major
(x fastest), markers are (m, 2) interleaved x, y.  3-D fields are
(nz, ny, nx) with the reference's idx = x + nx*(y + ny*z).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libafmref.so")

_D = C.c_double
_U = C.c_uint
_I = C.c_int
_L = C.c_long
_P = C.c_void_p


def build(force=False):
    """Compile liboracle.so (and _ref/ when the reference tree is present)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", HERE, "-s", "all"], check=True)


def _ptr(a):
    return a.ctypes.data_as(_P)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        sig = {
            "or_delta": (_D, [_D, _U, _I]),
            "or_delta_arg": (_D, [_D, _U]),
            "or_cos_glibc": (_D, [_D]),
            "or_cos_cr": (_D, [_D]),
            "or_delta_many": (None, [_P, _L, _U, _I, _P]),
            "or_delta_arg_many": (None, [_P, _L, _U, _P]),
            "or_interp2": (None, [_I, _P, _U, _U, _U, _U, _U, _P, _L, _P, _I]),
            "or_spread2": (None, [_I, _P, _P, _P, _L, _U, _U, _U, _U, _U, _P, _I]),
            "or_spread2_gather": (None, [_I, _P, _P, _P, _L, _U, _U, _U, _U, _U, _P, _I]),
            "or_set_spread_gather": (None, [_I]),
            "or_window2": (None, [_I, _D, _D, _U, _U, _U, _U, _U, _I, _P, _P]),
            "or_interp3": (None, [_I, _P, _U, _U, _U, _U, _U, _U, _U, _P, _L, _P, _I]),
            "or_spread3": (None, [_I, _P, _P, _P, _L, _U, _U, _U, _U, _U, _U, _U, _P, _I]),
            "or_stencil_nnz": (_L, [_U, _U, _U, _U, _I, _D]),
            "or_stencil_csr": (None, [_U, _U, _U, _U, _I, _D, _P, _P, _P]),
            "or_stencil_apply": (None, [_U, _U, _U, _U, _I, _D, _P, _P]),
            "or_smooth": (None, [_U, _U, _U, _U, _I, _P, _P, _I]),
            "or_cg": (_U, [_U, _U, _U, _U, _I, _D, _P, _P, _D, _U, _P]),
            "or_restrict3": (None, [_U, _U, _U, _P, _P]),
            "or_restrict2": (None, [_U, _U, _P, _P]),
            "or_prolong3": (None, [_U, _U, _U, _P, _P]),
            "or_prolong2": (None, [_U, _U, _P, _P]),
            "or_mg_solve": (_U, [_I, _U, _U, _U, _U, _D, _I, _U, _U, _I, _U, _P, _P]),
            "or_mg_vcycle": (None, [_I, _U, _U, _U, _U, _D, _U, _U, _I, _P]),
            "or_div2": (None, [_U, _U, _U, _P, _P, _P]),
            "or_grad2": (None, [_U, _U, _U, _P, _P, _P]),
            "or_advect2": (None, [_U, _U, _U, _P, _P, _P, _P]),
            "or_ib_run": (_U, [_U, _U, _U, _D, _D, _D, _D, _I, _I, _I, _P, _P, _L, _P, _P, _P, _P, _I, _P, _P]),
            "or_last_step_seconds": (_D, []),
            "or_ib_grad": (_D, [_U, _U, _U, _D, _D, _D, _D, _I, _I, _I, _P, _P, _L, _P, _P, _P, _P, _L, _I, _P,
                                _P]),
            "or_ns_run": (_U, [_U, _U, _U, _D, _D, _D, _D, _I, _I, _P, _P, _P, _P, _I, _P]),
            "or_div3": (None, [_U, _U, _U, _U, _P, _P, _P, _P, _P, _P]),
            "or_grad3": (None, [_U, _P, _P, _P, _P]),
            "or_advect3": (None, [_U, _P, _P, _P, _P, _P, _P]),
            "or_hbound3": (None, [_I, _U, _D, _P, _P, _P]),
            "or_proj_boundary3": (None, [_U, _P, _P, _P, _P, _P, _P]),
            "or_ns3_run": (None, [_U, _D, _D, _D, _D, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _P, _P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def have_ref():
    return os.path.exists(REF_PATH)


def ref():
    """The reference's own templates, instantiated by ref_shim.cc."""
    global _ref
    if _ref is None:
        R = C.CDLL(REF_PATH)
        sig = {
            "ref_delta": (_D, [_D, _U]),
            "ref_markers": (_U, []),
            "ref_interp": (_I, [_I, _I, _P, _P, _P]),
            "ref_spread": (_I, [_I, _I, _P, _P, _P, _P]),
            "ref_lap_apply": (_I, [_I, _P, _P]),
            "ref_smooth": (_I, [_I, _P, _P, _I]),
            "ref_cg": (_L, [_I, _I, _D, _P, _P, _D]),
            "ref_csr": (_L, [_I, _I, _D, _P, _P, _P]),
            "ref_restrict": (_I, [_I, _P, _P]),
            "ref_prolong": (_I, [_I, _P, _P]),
            "ref_mg_solve": (_L, [_I, _D, _U, _U, _P]),
            "ref_div_grad": (_I, [_I, _P, _P, _P, _P, _P, _P, _P, _P]),
            "ref_ns3_run": (_I, [_I, _D, _D, _D, _D, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _P]),
            "ref_advect": (_I, [_P, _P, _P, _P, _P, _P]),
            "ref_ib_test": (_I, [_D, _D, _D, _P, _P]),
            "ref_restrict_g1": (_I, [_P, _P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(R, name)
            f.restype = res
            f.argtypes = args
        _ref = R
    return _ref


# ------------------------------------------------------------------ delta
def delta(r, n, mode=1):
    r = _f64(np.atleast_1d(r))
    out = np.empty_like(r)
    lib().or_delta_many(_ptr(r), r.size, n, mode, _ptr(out))
    return out


def delta_arg(r, n):
    r = _f64(np.atleast_1d(r))
    out = np.empty_like(r)
    lib().or_delta_arg_many(_ptr(r), r.size, n, _ptr(out))
    return out


# ------------------------------------------------------------- 2-D IB
def interp2(comp, field, X, n, mode=1):
    field = _f64(field)
    ny, nx = field.shape
    X = _f64(X)
    m = X.shape[0]
    U = np.empty(m)
    lib().or_interp2(comp, _ptr(field), nx, ny, n, nx // n, ny // n, _ptr(X), m, _ptr(U), mode)
    return U


def spread2(comp, F, ds, X, nx, ny, n, mode=1):
    F, ds, X = _f64(F), _f64(ds), _f64(X)
    f = np.empty((ny, nx))
    lib().or_spread2(comp, _ptr(F), _ptr(ds), _ptr(X), X.shape[0], nx, ny, n, nx // n, ny // n, _ptr(f), mode)
    return f


def spread2_gather(comp, F, ds, X, nx, ny, n, mode=1):
    """Spreading as the reference computes it: a kd-tree box gather per face."""
    F, ds, X = _f64(F), _f64(ds), _f64(X)
    f = np.empty((ny, nx))
    lib().or_spread2_gather(comp, _ptr(F), _ptr(ds), _ptr(X), X.shape[0], nx, ny, n, nx // n, ny // n, _ptr(f), mode)
    return f


def window2(comp, x, y, nx, ny, n, mode=1):
    idx = np.empty(16, dtype=np.int64)
    w = np.empty(16)
    lib().or_window2(comp, x, y, nx, ny, n, nx // n, ny // n, mode, _ptr(idx), _ptr(w))
    return idx, w


def interp3(comp, field, X, n, W=1, H=1, D=1, mode=0):
    field = _f64(field).ravel()
    X = _f64(X)
    nx, ny, nz = W * n, H * n, D * n
    U = np.empty(X.shape[0])
    lib().or_interp3(comp, _ptr(field), nx, ny, nz, n, W, H, D, _ptr(X), X.shape[0], _ptr(U), mode)
    return U


def spread3(comp, F, ds, X, n, W=1, H=1, D=1, mode=0):
    F, ds, X = _f64(F), _f64(ds), _f64(X)
    nx, ny, nz = W * n, H * n, D * n
    N = nx * (ny - 1 if comp == 1 else ny) * nz
    f = np.empty(N)
    lib().or_spread3(comp, _ptr(F), _ptr(ds), _ptr(X), X.shape[0], nx, ny, nz, n, W, H, D, _ptr(f), mode)
    return f


# ---------------------------------------------------------- operators
def stencil_apply(x, nx, ny, nz, n, kind=0, scale=0.0):
    x = _f64(x).ravel()
    y = np.empty_like(x)
    lib().or_stencil_apply(nx, ny, nz, n, kind, scale, _ptr(x), _ptr(y))
    return y


def stencil_csr(nx, ny, nz, n, kind=0, scale=0.0):
    nnz = lib().or_stencil_nnz(nx, ny, nz, n, kind, scale)
    N = nx * ny * nz
    rs = np.empty(N + 1, dtype=np.int64)
    col = np.empty(nnz, dtype=np.int64)
    val = np.empty(nnz)
    lib().or_stencil_csr(nx, ny, nz, n, kind, scale, _ptr(rs), _ptr(col), _ptr(val))
    return rs, col, val


def smooth(x, b, nx, ny, nz, n, smoother=0, sweeps=1):
    x = _f64(x).ravel().copy()
    b = _f64(b).ravel()
    lib().or_smooth(nx, ny, nz, n, smoother, _ptr(x), _ptr(b), sweeps)
    return x


def cg(x, b, nx, ny, nz, n, kind=0, scale=0.0, tol=1e-10, max_iter=100000):
    x = _f64(x).ravel().copy()
    b = _f64(b).ravel()
    fn = np.zeros(1)
    it = lib().or_cg(nx, ny, nz, n, kind, scale, _ptr(x), _ptr(b), tol, max_iter, _ptr(fn))
    return x, it, fn[0]


def restrict2(fine):
    fine = _f64(fine)
    ny, nx = fine.shape
    out = np.empty((ny // 2, nx // 2))
    lib().or_restrict2(nx, ny, _ptr(fine), _ptr(out))
    return out


def prolong2(coarse):
    coarse = _f64(coarse)
    cy, cx = coarse.shape
    out = np.zeros((2 * cy, 2 * cx))
    lib().or_prolong2(2 * cx, 2 * cy, _ptr(coarse), _ptr(out))
    return out


def restrict3(fine, nx, ny, nz):
    fine = _f64(fine).ravel()
    out = np.empty(nx * ny * nz // 8)
    lib().or_restrict3(nx, ny, nz, _ptr(fine), _ptr(out))
    return out


def prolong3(coarse, fine, nx, ny, nz):
    coarse = _f64(coarse).ravel()
    fine = _f64(fine).ravel().copy()
    lib().or_prolong3(nx, ny, nz, _ptr(coarse), _ptr(fine))
    return fine


def mg_solve(rhs, levels, dim=2, W=1, H=1, D=1, tol=1e-10, tol_mode=0, nu1=2, nu2=2, smoother=1,
             max_cycles=1000):
    r = _f64(rhs).ravel().copy()
    fr = np.zeros(1)
    cyc = lib().or_mg_solve(dim, levels, W, H, D, tol, tol_mode, nu1, nu2, smoother, max_cycles, _ptr(r), _ptr(fr))
    return r.reshape(np.shape(rhs)), cyc, fr[0]


def mg_vcycle(rhs, levels, dim=2, W=1, H=1, D=1, tol=1e-10, nu1=2, nu2=2, smoother=1):
    r = _f64(rhs).ravel().copy()
    lib().or_mg_vcycle(dim, levels, W, H, D, tol, nu1, nu2, smoother, _ptr(r))
    return r.reshape(np.shape(rhs))


def div2(u, v, n):
    u, v = _f64(u), _f64(v)
    ny, nx = u.shape
    d = np.empty_like(u)
    lib().or_div2(nx, ny, n, _ptr(u), _ptr(v), _ptr(d))
    return d


def grad2(p, n):
    p = _f64(p)
    ny, nx = p.shape
    a, b = np.empty_like(p), np.empty_like(p)
    lib().or_grad2(nx, ny, n, _ptr(p), _ptr(a), _ptr(b))
    return a, b


def advect2(u, v, n):
    u, v = _f64(u), _f64(v)
    ny, nx = u.shape
    a, b = np.empty_like(u), np.empty_like(u)
    lib().or_advect2(nx, ny, n, _ptr(u), _ptr(v), _ptr(a), _ptr(b))
    return a, b


def ns_run(u, v, fx, fy, n, mu, rho, dt, tol, tol_mode=1, smoother=1, calls=2):
    u, v = _f64(u).copy(), _f64(v).copy()
    fx, fy = _f64(fx), _f64(fy)
    ny, nx = u.shape
    p = np.empty_like(u)
    cyc = lib().or_ns_run(nx, ny, n, mu, rho, dt, tol, tol_mode, smoother, _ptr(u), _ptr(v), _ptr(fx), _ptr(fy),
                          calls, _ptr(p))
    return u, v, p, cyc


# ---- 3-D wall-bounded mode: n^3 cells, arrays indexed [z, y, x]; u, w, p
# (n, n, n), v (n, n-1, n), wall vectors (n, n) as [z, x] ----
def div3(u, v, w, top=None, bottom=None):
    u, v, w = _f64(u), _f64(v), _f64(w)
    nz, ny, nx = u.shape
    top = np.zeros((nz, nx)) if top is None else _f64(top)
    bottom = np.zeros((nz, nx)) if bottom is None else _f64(bottom)
    d = np.empty_like(u)
    lib().or_div3(nx, ny, nz, nx, _ptr(u), _ptr(v), _ptr(w), _ptr(top), _ptr(bottom), _ptr(d))
    return d


def grad3(p):
    p = _f64(p)
    n = p.shape[0]
    px, pz = np.empty_like(p), np.empty_like(p)
    py = np.empty((n, n - 1, n))
    lib().or_grad3(n, _ptr(p), _ptr(px), _ptr(py), _ptr(pz))
    return px, py, pz


def advect3(u, v, w):
    u, v, w = _f64(u), _f64(v), _f64(w)
    n = u.shape[0]
    a, c = np.empty_like(u), np.empty_like(u)
    b = np.empty((n, n - 1, n))
    lib().or_advect3(n, _ptr(u), _ptr(v), _ptr(w), _ptr(a), _ptr(b), _ptr(c))
    return a, b, c


def hbound3(kind, n, scale, top, bottom):
    """helmholtz_periodic (kind 1) / helmholtz_dirichlet (kind 3) boundary()."""
    top, bottom = _f64(top), _f64(bottom)
    b = np.empty((n, n, n) if kind == 1 else (n, n - 1, n))
    lib().or_hbound3(kind, n, scale, _ptr(top), _ptr(bottom), _ptr(b))
    return b


def proj_boundary3(phx, phz, ut, ub, wt, wb):
    phx, phz = _f64(phx), _f64(phz)
    ut, ub, wt, wb = (_f64(a).copy() for a in (ut, ub, wt, wb))
    lib().or_proj_boundary3(phx.shape[0], _ptr(phx), _ptr(phz), _ptr(ut), _ptr(ub), _ptr(wt), _ptr(wb))
    return ut, ub, wt, wb


def ns3_run(L, u, v, w, walls, f, mu, rho, dt, tol, steps, smoother=1):
    """ns::solver<geometry<L,1,1,1>>: construct over (u, v, w), `steps`
    step() calls with forces f = (fx, fy, fz) and wall vectors
    walls = (ut, ub, wt, wb) (updated as the reference updates them).
    Returns (u, v, w, walls, p, (last V-cycles, last step's CG iterations))."""
    u, v, w = (_f64(a).copy() for a in (u, v, w))
    ut, ub, wt, wb = (_f64(a).copy() for a in walls)
    fx, fy, fz = (_f64(a) for a in f)
    p = np.empty_like(u)
    st = np.zeros(2, dtype=np.uint32)
    lib().or_ns3_run(L, mu, rho, dt, tol, smoother, _ptr(u), _ptr(v), _ptr(w), _ptr(ut), _ptr(ub), _ptr(wt), _ptr(wb),
                     _ptr(fx), _ptr(fy), _ptr(fz), steps, _ptr(p), st.ctypes.data_as(C.c_void_p))
    return u, v, w, (ut, ub, wt, wb), p, (int(st[0]), int(st[1]))


def ib_run(u, v, X, R, k, ds, n, mu, rho, dt, tol, steps, tol_mode=1, smoother=1, delta_mode=1, want_phi=False):
    """IB steps; returns (u, v, X, cycles, cg_iters) and, with want_phi, the
    potential phi of the last projection as a sixth value."""
    u, v = _f64(u).copy(), _f64(v).copy()
    X = _f64(X).copy()
    R, k, ds = _f64(R), _f64(k), _f64(ds)
    ny, nx = u.shape
    cgi = C.c_uint(0)
    phi = np.zeros_like(u) if want_phi else None
    cyc = lib().or_ib_run(nx, ny, n, mu, rho, dt, tol, tol_mode, smoother, delta_mode, _ptr(u), _ptr(v),
                          X.shape[0], _ptr(X), _ptr(R), _ptr(k), _ptr(ds), steps, C.byref(cgi),
                          _ptr(phi) if want_phi else None)
    return (u, v, X, cyc, cgi.value, phi) if want_phi else (u, v, X, cyc, cgi.value)


def ib_grad(u, v, X, R, k, ds, n, mu, rho, dt, tol, steps, msize, tol_mode=1, smoother=1, delta_mode=1,
            grad=True):
    """J (kinetic energy + centroid displacement) after `steps` IB steps and,
    with grad=True, the adjoint gradients dJ/dk (m,) and dJ/dR (m, 2)."""
    u, v, X, R, k, ds = (_f64(a) for a in (u, v, X, R, k, ds))
    ny, nx = u.shape
    m = X.shape[0]
    gk = np.zeros(m) if grad else None
    gR = np.zeros((m, 2)) if grad else None
    J = lib().or_ib_grad(nx, ny, n, mu, rho, dt, tol, tol_mode, smoother, delta_mode, _ptr(u), _ptr(v), m, _ptr(X),
                         _ptr(R), _ptr(k), _ptr(ds), msize, steps, _ptr(gk) if grad else None,
                         _ptr(gR) if grad else None)
    return (J, gk, gR) if grad else J
