"""Text formats of the reference and a checkpoint format for IB runs.

The reference prints a `dt::vector` as ``[v0, v1, ...]'`` (dt/vector.h:153-165)
and a `dt::point` as ``index (x, y, z)`` (dt/point.h:50-54), through a default
`std::ostream`: six significant digits, ``%g`` style (``1e-05``, ``0.333333``,
``inf``, ``-nan``).  `parse_vector` / `parse_points` read such printouts (one or
several in a text), so output of the reference's own test programs can serve
as fixtures; `format_vector` / `format_point` write the same form (the default
precision reproduces the reference's text exactly).

A checkpoint holds what an IB run needs to continue: the velocity fields, the
marker state (positions, tether points, stiffness, source volumes), the grid
and fluid parameters and the step count, and -- for the CG forms of the
projection (precision 1, 3) -- the potentials those solves start from (per
step kind the next start and the last potential it is extrapolated from).  Between IB steps nothing else carries over (each step's
backward-Euler half step takes its old velocity from u, v).  It is one
``.npz`` file, written and read with numpy alone and never unpickled; the
parameters sit in a JSON string array inside it.
"""
import json
import math
import re

import numpy as np

_NUM = r"[-+]?(?:nan|inf|(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?)"
_VEC = re.compile(r"\[\s*((?:" + _NUM + r"\s*(?:,\s*" + _NUM + r"\s*)*)?)\]'")
_PT = re.compile(r"(\d+)\s*\(\s*(" + _NUM + r")\s*,\s*(" + _NUM + r")\s*,\s*(" + _NUM + r")\s*\)")
CHECKPOINT_VERSION = 3        # 2: optional start potentials (2); 3: four (versions 1, 2 are read too)


def _num(tok):
    t = tok.strip().lower()
    if t.endswith("nan"):
        return math.nan
    return float(t)


def parse_vector(text):
    """The first ``[a, b, ...]'`` in text as a float64 array (ValueError if
    there is none)."""
    vs = parse_vectors(text)
    if not vs:
        raise ValueError("no [ ... ]' vector in the text")
    return vs[0]


def parse_vectors(text):
    """Every ``[a, b, ...]'`` in text, in order."""
    out = []
    for m in _VEC.finditer(text):
        body = m.group(1).strip()
        out.append(np.array([_num(t) for t in body.split(",")] if body else [], dtype=np.float64))
    return out


def _g(v, precision):
    if math.isnan(v):
        return "-nan" if math.copysign(1.0, v) < 0 else "nan"
    if math.isinf(v):
        return "inf" if v > 0 else "-inf"
    return f"{v:.{precision}g}"


def format_vector(values, precision=6):
    """``[v0, v1, ...]'`` as a default std::ostream prints it (precision 6)."""
    return "[" + ", ".join(_g(float(v), precision) for v in np.asarray(values, dtype=np.float64).ravel()) + "]'"


def parse_points(text):
    """Every ``index (x, y, z)`` in text: (indices int64 array, xyz (k, 3))."""
    idx, xyz = [], []
    for m in _PT.finditer(text):
        idx.append(int(m.group(1)))
        xyz.append([_num(m.group(2)), _num(m.group(3)), _num(m.group(4))])
    return np.array(idx, dtype=np.int64), np.array(xyz, dtype=np.float64).reshape(-1, 3)


def format_point(index, x, y, z, precision=6):
    return f"{int(index)} ({_g(float(x), precision)}, {_g(float(y), precision)}, {_g(float(z), precision)})"


# ------------------------------------------------------------ checkpoints --
_ARRAYS = ("u", "v", "X", "R", "k", "ds")


def save_checkpoint(path, u, v, X, R, k, ds, *, n, mu, rho, dt, tol, tol_mode=1, precision=0, step=0,
                    starts=(None, None, None, None)):
    """Write an IB state.  Arrays may be torch tensors (any device) or numpy
    arrays; they are stored as float64.  starts: the solver's potentials
    NSSolver.get_start(0) .. get_start(3) (fewer are padded with None),
    None where unset."""
    starts = tuple(starts) + (None,) * (4 - len(starts))
    def host(a):
        if hasattr(a, "detach"):
            a = a.detach().cpu().numpy()
        return np.ascontiguousarray(a, dtype=np.float64)
    arrays = {name: host(a) for name, a in zip(_ARRAYS, (u, v, X, R, k, ds))}
    for j, phi in enumerate(starts):
        if phi is not None:
            arrays[f"start{j}"] = host(phi)
            if arrays[f"start{j}"].shape != (n, n):
                raise ValueError(f"start potential {j} must be {n} x {n}")
    m = arrays["X"].shape[0]
    if arrays["u"].shape != (n, n) or arrays["v"].shape != (n, n):
        raise ValueError(f"u and v must be {n} x {n}")
    if arrays["X"].shape != (m, 2) or arrays["R"].shape != (m, 2) or arrays["k"].shape != (m,) \
            or arrays["ds"].shape != (m,):
        raise ValueError("X, R must be (m, 2) and k, ds (m,)")
    meta = {"format": "afmibins-ib-checkpoint", "version": CHECKPOINT_VERSION, "n": int(n), "mu": float(mu),
            "rho": float(rho), "dt": float(dt), "tol": float(tol), "tol_mode": int(tol_mode),
            "precision": int(precision), "step": int(step), "markers": int(m),
            "starts": [s is not None for s in starts]}
    np.savez(path, meta=np.array([json.dumps(meta)]), **arrays)


def load_checkpoint(path):
    """(meta dict, arrays dict of numpy float64) from save_checkpoint's file."""
    with np.load(path, allow_pickle=False) as z:
        meta = json.loads(str(z["meta"][0]))
        if meta.get("format") != "afmibins-ib-checkpoint":
            raise ValueError(f"{path} is not an IB checkpoint")
        if meta.get("version") not in (1, 2, CHECKPOINT_VERSION):
            raise ValueError(f"{path}: checkpoint version {meta.get('version')} (this build reads 1 to "
                             f"{CHECKPOINT_VERSION})")
        arrays = {name: np.array(z[name], dtype=np.float64) for name in _ARRAYS}
        for j, present in enumerate(meta.get("starts", [False, False])):
            if present:
                arrays[f"start{j}"] = np.array(z[f"start{j}"], dtype=np.float64)
    return meta, arrays


def restore(path, device="cuda:0"):
    """Build (ctx, ns, ib, u, v, meta) from a checkpoint, ready to step on."""
    import torch
    from . import api as A
    meta, a = load_checkpoint(path)
    t = lambda x: torch.tensor(x, dtype=torch.float64, device=device)
    ctx = A.Context(meta["n"], torch.device(device).index or 0)
    u, v = t(a["u"]), t(a["v"])
    ns = A.NSSolver(ctx, meta["mu"], meta["rho"], meta["dt"], meta["tol"], u, v, tol_mode=meta["tol_mode"],
                    precision=meta["precision"])
    # building the solver projects its initial velocity (ns/solver.h:71); the
    # saved fields are already the projected ones of a running simulation, so
    # they are written back into the solver's (aliased) u, v
    u.copy_(t(a["u"]))
    v.copy_(t(a["v"]))
    # the starts first: setting one unsets the last potential of its kind
    for j in range(4):
        if f"start{j}" in a:
            ns.set_start(j, t(a[f"start{j}"]))
    ib = A.IBSolver(ns, t(a["X"]), t(a["R"]), t(a["k"]), t(a["ds"]))
    return ctx, ns, ib, u, v, meta
