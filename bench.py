#!/usr/bin/env python
"""IB time steps per second on one MI355X.

Workload (the reference's 2048^2 benchmark configuration): doubly periodic
2048x2048 MAC grid; 64 elastic ellipses (centres U[0.1,0.9]^2, semi-axes
U[0.02,0.05]) of 256 markers each (16,384 markers) tethered with k = 1 to
equal-area circles; initial velocity from seeded Fourier modes |k| <= 8
(|u_k| ~ |k|^-2, exactly divergence-free, RMS 1); rho = 1, mu = 1e-3,
dt = 2e-5.  One step = spread, backward-Euler half step (CG Helmholtz x2 +
multigrid projection), interpolation + marker update, spread, Crank-Nicolson
full step, interpolation + marker update.  Solves stop at 1e-10 relative
residual (multigrid: RMS; CG: 2-norm).

Two forms are timed on identical inputs: plain FP64 (the reference's
multigrid iteration: nested start, then V(2,2) cycles until the residual test
holds), and the mixed form (`value`, the headline): flexible CG in FP64 --
true residual, search directions, inner products, solution -- preconditioned
by one FP32 V(2,2) cycle per iteration, each step's projection starting from
the previous potential of the same step kind.  `--forms` also takes mixed_ir
(the reference's iteration with FP32 cycles, FP64 iterative refinement) and
fp64_cg (the CG form with FP64 cycles), which split the mixed form's gain
into its precision and Krylov parts.  Inputs are resident in HBM before the
timed region; the timed region is K full steps bracketed by device
synchronisation.

`roofline` is the headline form's largest FP64 launch, timed with HIP events
on the stream it runs on during extra steps of the same workload: the CG
update k_fcg_update (iterations after the first), which reads z and p
(stored in the cycle's FP32, 4 + 4 B), x and b (FP64, 8 + 8) and writes p
(4), x (8) and r rounded to FP32 (4): 40 B per finest cell (56 B with FP64
cycles).  The FP64
form's is the finest-level post-smoothing fused with the outer update
(k_smooth_update): read e_pre and b (8 + 8), the coarse correction (8 per 4
cells = 2), xs and rhs (8 + 8); write xs and r (8 + 8): 50 B (37 B with FP32
cycles).  `roofline_smoothing` times the same way each form's finest-level
pre-smoothing + residual + restriction launch and, for the CG forms, the
preconditioner cycle's last post-smoothing (bytes per cell at SMOOTHING).
`cpu_baseline` is the oracle's restatement of the reference path
(kd-tree gather spreading, lexicographic multigrid, CSR CG, glibc cos) on one
host core over a three-step sample of the same workload.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# form: (precision code of afm_mg_solve, timed launch, its bytes per finest cell, counter summary)
FORMS = {
    "fp64": (0, "k_smooth_update<double, double, 2>", 50.0, None),
    "mixed": (1, "k_fcg_update<float>", 40.0, "k_fcg_update_f32_2048_pmc.txt"),
    "mixed_ir": (2, "k_smooth_update<float, float, 2>", 37.0, "k_smooth_update_f32_2048_pmc.txt"),
    "fp64_cg": (3, "k_fcg_update<double>", 56.0, None),
}
# the finest-level smoothing launches of each form (timed like the update):
# the pre-smoothing from zero + residual + restriction reads b and writes x
# and the coarse right-hand side (a quarter cell): 4 + 4 + 1 B per cell in
# FP32, 18 in FP64; the CG forms' last post-smoothing reads e_pre, b, r_prev
# and a quarter cell of the coarse correction and writes z: 17 B in FP32, 34
# in FP64 (the other forms fuse it with the update above)
SMOOTHING = {
    "fp64": (("pre", "k_smooth<double,2,M_ZERO,RESTRICT>", 18.0),),
    "mixed": (("pre", "k_smooth<float,2,M_ZERO,RESTRICT>", 9.0), ("post", "k_smooth_fcg<float,2>", 17.0)),
    "mixed_ir": (("pre", "k_smooth<float,2,M_ZERO,RESTRICT>", 9.0),),
    "fp64_cg": (("pre", "k_smooth<double,2,M_ZERO,RESTRICT>", 18.0), ("post", "k_smooth_fcg<double,2>", 34.0)),
}
DTYPE = {"fp64": "f64", "mixed": "f64 (f32 V-cycles preconditioning f64 CG)",
         "mixed_ir": "f64 (f32 V-cycles in f64 refinement)", "fp64_cg": "f64 (f64 V-cycles preconditioning f64 CG)"}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=100)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--n", type=int, default=2048)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--forms", default="fp64,mixed,mixed_ir,fp64_cg")
    p.add_argument("--no-extras", action="store_true", help="skip the smoother, transfer and solver-only measurements")
    p.add_argument("--kernel-only", action="store_true", help="a few mixed steps and a calibration copy (counter runs)")
    p.add_argument("--transfer-only", action="store_true",
                   help="spreading and interpolation of the 4096^2 / 262,144-marker stress case and a calibration "
                        "copy (counter runs)")
    p.add_argument("--grad", action="store_true",
                   help="only the differentiation check: 1024^2, 32 membranes (8,192 markers), 200 steps, "
                        "checkpoints every 20, FD along 32 directions")
    p.add_argument("--stress", action="store_true",
                   help="only the combined stress case: 4096^2, 262,144 clustered markers, mixed precision, "
                        "50 steps with checkpoints every 10, dt 1e-5, FD along 4 directions")
    return p.parse_args()


def hbm_roofline(nbytes, us, kernel, extra=None):
    gbs = nbytes / (us * 1e-6) / 1e9
    r = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kernel,
         "bytes_per_launch": nbytes, "avg_launch_us": round(us, 2)}
    if extra:
        r.update(extra)
    return r


def pmc_traffic(name):
    """HBM bytes per launch of a form's roofline kernel from its committed
    counter summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None."""
    if not name:
        return None
    try:
        for line in open(os.path.join(ROOT, "profiles", name)):
            if line.startswith("hbm_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def run_form(A, ctx, precision, args, inputs, dev, timing_steps=0):
    import torch
    u0, v0, X, R, k, ds = inputs
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=dev)
    u, v = t(u0), t(v0)
    ns = A.NSSolver(ctx, 1e-3, 1.0, 2e-5, 1e-10, u, v, tol_mode=1, precision=precision)
    ib = A.IBSolver(ns, t(X), t(R), t(k), t(ds))
    ib.step(args.warmup)
    ctx.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ib.step(args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ctx.sync()
    st = ns.stats()
    out = {"ms_per_step": 1e3 * (t1 - t0) / args.steps, "steps_per_s": args.steps / (t1 - t0),
           "last_mg_cycles": st["cycles"], "last_cg_iters": [st["cg_u"], st["cg_v"]]}
    if timing_steps:
        ctx.update_timing(True)
        ib.step(timing_steps)
        torch.cuda.synchronize()
        out["update_us"], out["update_launches"] = ctx.update_timing_read()
        out["post_us"], out["post_launches"] = ctx.kernel_timing_read(1)
        out["pre_us"], out["pre_launches"] = ctx.kernel_timing_read(2)
        ctx.update_timing(False)
    out["u_final"] = u.clone()
    ib.close(); ns.close()
    return out


def _events():
    import torch
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def time_launches(fn, reps, warm=3):
    """Average duration of fn() launches, HIP events on torch's current
    stream (the stream the library is bound to)."""
    import torch
    for i in range(warm):
        fn(i)
    s = torch.cuda.current_stream()
    e0, e1 = _events()
    e0.record(s)
    for i in range(reps):
        fn(i)
    e1.record(s)
    e1.synchronize()
    return 1e3 * e0.elapsed_time(e1) / reps


def smoother_roofline(A, ctx, dev, reps=50):
    """The 2-sweep red-black smoother (FP64, plain mode) on the finest level."""
    import torch
    n = ctx.n
    b = torch.randn((n, n), dtype=torch.float64, device=dev)
    x = [torch.randn((n, n), dtype=torch.float64, device=dev) for _ in range(2)]

    def launch(i):
        ctx._call("afm_mg_smooth", ctx.levels, C.c_void_p(x[i & 1].data_ptr()), C.c_void_p(b.data_ptr()),
                  C.c_void_p(x[(i + 1) & 1].data_ptr()), 2, 0)
    us = time_launches(launch, reps)
    return hbm_roofline(24.0 * n * n, us, f"k_smooth<double,2,M_NORMAL> (2 fused red-black sweeps, {n}x{n}; "
                                          "read x, b, write x: 24 B/cell)")


def transfer_rates(A, W, dev, reps=10):
    """Spreading and interpolation of the clustered 262,144-marker case on a
    4096x4096 grid (both MAC components), timed per call with HIP events;
    algorithmic bytes: spread 8 B per cell (dense write) + 32 B per marker
    (X, F, ds); interpolation 16 + 16*8 + 8 = 152 B per marker."""
    import torch
    n = 4096
    X, R, k, ds = W.clustered_structures()
    m = X.shape[0]
    ctx = A.Context(n, 0)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=dev)
    Xd, dsd = t(X), t(ds)
    F = torch.randn(m, dtype=torch.float64, device=dev)
    f = [torch.empty((n, n), dtype=torch.float64, device=dev) for _ in range(2)]
    U = torch.randn((n, n), dtype=torch.float64, device=dev)
    out = {}
    for name, fn, nbytes in (
            ("spread", lambda i: ctx.spread(i & 1, F, dsd, Xd, out=f[i & 1]), 8.0 * n * n + 32.0 * m),
            ("interp", lambda i: ctx.interp(i & 1, U, Xd), 152.0 * m)):
        us = time_launches(fn, reps)
        out[name] = {"us_per_call": round(us, 2), "GB/s": round(nbytes / (us * 1e-6) / 1e9, 1),
                     "frac_hbm": round(nbytes / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
                     "marker_per_s": round(m / (us * 1e-6), 1)}
    ctx.sync()
    ctx.close()
    out["case"] = f"{n}x{n} grid, {m} markers (512 membranes + 512 filaments x 256, 8 Gaussian clusters)"
    return out


def solver_only(A, W, dev, n=4096, reps=3):
    """Poisson by multigrid (the FP64 reference iteration, the mixed CG form
    and FP32 cycles in the reference iteration) and
    Helmholtz (I - dt mu L, dt = 2e-5, mu = 1e-3) by CG, both to 1e-10
    relative residual, on a seeded 256-mode |k|^-2 right-hand side."""
    import torch
    ctx = A.Context(n, 0)
    f = W.fourier_rhs_device(n, 256, 32, 11, dev)
    res = {"case": f"{n}x{n} periodic, 256 Fourier modes |k|<=32, |k|^-2 amplitudes, mean removed"}
    sol = {}
    for name, prec in (("fp64", 0), ("mixed", 1), ("mixed_ir", 2)):
        x = f.clone()
        cyc, rms = ctx.mg_solve(x, 1e-10, 1, 2, 2, prec)        # warm (and the solution)
        sol[name] = x
        xs = [f.clone() for _ in range(reps)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(reps):
            ctx.mg_solve(xs[i], 1e-10, 1, 2, 2, prec)
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / reps
        res[f"poisson_{name}"] = {"ms_per_solve": round(ms, 3), "cycles": cyc, "final_rel_rms": rms}
    d = sol["mixed"] - sol["fp64"]
    d = d - d.mean()
    res["poisson_mixed_over_fp64"] = round(res["poisson_fp64"]["ms_per_solve"] / res["poisson_mixed"]["ms_per_solve"], 3)
    res["poisson_mixed_ir_over_fp64"] = round(res["poisson_fp64"]["ms_per_solve"]
                                              / res["poisson_mixed_ir"]["ms_per_solve"], 3)
    res["poisson_mixed_vs_fp64_rel_err"] = float(d.norm() / (sol["fp64"] - sol["fp64"].mean()).norm())
    scale = -2e-5 * 1e-3
    x = torch.zeros_like(f)
    it, nrm = ctx.helmholtz_cg(scale, f, x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        x.zero_()
        ctx.helmholtz_cg(scale, f, x)
    torch.cuda.synchronize()
    res["helmholtz_cg_fp64"] = {"ms_per_solve": round(1e3 * (time.perf_counter() - t0) / reps, 3), "iters": it}
    ctx.sync()
    ctx.close()
    return res


def gradient_check(A, W, dev, n, structures, msize, steps, every, precision, dt, mu, ndir, tol, label):
    """dJ/dk and dJ/dR by the adjoint (checkpoint every `every` steps) against
    central finite differences, h = 1e-6, along `ndir` seeded directions;
    J = 1/2 sum |u(T)|^2 h^2 + sum over structures |c(T) - c(0)|^2."""
    import numpy as np
    import torch
    from afmibins_amd.autograd import IBSetup, run
    X, R, k, ds = structures
    u0, v0 = W.fourier_velocity(n, 8, seed=3)
    t = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    ctx = A.Context(n, 0)
    st = IBSetup(ctx, t(u0), t(v0), t(X), t(ds), mu, 1.0, dt, tol, steps, msize, every, precision)
    kt, Rt = t(k), t(R)
    run(st, kt, Rt, False)                      # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J, gk, gR = run(st, kt, Rt, True)
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    t0 = time.perf_counter()
    run(st, kt, Rt, False)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    gk, gR = gk.cpu().numpy(), gR.cpu().numpy()
    print(f"[bench] {label}: forward {t_fwd:.2f} s, forward + adjoint {t_grad:.2f} s", file=sys.stderr, flush=True)
    rng = np.random.default_rng(17)
    h, errs = 1e-6, []
    for _ in range(ndir):
        dk, dR = rng.standard_normal(k.shape), rng.standard_normal(R.shape)
        Jp = run(st, t(k + h * dk), t(R + h * dR), False)[0]
        Jm = run(st, t(k - h * dk), t(R - h * dR), False)[0]
        fd = (Jp - Jm) / (2 * h)
        ad = float(gk @ dk + np.sum(gR * dR))
        errs.append(abs(fd - ad) / abs(ad))
        print(f"[bench] FD direction {len(errs)}: fd {fd:.12e} adjoint {ad:.12e} rel {errs[-1]:.2e}",
              file=sys.stderr, flush=True)
    ctx.close()
    m = X.shape[0]
    return {"case": label, "grid": n, "markers": m, "steps": steps, "checkpoint_every": every,
            "precision": "mixed" if precision else "fp64", "solver_tol_rel": tol, "J": J,
            "forward_s": round(t_fwd, 3), "forward_steps_per_s": round(steps / t_fwd, 2),
            "forward_plus_adjoint_s": round(t_grad, 3), "adjoint_over_forward": round(t_grad / t_fwd, 2),
            "fd_directions": ndir, "fd_h": h, "fd_rel_err_max": max(errs), "fd_rel_err_median": float(np.median(errs)),
            "grad_k_norm": float(np.linalg.norm(gk)), "grad_R_norm": float(np.linalg.norm(gR))}


def kernel_only(A, W, dev):
    """For counter runs: three mixed-form steps of the benchmark workload and
    a plain 256 MiB copy (the calibration kernel for FETCH_SIZE / WRITE_SIZE)."""
    import torch
    inputs = W.config2_inputs(n=2048)
    ctx = A.Context(2048, 0)

    class Args:
        warmup, steps = 1, 2
    r = run_form(A, ctx, 1, Args, inputs, dev)
    ctx.close()
    print(json.dumps({"ms_per_step": r["ms_per_step"], "copy_bytes_each_way": calibration_copy(dev)}), flush=True)


def calibration_copy(dev):
    """Five copies of a 256 MiB buffer: the known traffic that scales the
    FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    a = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device=dev).fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    return a.numel() * 8


def cpu_baseline(inputs, n, steps=3):
    """The oracle's restatement of the reference path on one host core:
    lexicographic Gauss-Seidel V(2,2) multigrid with the nested start, CSR CG
    for the Helmholtz solves, the delta with glibc's cos of the x87 argument,
    and spreading as the reference does it (every face gathers the markers
    within 2/n from a kd-tree); interpolation by the marker windows."""
    from oracle import oracle as O
    u0, v0, X, R, k, ds = inputs
    O.lib().or_set_spread_gather(1)
    try:
        O.ib_run(u0, v0, X, R, k, ds, n, 1e-3, 1.0, 2e-5, 1e-10, steps, tol_mode=1, smoother=0, delta_mode=0)
    finally:
        O.lib().or_set_spread_gather(0)
    sec = O.lib().or_last_step_seconds() / steps
    return {"value": round(1.0 / sec, 5), "unit": "IB steps/s", "cores": 1, "kind": "port",
            "sample": f"{steps} IB steps of the same {n}x{n} / {X.shape[0]}-marker workload on one host core "
                      f"({sec:.1f} s per step): the oracle's restatement of the reference path (lexicographic "
                      f"Gauss-Seidel V(2,2) multigrid, CSR CG, glibc-cos delta, kd-tree gather spreading)",
            "host_nproc": os.cpu_count()}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              f"(it times one GPU's share of the per-timestep path, a whole grid per GPU)", file=sys.stderr)
        sys.exit(2)
    import torch
    import afmibins_amd as A
    from afmibins_amd import workloads as W
    dev = torch.device("cuda", 0)
    if args.kernel_only:
        kernel_only(A, W, dev)
        return
    if args.transfer_only:
        out = transfer_rates(A, W, dev)
        print(json.dumps({"transfer": out, "copy_bytes_each_way": calibration_copy(dev)}), flush=True)
        return
    if args.grad:
        X, R, k, ds = W.membranes(32, 256, seed=2)
        print(json.dumps(gradient_check(A, W, dev, 1024, (X, R, k, ds), 256, 200, 20, 0, 2e-5, 1e-3, 32, 1e-12,
                                        "differentiation: 1024^2, 32 membranes x 256")), flush=True)
        return
    if args.stress:
        print(json.dumps(gradient_check(A, W, dev, 4096, W.clustered_structures(), 256, 50, 10, 1, 1e-5, 1e-3, 4,
                                        1e-10, "stress: 4096^2, 512 membranes + 512 filaments x 256, 8 clusters")),
              flush=True)
        return
    n = args.n
    inputs = W.config2_inputs(n=n)
    ctx = A.Context(n, 0)
    forms = {}
    for name in args.forms.split(","):
        forms[name] = run_form(A, ctx, FORMS[name][0], args, inputs, dev, timing_steps=5)
        print(f"[bench] {name}: {forms[name]['ms_per_step']:.3f} ms/step, cycles {forms[name]['last_mg_cycles']}, "
              f"cg {forms[name]['last_cg_iters']}, update launch {forms[name]['update_us']:.2f} us",
              file=sys.stderr, flush=True)
    head_name = "mixed" if "mixed" in forms else "fp64"
    head = forms[head_name]
    m = inputs[2].shape[0]
    cells = float(n) * n
    def form_roof(name):
        _, kern, bpc, pmc = FORMS[name]
        r = hbm_roofline(bpc * cells, forms[name]["update_us"], f"{kern} ({n}x{n}, {bpc:.0f} B/cell)",
                         {"launches_timed": forms[name]["update_launches"]})
        r["traffic"] = pmc_traffic(pmc)
        return r
    roof = form_roof(head_name)
    line = {
        "metric": "IB steps/s @2048² grid, 16k markers: FP64 vs FP32-MG+FP64 refine at 1e-10 resid",
        "value": round(head["steps_per_s"], 3),
        "unit": "IB steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[head_name],
        "data": "synthetic (seeded membranes and Fourier velocity)",
        "config": {"workload": f"2-D periodic IB, {n}x{n} MAC grid, {m} markers (64 membranes x 256), "
                               "mu=1e-3, rho=1, dt=2e-5, V(2,2) MG + CG Helmholtz to 1e-10 relative",
                   "grid": n, "markers": m},
        "forms": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in f.items() if kk != "u_final"}
                  for k, f in forms.items()},
        "marker_steps_per_s": round(head["steps_per_s"] * m, 1),
        "roofline": roof,
    }
    if "fp64" in forms and "mixed" in forms:
        line["mixed_over_fp64"] = round(forms["mixed"]["steps_per_s"] / forms["fp64"]["steps_per_s"], 3)
        a, b = forms["mixed"]["u_final"], forms["fp64"]["u_final"]
        line["mixed_vs_fp64_u_rel_err"] = float((a - b).norm() / b.norm())
    line["roofline_smoothing"] = {
        name: {which: hbm_roofline(bpc * cells, forms[name][f"{which}_us"], f"{kern} ({n}x{n}, {bpc:.0f} B/cell)",
                                   {"launches_timed": forms[name][f"{which}_launches"]})
               for which, kern, bpc in SMOOTHING[name]}
        for name in forms}
    for name in forms:
        if name != head_name:
            line[f"roofline_{name}_form"] = form_roof(name)
        if name not in ("fp64", "mixed") and "fp64" in forms:
            line[f"{name}_over_fp64"] = round(forms[name]["steps_per_s"] / forms["fp64"]["steps_per_s"], 3)
    if not args.no_extras:
        line["smoother_2048"] = smoother_roofline(A, ctx, dev)
        big = A.Context(8192, 0)
        line["smoother_8192"] = smoother_roofline(A, big, dev, reps=20)
        big.close()
        print(f"[bench] smoother: {line['smoother_2048']['achieved']} GB/s at 2048^2, "
              f"{line['smoother_8192']['achieved']} GB/s at 8192^2", file=sys.stderr, flush=True)
        line["transfer"] = transfer_rates(A, W, dev)
        print(f"[bench] transfer: {line['transfer']}", file=sys.stderr, flush=True)
        line["solver_4096"] = solver_only(A, W, dev)
        print(f"[bench] solver-only: {line['solver_4096']}", file=sys.stderr, flush=True)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(inputs, n)
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
