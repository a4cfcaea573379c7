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

Two forms are timed on identical inputs: plain FP64, and FP32 V(2,2) cycles
inside FP64 iterative refinement (`value`, the headline).  Inputs are resident
in HBM before the timed region; the timed region is K full steps bracketed by
device synchronisation.

The JSON line also carries the roofline of the dominant kernel (the fused
2-sweep red-black smoother at 2048^2, timed with HIP events on the stream it
runs on) and the CPU baseline (the oracle's restatement of the reference path:
lexicographic Gauss-Seidel V(2,2), CSR CG, one core) on a one-step sample.
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


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=100)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--n", type=int, default=2048)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--forms", default="fp64,mixed")
    p.add_argument("--no-extras", action="store_true", help="skip the large-grid and transfer measurements")
    p.add_argument("--kernel-only", action="store_true", help="only the smoother loop and a calibration copy")
    return p.parse_args()


def run_form(A, W, ctx, precision, args, inputs, dev):
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
    ib.close(); ns.close()
    return out


def _events(dev):
    import torch
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def smoother_roofline(A, ctx, dev, reps=50):
    """Average duration of the fused 2-sweep red-black smoother on the
    context's finest level, timed with HIP events on the stream it is launched
    on (torch's current stream, to which the context is bound)."""
    import torch
    n = ctx.n
    b = torch.randn((n, n), dtype=torch.float64, device=dev)
    x = [torch.randn((n, n), dtype=torch.float64, device=dev) for _ in range(2)]

    def launch(i):
        ctx._call("afm_mg_smooth", ctx.levels, C.c_void_p(x[i & 1].data_ptr()), C.c_void_p(b.data_ptr()),
                  C.c_void_p(x[(i + 1) & 1].data_ptr()), 2, 0)
    for i in range(5):
        launch(i)
    s = torch.cuda.current_stream(dev)
    e0, e1 = _events(dev)
    e0.record(s)
    for i in range(reps):
        launch(i)
    e1.record(s)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    bytes_per_launch = 24.0 * n * n        # read x, read b, write x (2 sweeps fused)
    gbs = bytes_per_launch / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": f"k_smooth<double,2,M_NORMAL> (2 fused red-black sweeps, {n}x{n})",
            "bytes_per_launch": bytes_per_launch, "avg_launch_us": round(ms * 1e3, 2)}


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
        for i in range(3):
            fn(i)
        s = torch.cuda.current_stream(dev)
        e0, e1 = _events(dev)
        e0.record(s)
        for i in range(reps):
            fn(i)
        e1.record(s)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[name] = {"us_per_call": round(ms * 1e3, 2), "GB/s": round(nbytes / (ms * 1e-3) / 1e9, 1),
                     "frac_hbm": round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "marker_per_s": round(m / (ms * 1e-3), 1)}
    ctx.sync()
    ctx.close()
    out["case"] = f"{n}x{n} grid, {m} markers (512 membranes + 512 filaments x 256, 8 Gaussian clusters)"
    return out


def kernel_only(A, dev):
    """For counter runs: the smoother at 2048^2 (50 launches) and a plain copy
    of 256 MiB (the calibration kernel for FETCH_SIZE / WRITE_SIZE)."""
    import torch
    ctx = A.Context(2048, 0)
    r = smoother_roofline(A, ctx, dev)
    a = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device=dev).fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    ctx.close()
    print(json.dumps({"smoother": r, "copy_bytes_each_way": a.numel() * 8}), flush=True)


def cpu_baseline(inputs, n):
    from oracle import oracle as O
    u0, v0, X, R, k, ds = inputs
    O.ib_run(u0, v0, X, R, k, ds, n, 1e-3, 1.0, 2e-5, 1e-10, 1, tol_mode=1, smoother=0, delta_mode=0)
    sec = O.lib().or_last_step_seconds()
    return {"value": round(1.0 / sec, 4), "unit": "IB steps/s", "cores": 1, "kind": "port",
            "sample": f"1 IB step of the same {n}x{n} / {X.shape[0]}-marker workload on one host core "
                      f"({sec:.1f} s): the oracle's restatement of the reference path (lexicographic "
                      f"Gauss-Seidel V(2,2) multigrid, CSR CG, glibc-cos delta)",
            "host_nproc": os.cpu_count()}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              f"(the path is single-GPU; see DESIGN.md for how it would shard)", file=sys.stderr)
        sys.exit(2)
    import torch
    import afmibins_amd as A
    from afmibins_amd import workloads as W
    dev = torch.device("cuda", 0)
    if args.kernel_only:
        kernel_only(A, dev)
        return
    n = args.n
    inputs = W.config2_inputs(n=n)
    ctx = A.Context(n, 0)
    forms = {}
    for name in args.forms.split(","):
        prec = 0 if name == "fp64" else 1
        forms[name] = run_form(A, W, ctx, prec, args, inputs, dev)
        print(f"[bench] {name}: {forms[name]['ms_per_step']:.3f} ms/step, "
              f"cycles {forms[name]['last_mg_cycles']}, cg {forms[name]['last_cg_iters']}", file=sys.stderr, flush=True)
    roof = smoother_roofline(A, ctx, dev)
    print(f"[bench] smoother: {roof['avg_launch_us']} us, {roof['achieved']} GB/s", file=sys.stderr, flush=True)
    head = forms.get("mixed", forms.get("fp64"))
    m = inputs[2].shape[0]
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
        "dtype": "f64 (f32 V-cycles in f64 refinement)" if "mixed" in forms else "f64",
        "data": "synthetic (seeded membranes and Fourier velocity)",
        "config": {"workload": f"2-D periodic IB, {n}x{n} MAC grid, {m} markers (64 membranes x 256), "
                               "mu=1e-3, rho=1, dt=2e-5, V(2,2) MG + CG Helmholtz to 1e-10 relative",
                   "grid": n, "markers": m},
        "forms": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in f.items()}
                  for k, f in forms.items()},
        "marker_steps_per_s": round(head["steps_per_s"] * m, 1),
        "roofline": roof,
    }
    if "fp64" in forms and "mixed" in forms:
        line["mixed_over_fp64"] = round(forms["mixed"]["steps_per_s"] / forms["fp64"]["steps_per_s"], 3)
    if not args.no_extras:
        big = A.Context(8192, 0)
        line["roofline_hbm_8192"] = smoother_roofline(A, big, dev, reps=20)
        big.close()
        print(f"[bench] smoother 8192^2: {line['roofline_hbm_8192']['achieved']} GB/s", file=sys.stderr, flush=True)
        line["transfer"] = transfer_rates(A, W, dev)
        print(f"[bench] transfer: {line['transfer']}", file=sys.stderr, flush=True)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(inputs, n)
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
