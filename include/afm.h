/*
 * afm.h -- C ABI of libafm, the gfx950 (MI355X) implementation of the
 * immersed-boundary hot path of the reference IB/Navier-Stokes library.
 *
 * Scope: the doubly periodic 2-D MAC grid (the reference's periodic (x, z)
 * plane; u on x-faces at (i, j+1/2)/n, v on y-faces at (i+1/2, j)/n,
 * pressure cell-centred), unit-square domains, n a power of two.
 *
 * Conventions
 *   - Every pointer argument is DEVICE memory unless the name ends in _host.
 *     Fields are nx*ny doubles, x fastest (idx = i + nx*j), as the reference
 *     stores them (ib/eulerian_grid.h:54-62).  Markers are m (x, y) pairs,
 *     interleaved.
 *   - Calls are ordered on the context's stream.  Functions that return an
 *     iteration count synchronise; the others are asynchronous.
 *   - Return value 0 = success; -1 invalid argument; -3 HIP error;
 *     -4 iteration limit reached (the reference loops forever instead).
 *     afm_last_error() describes the last failure of the calling thread.
 *   - One context per host thread; a context is not re-entrant.
 *
 * Each entry point names the reference interface it replaces (file:line in
 * the reference's src/ tree).
 */
#ifndef AFM_H
#define AFM_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct afm_ctx afm_ctx;
typedef struct afm_ns afm_ns;
typedef struct afm_ib afm_ib;
typedef struct afm_ns3 afm_ns3;

/* dt::geometry<Level, Width, Height, Depth> (dt/geometry.h:29-44), chosen at
 * run time.  dim must be 2; n = 2^Level >= 8; width = depth = 1. */
typedef struct afm_geom {
	int dim;
	unsigned n;
	unsigned width;
	unsigned depth;
} afm_geom;

/* Context: geometry, multigrid hierarchy and scratch on one device. */
int afm_ctx_create(const afm_geom* g, int device, afm_ctx** out);
int afm_ctx_destroy(afm_ctx* c);
int afm_ctx_info(afm_ctx* c, unsigned* nx, unsigned* ny, unsigned* levels, unsigned* coarse_level);
int afm_sync(afm_ctx* c);
/* Order the context's work on an existing HIP stream (NULL: HIP's null
 * stream, which is torch's default stream).  Pending work on the previous
 * stream is waited for first.  A context starts on its own non-blocking
 * stream, returned by afm_own_stream. */
int afm_set_stream(afm_ctx* c, void* hip_stream);
void* afm_own_stream(afm_ctx* c);
const char* afm_last_error(void);

/* Measurement (no reference counterpart): HIP-event timing, on the context's
 * stream, of the outer-update launch of afm_mg_solve (precision 0 and 2: the
 * finest-level post-smoothing fused with the update; 1 and 3: the CG update
 * of iterations after the first), counting only launches that ran.  afm_update_timing(c, 1) enables and resets; 0 disables. */
int afm_update_timing(afm_ctx* c, int enable);
int afm_update_timing_read(afm_ctx* c, double* avg_us, long* launches);
/* The same timing for three finest-level launches of afm_mg_solve above the
 * LDS levels (no reference counterpart): kind 0 the outer update (as
 * afm_update_timing_read), kind 1 the last post-smoothing of the CG forms'
 * preconditioner cycle (precision 1 and 3), kind 2 the pre-smoothing +
 * residual + restriction launch; launches of cycles that ran.  Enabled and
 * reset by afm_update_timing. */
int afm_kernel_timing_read(afm_ctx* c, int kind, double* avg_us, long* launches);

/* Device memory on the context's GPU for callers that hold no HIP headers
 * (the C++ drop-in classes of include/afm/).  Copies are synchronous. */
int afm_alloc(afm_ctx* c, unsigned long bytes, void** out);
int afm_free(afm_ctx* c, void* p);
int afm_copy_to_device(afm_ctx* c, void* dst, const void* src_host, unsigned long bytes);
int afm_copy_to_host(afm_ctx* c, void* dst_host, const void* src, unsigned long bytes);

/* ib::grid::delta_function (ib/grid.h:51-55) for n = the context's n:
 * out[i] = delta(r[i]).  afm_delta_host evaluates the same code on the CPU. */
int afm_delta(afm_ctx* c, const double* r, long count, double* out);
int afm_delta_host(const double* r_host, long count, unsigned n, double* out_host);

/* ib::grid::interpolate with marker targets (ib/grid.h:68-84 through
 * ib/mac_grid.h:49-73): U[k] = sum over the 4x4 window of
 * (delta_x delta_y h^2) field[idx].  comp 0 = u (x-faces), 1 = v (y-faces). */
int afm_interp(afm_ctx* c, int comp, const double* field, const double* X, long m, double* U);

/* ib::grid::interpolate with grid targets (ib/grid.h:68-84 with a
 * lagrangian_grid source, ib/lagrangian_grid.h:58-67): f = sum_k
 * (delta_x delta_y ds_k) F_k, summed per cell in ascending marker order
 * (bit-reproducible; no float atomics).  f is overwritten. */
int afm_spread(afm_ctx* c, int comp, const double* F, const double* ds, const double* X, long m, double* f);

/* mg::laplacian<G> * x (mg/laplacian.h:46-153, dt/matrix.h:81-94) at a
 * multigrid level (0 = 1x1 ... levels = finest). */
int afm_laplacian_apply(afm_ctx* c, unsigned level, const double* x, double* y);

/* Multigrid building blocks at a global level (n >= 64):
 * red-black Gauss-Seidel sweeps of L x = b (alg/smooth.h:35-51 per row),
 * x_out != x_in; full-weighting restriction (mg/solver.h:153-181);
 * bilinear prolongation, overwriting fine (mg/solver.h:71-151);
 * one V-cycle on the residual equation, rhs replaced by the correction
 * (mg/solver.h:199-216). */
int afm_mg_smooth(afm_ctx* c, unsigned level, const double* x_in, const double* b, double* x_out, int sweeps,
                  int zero_init);
int afm_mg_restrict(afm_ctx* c, unsigned level, const double* fine, double* coarse);
int afm_mg_prolong(afm_ctx* c, unsigned level, const double* coarse, double* fine);
int afm_mg_vcycle(afm_ctx* c, double* rhs_inout, int nu1, int nu2);

/* mg::solver<G>(tol, nu1, nu2).solve(rhs) (mg/solver.h:218-237): rhs is
 * replaced by the solution (its mean is arbitrary, as in the reference).
 * tol_mode 0: stop when RMS(rhs - L x) <= tol (the reference's test);
 * tol_mode 1: when RMS(r) <= tol * RMS(rhs - mean(rhs)).
 * precision 0: FP64 throughout, the reference's iteration (nested start,
 * then x += V(r) until the test holds);
 * 1: the mixed form -- flexible CG in FP64 (true residual, search
 * directions, inner products, solution) preconditioned by one FP32 V-cycle
 * per iteration, from x = 0; the coarsest levels of every cycle in FP64;
 * 2: the reference's iteration with FP32 V-cycles (FP64 iterative
 * refinement: FP64 residual, update, reductions and coarsest solve);
 * 3: as 1 with FP64 V-cycles.
 * Forms 1 and 3 reach the same residual test with fewer V-cycles; their
 * solution differs from the reference's by the arbitrary constant of the
 * periodic problem (the mean) and by at most the tolerance otherwise.
 * nu1, nu2 in {1, 2}.  *cycles = V-cycles (the reference's return; for 1
 * and 3 the CG iterations, one V-cycle each). */
int afm_mg_solve(afm_ctx* c, double* rhs_inout, double tol, int tol_mode, int nu1, int nu2, int precision,
                 int max_cycles, int* cycles, double* final_rms);

/* The reference's native 3-D geometry for the multigrid solve:
 * geometry<levels, width, width, width> (dt/geometry.h:29-60): a cube of
 * s = width * 2^levels cells per axis (width a power of two, 2 <= s <= 512),
 * laplacian scale n = 2^levels, periodic in x and z, walls (homogeneous
 * Neumann) in y -- mg::solver<G> (mg/solver.h:35-271); width 1 is the
 * geometry of the reference's test/mg_test3.cc, geometry<1,8,8,8> that of
 * test/mg_test2.cc.  Fields are x fastest, then y, then z.  The smoother is
 * red-black Gauss-Seidel with the reference's row arithmetic
 * (alg/smooth.h:35-51); restriction, prolongation, the nested start, the
 * solve loop and the coarsest level (mean removal and CG on width^3 cells,
 * mg/solver.h:264-271) are the reference's.  Device pointers on the
 * context's device and stream; level scratch is allocated per call (a parity
 * path, not tuned).
 *   laplacian3_apply: y = L x (mg/laplacian.h:46-153, dt/matrix.h:81-94);
 *   mg3_vcycle: one recurse(levels, rhs), rhs replaced by the result (tol:
 *     the coarsest level's CG tolerance, unused for width 1);
 *   mg3_solve: as afm_mg_solve's precision 0 on this geometry;
 *   mg3_restrict: mg::solver::restriction (mg/solver.h:153-181), level
 *     `levels` (s^3 values) to levels - 1;
 *   mg3_prolong: mg::solver::interpolation (mg/solver.h:71-151), levels - 1
 *     to levels, every fine value written;
 *   mg3_smooth: `sweeps` red-black sweeps of L x = b in place;
 *   laplacian3_cg: alg::conjugate_gradient (alg/cg.h:32-56) on L: x is the
 *     start and the result, stop when |r|_2 <= tol. */
int afm_laplacian3_apply(afm_ctx* c, unsigned levels, unsigned width, const double* x, double* y);
int afm_mg3_vcycle(afm_ctx* c, unsigned levels, unsigned width, double* rhs_inout, int nu1, int nu2, double tol);
int afm_mg3_solve(afm_ctx* c, unsigned levels, unsigned width, double* rhs_inout, double tol, int tol_mode, int nu1,
                  int nu2, int max_cycles, int* cycles, double* final_rms);
int afm_mg3_restrict(afm_ctx* c, unsigned levels, unsigned width, const double* fine, double* coarse);
int afm_mg3_prolong(afm_ctx* c, unsigned levels, unsigned width, const double* coarse, double* fine);
int afm_mg3_smooth(afm_ctx* c, unsigned levels, unsigned width, double* x, const double* b, int sweeps);
int afm_laplacian3_cg(afm_ctx* c, unsigned levels, unsigned width, const double* b, double* x, double tol,
                      int max_iters, int* iters, double* final_norm);

/* The reference's 3-D wall-bounded Navier-Stokes solver on the same geometry
 * (1 <= levels <= 8): ns::solver<geometry<levels,1,1,1>> (ns/solver.h:37-236)
 * with its projection (ns/projection.h:60-122).  u0, w0 hold n^3 x- and
 * z-face velocities, v0 the n*(n-1)*n interior y-face velocities; they are
 * aliased, projected at creation, and advanced in place by step(), which
 * alternates the backward-Euler half step and the Crank-Nicolson full step.
 * ut, ub, wt, wb are the n*n top and bottom wall values of u and w (index
 * x + n*z); step() adds projection::boundary's correction to them in place,
 * as the reference does to the caller's vectors.  fx, fz: n^3, fy:
 * n*(n-1)*n.  tol is the reference's absolute bound for the multigrid (RMS)
 * and the CG (|r|_2) solves; the multigrid smoother is red-black.  get_p:
 * p = H_s phi (ns/solver.h:83).  Level and solver scratch is allocated at
 * creation. */
int afm_ns3_create(afm_ctx* c, unsigned levels, double mu, double rho, double dt, double tol, double* u0, double* v0,
                   double* w0, afm_ns3** out);
int afm_ns3_step(afm_ns3* s, double* ut, double* ub, double* wt, double* wb, const double* fx, const double* fy,
                 const double* fz);
int afm_ns3_get_p(afm_ns3* s, double* p);
int afm_ns3_stats(afm_ns3* s, int* last_cycles, int* last_cg, int* have_half_stepped);
int afm_ns3_destroy(afm_ns3* s);
/* Its operators on the n^3 cube.  div3: alg::div (alg/div.h:29-58) with wall
 * values top/bottom (n*n); grad3: alg::grad (alg/grad.h:32-43), py has
 * n*(n-1)*n entries; advect3: ns::solver::advection (ns/solver.h:142-202);
 * helmholtz3: kind 1 ns::helmholtz_periodic (ns/helmholtz.h:61-175, n^3
 * rows), kind 3 ns::helmholtz_dirichlet (ns/helmholtz.h:218-319, n*(n-1)*n
 * rows), and alg::conjugate_gradient on it (alg/cg.h:32-56; x is the
 * initial guess and the result, stop when |r|_2 <= tol). */
int afm_div3(afm_ctx* c, unsigned levels, const double* u, const double* v, const double* w, const double* top,
             const double* bottom, double* d);
int afm_grad3(afm_ctx* c, unsigned levels, const double* p, double* px, double* py, double* pz);
int afm_advect3(afm_ctx* c, unsigned levels, const double* u, const double* v, const double* w, double* u_adv,
                double* v_adv, double* w_adv);
int afm_helmholtz3_apply(afm_ctx* c, unsigned levels, int kind, double scale, const double* x, double* y);
int afm_helmholtz3_cg(afm_ctx* c, unsigned levels, int kind, double scale, const double* b, double* x, double tol,
                      int max_iters, int* iters, double* final_norm);
/* helmholtz_periodic::boundary (kind 1, ns/helmholtz.h:177-196, n^3 values)
 * and helmholtz_dirichlet::boundary (kind 3, ns/helmholtz.h:321-337,
 * n*(n-1)*n values) of the wall vectors top, bottom (n*n) into out;
 * projection::boundary (ns/projection.h:75-106) with the gradient
 * components phx, phz (n^3) of the last projection's potential: ut, ub, wt,
 * wb (n*n) are corrected in place. */
int afm_helmholtz3_boundary(afm_ctx* c, unsigned levels, int kind, double scale, const double* top,
                            const double* bottom, double* out);
int afm_projection3_boundary(afm_ctx* c, unsigned levels, const double* phx, const double* phz, double* ut,
                             double* ub, double* wt, double* wb);

/* ib::grid::interpolate (ib/grid.h:68-84) on the same 3-D MAC grid, for
 * comp 0 (u), 1 (v: the mac_grid<y> window of ib/mac_grid.h:114-132 on the
 * n*(n-1)*n interior y-faces), 2 (w), or 3 (the node grid eulerian_grid of
 * ib/eulerian_grid.h:54-93, n^3 nodes at (i, j, k)/n); X holds m markers as
 * (x, y, z).
 * interp3: U[k] = sum over the 64 taps of (w vol) field[idx], in the
 * reference's tap order.  spread3: f = sum over markers of (w ds_k) F_k per
 * face, summed per face in marker order (run-to-run identical, no float
 * atomics); f is overwritten.  Taps the grid does not store -- the v
 * window's wall faces and the u / w windows' ghost row above the top wall,
 * which the reference indexes out of range -- read and receive nothing. */
int afm_interp3(afm_ctx* c, unsigned levels, int comp, const double* field, const double* X, long m, double* U);
int afm_spread3(afm_ctx* c, unsigned levels, int comp, const double* F, const double* ds, const double* X, long m,
                double* f);

/* A dt::matrix (dt/matrix.h:35-94) in CSR form in device memory -- rows + 1
 * row starts, column indices in ascending order per row, values -- for
 * operators that have no stencil kernel here (a matrix a program fills):
 *   csr_spmv: y = A x, each row summed from 0 in column order
 *     (dt/matrix.h:81-94);
 *   csr_cg: alg::conjugate_gradient (alg/cg.h:32-56): x is the start and the
 *     result, stop when |r|_2 <= tol, -4 past max_iters;
 *   csr_smooth: `sweeps` lexicographic Gauss-Seidel sweeps (alg/smooth.h:
 *     35-51) in the reference's row order, on one lane (not a fast path).
 * axpy: y += a x on `count` doubles (the drop-in classes' vector updates). */
int afm_csr_spmv(afm_ctx* c, unsigned long rows, const unsigned long* row_start, const unsigned long* column,
                 const double* values, const double* x, double* y);
int afm_csr_cg(afm_ctx* c, unsigned long rows, const unsigned long* row_start, const unsigned long* column,
               const double* values, const double* b, double* x, double tol, int max_iters, int* iters,
               double* final_norm);
int afm_csr_smooth(afm_ctx* c, unsigned long rows, const unsigned long* row_start, const unsigned long* column,
                   const double* values, double* x, const double* b, int sweeps);
int afm_axpy(afm_ctx* c, unsigned long count, double a, const double* x, double* y);

/* (I + scale L5) x (ns/helmholtz.h:61-175 for the periodic plane) and
 * alg::conjugate_gradient on it (alg/cg.h:32-56): x is the initial guess and
 * the result.  tol_mode 0: stop when |r|_2 <= tol (the reference's test);
 * 1: when |r|_2 <= tol |b|_2. */
int afm_helmholtz_apply(afm_ctx* c, double scale, const double* x, double* y);
int afm_helmholtz_cg(afm_ctx* c, double scale, const double* b, double* x, double tol, int tol_mode, int max_iters,
                     int* iters, double* final_norm);
/* afm_helmholtz_cg without any allocation (no counterpart in the reference:
 * for callers that cannot allocate device memory -- code that prepares HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * CG's residual and search directions live in a buffer the library holds on
 * each GPU for the life of the process (3 * 2048^2 doubles, 96 MiB, set up
 * when the first context on that GPU is created), so the grid must have at
 * most 2048 x 2048 cells (-1 otherwise, with nothing run).  Nothing in the
 * buffer carries over between calls, but calls that overlap on one GPU
 * (from several contexts or threads) race on it.  Same arguments, kernels
 * and results as afm_helmholtz_cg. */
int afm_helmholtz_cg_noalloc(afm_ctx* c, double scale, const double* b, double* x, double tol, int tol_mode,
                             int max_iters, int* iters, double* final_norm);

/* alg::div (alg/div.h:29-60), alg::grad + subtraction (alg/grad.h:29-44,
 * ns/projection.h:69-72), ns::solver::advection (ns/solver.h:142-236). */
int afm_div(afm_ctx* c, const double* u, const double* v, double* d);
int afm_grad_sub(afm_ctx* c, const double* p, double* u, double* v);
int afm_advect(afm_ctx* c, const double* u, const double* v, double* u_adv, double* v_adv);

/* ns::solver<G>(mu, rho, dt, tol, u0, v0, w0) (ns/solver.h:61-72): u0, v0 are
 * aliased and projected at creation; step() advances them in place
 * (ns/solver.h:90-140: the first call is the backward-Euler half step, the
 * second the Crank-Nicolson full step, alternating).  tol/tol_mode apply to
 * both the multigrid and the CG solves; precision as for afm_mg_solve. */
int afm_ns_create(afm_ctx* c, double mu, double rho, double dt, double tol, int tol_mode, int precision, double* u0,
                  double* v0, afm_ns** out);
int afm_ns_step(afm_ns* s, const double* fx, const double* fy);
int afm_ns_project(afm_ns* s, double* u, double* v);             /* ns/projection.h:60-73 */
int afm_ns_get_p(afm_ns* s, double* p);                         /* ns/solver.h:86 */
int afm_ns_stats(afm_ns* s, int* last_cycles, int* last_cg_u, int* last_cg_v, int* have_half_stepped);
int afm_ns_old(afm_ns* s, const double** u_old, const double** v_old);
/* The last projection's multigrid solve (no reference counterpart: the
 * reference's solve returns only its cycle count, mg/solver.h:218-237):
 * *final_rms = RMS(div - L phi) of the returned potential, which is the RMS
 * divergence left in the projected velocity; *threshold = the bound it was
 * held to (tol, or tol * RMS(div - mean(div)) for tol_mode 1). */
int afm_ns_last_residual(afm_ns* s, double* final_rms, double* threshold);
/* The CG forms' start potentials (precision 1 and 3; no reference
 * counterpart: the reference's solve starts from its nested iteration).
 * Such a solver starts a step's projection from the potentials of the
 * previous steps of the same kind -- the backward-Euler half step's or the
 * Crank-Nicolson full step's: kind 0 / 1 is the next start of the half /
 * full step (the linear extrapolation 2 phi_n - phi_{n-1} of the kind's
 * last two potentials; phi_n alone after the first), kind 2 / 3 the kind's
 * last potential phi_n.  get copies one to phi (n x n on the device) and
 * sets *valid (0: unset, phi untouched); set copies phi in (NULL: unset).
 * Setting kind 0 / 1 unsets kind 2 / 3 (a start alone has nothing to
 * extrapolate from); set 2 / 3 after it.  A checkpoint of such a solver
 * saves and restores all four to continue bit for bit. */
int afm_ns_get_start(afm_ns* s, int kind, double* phi, int* valid);
int afm_ns_set_start(afm_ns* s, int kind, const double* phi);
int afm_ns_destroy(afm_ns* s);

/* Immersed-boundary time stepping (no reference counterpart: the reference
 * stops at the coupling comment of test/ib_ns_test.cc:80-90).  Markers with
 * tether springs F = -k (X - R) and arclength weights ds; one step is
 *   f = S(X) F(X);  BE half step;  X_h = X + dt/2 J(X) u;
 *   f = S(X_h) F(X_h);  CN full step;  X += dt/2 (J(X_h) u_old + J(X_h) u).
 * X, R, k, ds are copied at creation. */
int afm_ib_create(afm_ns* s, long m, const double* X, const double* R, const double* k, const double* ds,
                  afm_ib** out);
int afm_ib_step(afm_ib* b, int steps);
int afm_ib_markers(afm_ib* b, double* X_out);

/* Differentiation (no reference counterpart).  Runs `steps` IB steps from the
 * current state (which must be at a full step) and evaluates
 *   J = 1/2 sum (u^2 + v^2) h^2 + sum_m |c_m(T) - c_m(0)|^2,
 * c_m the centroid of membrane m = markers [m*msize, (m+1)*msize).  With
 * grad_k / grad_R non-NULL (device, m and 2m doubles) it also returns dJ/dk
 * and dJ/dR by the discrete adjoint of the step (adjoint transfer kernels,
 * the same symmetric CG and multigrid solves), checkpointing (u, v, X) every
 * `checkpoint_every` steps and recomputing each segment in the reverse
 * sweep.  *J_host receives J.  The solver is left at the final state. */
int afm_ib_grad(afm_ib* b, int steps, int checkpoint_every, long msize, double* J_host, double* grad_k,
                double* grad_R);
int afm_ib_destroy(afm_ib* b);

#ifdef __cplusplus
}
#endif
#endif
