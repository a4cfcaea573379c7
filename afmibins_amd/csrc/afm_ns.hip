// afm_ns.hip -- the Navier-Stokes step around the two solves, on gfx950.
//
//   k_advect_cg_init  conservative centred advection (ns/solver.h:204-234,
//                  the 2-D branch) fused with the right-hand-side assembly of
//                  the backward-Euler half step or Crank-Nicolson full step
//                  (ns/solver.h:105-135) and the start of both CG solves;
//   k_cg_*         CG on the periodic Helmholtz operator I + s L
//                  (alg/cg.h:32-56, ns/helmholtz.h), both velocity components
//                  in one launch (blockIdx.y = component), the stencil
//                  product in the CSR entry order, dot products reduced
//                  deterministically, the stop test on the GPU;
//   k_div, k_grad_sub   projection (ns/projection.h:60-73, alg/div.h,
//                  alg/grad.h) around the multigrid solve.
#include <mutex>
#include <string>

#include "afm_device.h"

namespace afm {

// ---------------------------------------------------------------- advection
// Cell indices are 32-bit (fields have fewer than 2^32 cells): the 64-bit
// index arithmetic of the neighbour lookups was most of the kernel's work.
// Each value is loaded once; every sum keeps the operand order of
// uu(j) = (u[j] + u[xm j]) / 2 squared, ww(j) likewise in z, and
// uw(j) = (u[zm j] + u[j]) / 2 * (w[xm j] + w[j]) / 2.
// The advection terms of one cell from its neighbours' values: u at the
// cell, x - 1, x + 1, z - 1, z + 1 and (x + 1, z - 1); w at the cell, x - 1,
// x + 1, z - 1, z + 1 and (x - 1, z + 1).
__device__ inline void adv_vals(double ui, double uxm, double uxp, double uzm, double uzp, double uxpzm, double wi,
                                double wxm, double wxp, double wzm, double wzp, double wzpxm, unsigned n, double& ua,
                                double& wa)
{
	auto sq = [](double q) { return q * q; };
	const double uu_xp = sq(0.5 * (uxp + ui)), uu_i = sq(0.5 * (ui + uxm));
	const double ww_zp = sq(0.5 * (wzp + wi)), ww_i = sq(0.5 * (wi + wzm));
	const double uw_i = (0.5 * (uzm + ui)) * (0.5 * (wxm + wi));
	const double uw_zp = (0.5 * (ui + uzp)) * (0.5 * (wzpxm + wzp));
	const double uw_xp = (0.5 * (uxpzm + uxp)) * (0.5 * (wi + wxp));
	const double dv = 0.0 - 0.0;   // alg/div.h:54 with zero top/bottom
	ua = n * ((uu_xp - uu_i) + dv + (uw_zp - uw_i));
	wa = n * ((uw_xp - uw_i) + dv + (ww_zp - ww_i));
}

__device__ inline void adv_at(const double* __restrict__ u, const double* __restrict__ w, int nx, int ny, unsigned n,
                              unsigned i, int x, int z, double& ua, double& wa)
{
	const unsigned N = (unsigned) nx * (unsigned) ny, unx = (unsigned) nx;
	const unsigned ixm = x == 0 ? i + (unx - 1) : i - 1, ixp = x == nx - 1 ? i - (unx - 1) : i + 1;
	const unsigned izm = z == 0 ? i + (N - unx) : i - unx, izp = z == ny - 1 ? i - (N - unx) : i + unx;
	const unsigned ixp_zm = z == 0 ? ixp + (N - unx) : ixp - unx;   // zm(xp(i))
	const unsigned izp_xm = x == 0 ? izp + (unx - 1) : izp - 1;     // xm(zp(i))
	adv_vals(u[i], u[ixm], u[ixp], u[izm], u[izp], u[ixp_zm], w[i], w[ixm], w[ixp], w[izm], w[izp], w[izp_xm], n, ua,
	         wa);
}

__global__ void k_advect(const double* __restrict__ u, const double* __restrict__ v, double* __restrict__ ua,
                         double* __restrict__ va, int nx, int ny, unsigned n)
{
	size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i >= (size_t) nx * ny) return;
	int gx, gy;
	cell_xy(i, nx, gx, gy);
	adv_at(u, v, nx, ny, n, i, gx, gy, ua[i], va[i]);
}

// d = n ((u[x+1] - u) + dv + (v[y+1] - v))   (alg/div.h:38-59)
__global__ void k_div(const double* __restrict__ u, const double* __restrict__ v, double* __restrict__ d, int nx,
                      int ny, unsigned n)
{
	const size_t N = (size_t) nx * ny;
	// consecutive blocks per XCD: a row's z + 1 neighbours come from the same
	// L2 (17.3 -> 16.7 us at 2048^2; the gradient update lost 8% so)
	size_t i = xcd_remap(blockIdx.x, gridDim.x) * (size_t) blockDim.x + threadIdx.x;
	if (i >= N) return;
	int x, z;
	cell_xy(i, nx, x, z);
	const size_t ixp = x == nx - 1 ? i - (nx - 1) : i + 1, izp = z == ny - 1 ? (size_t) x : i + nx;
	const double dv = 0.0 - 0.0;
	d[i] = n * ((u[ixp] - u[i]) + dv + (v[izp] - v[i]));
}

// k_div fused with the first launch of the projection's solve (mg_solve's
// k_mean, pass 0): the divergence on cell pairs along x (grid-stride over
// the reduction grid, consecutive blocks per XCD), each cell by k_div's
// formula (the same bits), and per block the sum of d * 1 in cell order for
// the mean; block 0 starts the solve's control block as k_mean does.
__global__ __launch_bounds__(256) void k_div_sum(const double* __restrict__ u, const double* __restrict__ v,
                                                 double* __restrict__ d, int nx, int ny, unsigned n,
                                                 SolveCtl* __restrict__ ctl, double* __restrict__ partials,
                                                 double tol, int max_iter)
{
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		ctl->tol = tol;
		ctl->max_iter = max_iter;
		ctl->count = 0;
		ctl->done = 0;
		ctl->err = 0;
		ctl->res = 0.0;
	}
	const unsigned N = (unsigned) nx * (unsigned) ny, unx = (unsigned) nx, P = N / 2;
	double acc[1] = {0.0};
	for (unsigned k = xcd_remap(blockIdx.x, gridDim.x) * 256 + threadIdx.x; k < P; k += gridDim.x * 256) {
		const unsigned i = 2 * k;
		const int x = (int) (i & (unx - 1)), z = (int) (i >> __builtin_ctz(unx));
		const unsigned rz = i - (unsigned) x;
		const unsigned rp = z == ny - 1 ? rz - (N - unx) : rz + unx;
		const unsigned x2 = x + 2 == nx ? 0u : (unsigned) x + 2;
		const double2 U = *(const double2*) (u + i), V = *(const double2*) (v + i),
		              Vp = *(const double2*) (v + rp + x);
		const double uE = u[rz + x2];
		const double dv = 0.0 - 0.0;
		const double d0 = n * ((U.y - U.x) + dv + (Vp.x - V.x));
		const double d1 = n * ((uE - U.y) + dv + (Vp.y - V.y));
		*(double2*) (d + i) = make_double2(d0, d1);
		acc[0] += d0 * 1.0;
		acc[0] += d1 * 1.0;
	}
	reduce_partial<1>(acc, partials);
}

// u -= n (p - p[x-1]),  v -= n (p - p[y-1])   (alg/grad.h:36-43, ns/projection.h:70-72)
// u -= n (p - p_west), v -= n (p - p_south); with pcopy, p is also copied
// there (the projection's potential kept for get_p)
// u -= grad p, v likewise; pcopy (when given) receives p.  xnext (the CG
// forms' next start of this step kind, when given) receives the linear
// extrapolation 2 p - pcopy_old from the previous potential of the kind
// (extrap), else p itself (2 p - p = p exactly).
__global__ void k_grad_sub(const double* __restrict__ p, double* __restrict__ u, double* __restrict__ v, int nx, int ny,
                           unsigned n, double* __restrict__ pcopy, double* __restrict__ xnext, int extrap)
{
	const size_t N = (size_t) nx * ny;
	size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i >= N) return;
	int x, z;
	cell_xy(i, nx, x, z);
	const size_t ixm = x == 0 ? i + (nx - 1) : i - 1, izm = z == 0 ? i + N - nx : i - nx;
	const double px = n * (p[i] - p[ixm]);
	const double pz = n * (p[i] - p[izm]);
	u[i] = u[i] - px;
	v[i] = v[i] - pz;
	if (pcopy) {
		const double pv = p[i];
		const double old = extrap ? pcopy[i] : pv;
		pcopy[i] = pv;
		if (xnext) xnext[i] = 2.0 * pv - old;
	}
}

__global__ void k_helm_apply(const double* __restrict__ x, double* __restrict__ y, int nx, int ny, double off, double dg)
{
	const size_t N = (size_t) nx * ny;
	size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i >= N) return;
	int gx, gy;
	cell_xy(i, nx, gx, gy);
	const size_t ie = gx == nx - 1 ? i - (nx - 1) : i + 1, iw = gx == 0 ? i + (nx - 1) : i - 1;
	const size_t in_ = gy == ny - 1 ? (size_t) gx : i + nx, is = gy == 0 ? i + N - nx : i - nx;
	y[i] = stencil5<double, true>(off, dg, x[i], x[ie], x[iw], x[in_], x[is], gx == nx - 1, gx == 0, gy == ny - 1, gy == 0);
}

// ---------------------------------------------------------------------- CG
// Per-field pointers, indexed by blockIdx.y (blockIdx.x in the one-block
// finalize kernels).  Partial sums per field: the start writes two values
// per block at pu (r.r) and pu + MAX_RED_BLOCKS (b.b), the p update one at
// pp, the x / r update one at pu -- so a kernel never overwrites the
// partials that it (or its finaliser step) is still reading.
struct CGArgs {
	double* x[2];
	const double* b[2];
	double* r[2];
	double* p[2][2];       // ping-pong
	SolveCtl* ctl[2];
	double* pu[2];
	double* pp[2];
	int tol_mode;
};

#define CG_NEIGHBOURS                                                                              \
	const int gx = (int) (i & (unsigned) (nx - 1)), gy = (int) (i >> __builtin_ctz((unsigned) nx));    \
	const unsigned ie = gx == nx - 1 ? i - (nx - 1) : i + 1, iw = gx == 0 ? i + (nx - 1) : i - 1;  \
	const unsigned in_ = gy == ny - 1 ? (unsigned) gx : i + nx, is = gy == 0 ? i + N - nx : i - nx; \
	const bool we = gx == nx - 1, ww = gx == 0, wn = gy == ny - 1, ws = gy == 0;

// Start field f's control block (block 0 of the start launch; the stop test
// of a field not being solved reads as fired).
__device__ inline void cg_ctl_start(CGArgs& a, int f, double tol, int max_iters)
{
	SolveCtl* c = a.ctl[f];
	c->tol = tol;
	c->max_iter = max_iters;
	c->count = 0;
	c->done = a.x[f] ? 0 : 1;
	c->err = 0;
}

// r = b - A x, n0 = r.r  (alg/cg.h:38-39); |b|^2 for a relative tol.  p = r
// is not stored: the first p update reads r.
__global__ __launch_bounds__(256) void k_cg_init(CGArgs a, int nx, int ny, double off, double dg, double tol,
                                                 int max_iters)
{
	const int f = blockIdx.y;
	if (blockIdx.x == 0 && threadIdx.x == 0) cg_ctl_start(a, f, tol, max_iters);
	if (!a.x[f]) return;
	const size_t N = (size_t) nx * ny;
	const double* x = a.x[f];
	const double* b = a.b[f];
	double acc[2] = {0.0, 0.0};
	for (unsigned i = blockIdx.x * 256 + threadIdx.x; i < (unsigned) N; i += gridDim.x * 256) {
		CG_NEIGHBOURS
		double Ax = stencil5<double, true>(off, dg, x[i], x[ie], x[iw], x[in_], x[is], we, ww, wn, ws);
		double r = b[i] - Ax;
		a.r[f][i] = r;
		acc[0] += r * r;
		acc[1] += b[i] * b[i];
	}
	reduce_partial<2>(acc, a.pu[f]);
}

__global__ __launch_bounds__(256) void k_cg_init_fin(CGArgs a, unsigned nblk)
{
	const int f = blockIdx.x;
	if (!a.x[f]) return;
	double tot[2];
	red_total<2>(a.pu[f], nblk, tot);
	if (threadIdx.x == 0) {
		SolveCtl* c = a.ctl[f];
		if (a.tol_mode == 1) c->tol = c->tol * sqrt(tot[1]);
		c->n0 = tot[0];
		c->n0p[0] = tot[0];
		c->count = 0;
		c->res = sqrt(tot[0]);
		c->done = !(sqrt(tot[0]) > c->tol && 0 < c->max_iter);
	}
}

// the stop flag as the block found it at its start, the same in all its waves
// (with fold, block 0 of the same launch may set it meanwhile)
__device__ inline bool block_done(const SolveCtl* c)
{
	__shared__ int d;
	if (threadIdx.x == 0) d = c->done;
	__syncthreads();
	return d != 0;
}

// the sum of a kernel's per-block partials, in every thread of the block
// (red_total's fixed order, so every block, and the one-block finalisers,
// get the same bits)
__device__ inline double block_total(const double* partials, unsigned nblk)
{
	__shared__ double t;
	double tot[1];
	red_total<1>(partials, nblk, tot);
	if (threadIdx.x == 0) t = tot[0];
	__syncthreads();
	return t;
}

// The stop test after update `it` (alg/cg.h:42, 48-50): n1 = r.r from the
// update's partials, mu = n1 / n0.  Written to ctl by one thread; n1 goes to
// n0p[it & 1] while the old n0 is read from n0p[(it - 1) & 1], so the other
// blocks of the same launch still read the old value.
__device__ inline bool cg_stop_test(CGArgs& a, int f, int it, double n1, double& mu, bool writer)
{
	SolveCtl* c = a.ctl[f];
	mu = n1 / c->n0p[(it - 1) & 1];
	const bool done = !(sqrt(n1) > c->tol && it < c->max_iter);
	if (writer) {
		c->mu = mu;
		c->n0 = n1;
		c->n0p[it & 1] = n1;
		c->count = it;
		c->res = sqrt(n1);
		c->done = done;
	}
	return done;
}

// Iteration `it` (0-based), first half: p' = (it == 0) ? r : p*mu + r
// (alg/cg.h:39 and 51-52 of the previous iteration), stored; partial p'.(A p').
// With fold, every block first does the previous iteration's stop test from
// the update's partials (block 0 records it), instead of a separate launch.
// TP, TR: the types p and r are stored in -- double for the reference's
// iteration; float in the mixed form, whose p' and r are formed in FP64 and
// stored rounded (every thread forms its neighbours' p' from the same stored
// values, and x and r are both updated with the stored p, so r stays b - A x
// up to the roundings of r, about 1e-7 of the residuals before it; the step
// length sees the rounding of p, about 1e-7 relative; r.r and the stop test
// take each new residual before its rounding).
template<typename TP, typename TR>
__global__ __launch_bounds__(256) void k_cg_p(CGArgs a, int nx, int ny, double off, double dg, int pin, int it,
                                              int fold, unsigned nblk)
{
	const int f = blockIdx.y;
	if (!a.x[f]) return;
	SolveCtl* c = a.ctl[f];
	if (block_done(c)) return;
	double mu = c->mu;
	if (fold) {
		const double n1 = block_total(a.pu[f], nblk);
		if (cg_stop_test(a, f, it, n1, mu, blockIdx.x == 0 && threadIdx.x == 0)) return;
	}
	const size_t N = (size_t) nx * ny;
	const TP* p = (const TP*) a.p[f][pin];
	TP* po = (TP*) a.p[f][pin ^ 1];
	const TR* r = (const TR*) a.r[f];
	const bool first = it == 0;
	double acc[1] = {0.0};
	for (unsigned i = blockIdx.x * 256 + threadIdx.x; i < (unsigned) N; i += gridDim.x * 256) {
		CG_NEIGHBOURS
		double pc, pe, pw, pn, ps;
		if (first) { pc = r[i]; pe = r[ie]; pw = r[iw]; pn = r[in_]; ps = r[is]; }   // p = r
		else {
			pc = (double) p[i] * mu + (double) r[i]; pe = (double) p[ie] * mu + (double) r[ie];
			pw = (double) p[iw] * mu + (double) r[iw]; pn = (double) p[in_] * mu + (double) r[in_];
			ps = (double) p[is] * mu + (double) r[is];
		}
		double z = stencil5<double, true>(off, dg, pc, pe, pw, pn, ps, we, ww, wn, ws);
		po[i] = (TP) pc;
		acc[0] += pc * z;
	}
	reduce_partial<1>(acc, a.pp[f]);
}

// Second half: nu = n0 / (p.z) from the p kernel's partials (every block,
// same bits), z = A p (recomputed, same bits), x += p nu, r -= z nu;
// partial r.r
template<typename TP, typename TR>
__global__ __launch_bounds__(256) void k_cg_update(CGArgs a, int nx, int ny, double off, double dg, int pcur, int it,
                                                   unsigned nblk)
{
	const int f = blockIdx.y;
	if (!a.x[f]) return;
	SolveCtl* c = a.ctl[f];
	if (c->done) return;
	const double nu = c->n0p[it & 1] / block_total(a.pp[f], nblk);
	const size_t N = (size_t) nx * ny;
	double* x = a.x[f];
	TR* r = (TR*) a.r[f];
	const TP* p = (const TP*) a.p[f][pcur];
	double acc[1] = {0.0};
	for (unsigned i = blockIdx.x * 256 + threadIdx.x; i < (unsigned) N; i += gridDim.x * 256) {
		CG_NEIGHBOURS
		const double pc = (double) p[i];
		double z = stencil5<double, true>(off, dg, pc, (double) p[ie], (double) p[iw], (double) p[in_], (double) p[is],
		                                  we, ww, wn, ws);
		x[i] += pc * nu;
		z *= nu;
		double ri = (double) r[i] - z;
		r[i] = (TR) ri;
		acc[0] += ri * ri;
	}
	reduce_partial<1>(acc, a.pu[f]);
}

// the stop test after update `it`, as its own launch (the last iteration
// of a batch, before the host reads the control block)
__global__ __launch_bounds__(256) void k_cg_update_fin(CGArgs a, int it, unsigned nblk)
{
	const int f = blockIdx.x;
	if (!a.x[f] || a.ctl[f]->done) return;
	const double n1 = block_total(a.pu[f], nblk);
	double mu;
	cg_stop_test(a, f, it, n1, mu, threadIdx.x == 0);
}

__global__ __launch_bounds__(256) void k_sum(const double* __restrict__ x, size_t N, double* partials)
{
	double acc[1] = {0.0};
	for (size_t i = blockIdx.x * 256 + threadIdx.x; i < N; i += (size_t) gridDim.x * 256) acc[0] += x[i];
	reduce_partial<1>(acc, partials);
}

__global__ __launch_bounds__(256) void k_sum_fin(unsigned nblk, SolveCtl* ctl, const double* partials)
{
	double tot[1];
	red_total<1>(partials, nblk, tot);
	if (threadIdx.x == 0) ctl->sum2 = tot[0];
}

// k_advect_rhs fused with k_cg_init for the two velocity solves of a step:
// the right-hand side b = rx (or ry) is formed per cell and never stored;
// r = b - A u with u the initial guess (ns/solver.h:105-135 then
// alg/cg.h:38-39), partial sums of r.r and b.b per field.  With STORE_B the
// right-hand sides are stored to bu, bv instead and only b.b is summed (the
// polynomial start, k_cg_start_fused, forms x0 and r; the r.r partials are
// written as zeros here).
template<bool STORE_B>
__global__ __launch_bounds__(256) void k_advect_cg_init(CGArgs a, const double* __restrict__ fx,
                                                        const double* __restrict__ fy, double* __restrict__ uo,
                                                        double* __restrict__ vo, int nx, int ny, unsigned n, double dt,
                                                        double rho, int be, double hoff, double hdg, double off,
                                                        double dg, double tol, int max_iters, double* __restrict__ bu,
                                                        double* __restrict__ bv)
{
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		cg_ctl_start(a, 0, tol, max_iters);
		cg_ctl_start(a, 1, tol, max_iters);
	}
	// 32-bit cell indices (fields have fewer than 2^32 cells)
	const unsigned N = (unsigned) nx * (unsigned) ny, unx = (unsigned) nx;
	const double* u = a.x[0];
	const double* v = a.x[1];
	double acc_u[2] = {0.0, 0.0}, acc_v[2] = {0.0, 0.0};
	for (unsigned i = blockIdx.x * 256 + threadIdx.x; i < N; i += gridDim.x * 256) {
		const int x = (int) (i & (unx - 1)), z = (int) (i >> __builtin_ctz(unx));
		double ua, va;
		adv_at(u, v, nx, ny, n, i, x, z, ua, va);
		const unsigned ie = x == nx - 1 ? i - (unx - 1) : i + 1, iw = x == 0 ? i + (unx - 1) : i - 1;
		const unsigned in_ = z == ny - 1 ? (unsigned) x : i + unx, is = z == 0 ? i + (N - unx) : i - unx;
		const bool we = x == nx - 1, ww = x == 0, wn = z == ny - 1, ws = z == 0;
		const double u0 = u[i], v0 = v[i];
		double rx, ry;
		if (be) {
			uo[i] = u0;
			vo[i] = v0;
			rx = u0 + (dt / 2) * (1 / rho * fx[i] - ua);
			ry = v0 + (dt / 2) * (1 / rho * fy[i] - va);
		} else {
			double hu = stencil5<double, true>(hoff, hdg, uo[i], uo[ie], uo[iw], uo[in_], uo[is], we, ww, wn, ws);
			double hv = stencil5<double, true>(hoff, hdg, vo[i], vo[ie], vo[iw], vo[in_], vo[is], we, ww, wn, ws);
			rx = hu + dt * (1 / rho * fx[i] - ua);
			ry = hv + dt * (1 / rho * fy[i] - va);
		}
		if (STORE_B) {
			bu[i] = rx;
			bv[i] = ry;
			acc_u[1] += rx * rx;
			acc_v[1] += ry * ry;
			continue;
		}
		const double Au = stencil5<double, true>(off, dg, u0, u[ie], u[iw], u[in_], u[is], we, ww, wn, ws);
		const double Av = stencil5<double, true>(off, dg, v0, v[ie], v[iw], v[in_], v[is], we, ww, wn, ws);
		const double r_u = rx - Au, r_v = ry - Av;
		a.r[0][i] = r_u;
		a.r[1][i] = r_v;
		acc_u[0] += r_u * r_u;
		acc_u[1] += rx * rx;
		acc_v[0] += r_v * r_v;
		acc_v[1] += ry * ry;
	}
	reduce_partial<2>(acc_u, a.pu[0]);
	reduce_partial<2>(acc_v, a.pu[1]);
}

// k_advect_cg_init on runs of C cells along x (nx a multiple of C): thread
// k forms cells Ck .. Ck + C - 1 of a row from 16-byte loads of the run in
// rows z - 1, z and z + 1 plus the single neighbours outside it, each value
// read once, and stores the run's results as 16-byte vectors: a fraction of
// the memory instructions and index arithmetic of one cell per thread.
// Every cell takes the arithmetic of adv_vals / stencil5 with its own wrap
// flags, so each cell's b and r are the same bits as k_advect_cg_init's; the
// r.r and b.b partials add the run's terms in cell order.
constexpr int ADV_BS = 256;   // threads per block of the right-hand-side launch

template<bool STORE_B, int C>
__global__ __launch_bounds__(ADV_BS) void k_advect_cg_init_runs(CGArgs a, const double* __restrict__ fx,
                                                             const double* __restrict__ fy, double* __restrict__ uo,
                                                             double* __restrict__ vo, int nx, int ny, unsigned n,
                                                             double dt, double rho, int be, double hoff, double hdg,
                                                             double off, double dg, double tol, int max_iters,
                                                             double* __restrict__ bu, double* __restrict__ bv)
{
	static_assert(C % 2 == 0, "runs of whole 16-byte vectors");
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		cg_ctl_start(a, 0, tol, max_iters);
		cg_ctl_start(a, 1, tol, max_iters);
	}
	const unsigned N = (unsigned) nx * (unsigned) ny, unx = (unsigned) nx, P = N / C;
	const double* __restrict__ u = a.x[0];
	const double* __restrict__ v = a.x[1];
	double acc_u[2] = {0.0, 0.0}, acc_v[2] = {0.0, 0.0};
	// the run's values with one extra slot each side: [0] = x - 1, [C + 1] = x + C
	auto load_run = [&](const double* __restrict__ f, unsigned row, int x, double (&r)[C]) {
#pragma unroll
		for (int j = 0; j < C; j += 2) {
			const double2 t = *(const double2*) (f + row + x + j);
			r[j] = t.x; r[j + 1] = t.y;
		}
	};
	// consecutive block indices on each XCD (blocks are dealt to the 8 XCDs
	// in turn): a run's neighbour rows are mostly read by the same XCD, from
	// its own L2 (72.0 -> 69.5 us at 2048^2)
	const unsigned lb = xcd_tile(blockIdx.x, gridDim.x);
	for (unsigned k = lb * ADV_BS + threadIdx.x; k < P; k += gridDim.x * ADV_BS) {
		const unsigned i = C * k;
		const int x = (int) (i & (unx - 1)), z = (int) (i >> __builtin_ctz(unx));
		const unsigned rz = i - (unsigned) x;                                 // row z
		const unsigned rm = z == 0 ? rz + (N - unx) : rz - unx;               // row z - 1
		const unsigned rp = z == ny - 1 ? rz - (N - unx) : rz + unx;          // row z + 1
		const unsigned xm = x == 0 ? unx - 1 : (unsigned) x - 1;              // column x - 1
		const unsigned xc = x + C == nx ? 0u : (unsigned) x + C;              // column x + C
		double U[C], Um[C], Up[C], W[C], Wm[C], Wp[C];
		load_run(u, rz, x, U); load_run(u, rm, x, Um); load_run(u, rp, x, Up);
		load_run(v, rz, x, W); load_run(v, rm, x, Wm); load_run(v, rp, x, Wp);
		const double uW = u[rz + xm], uE = u[rz + xc], umE = u[rm + xc];
		const double wW = v[rz + xm], wE = v[rz + xc], wpW = v[rp + xm];
		double ua[C], va[C];
#pragma unroll
		for (int j = 0; j < C; ++j) {
			const double uxm = j ? U[j - 1] : uW, uxp = j < C - 1 ? U[j + 1] : uE;
			const double uxpzm = j < C - 1 ? Um[j + 1] : umE;
			const double wxm = j ? W[j - 1] : wW, wxp = j < C - 1 ? W[j + 1] : wE;
			const double wzpxm = j ? Wp[j - 1] : wpW;
			adv_vals(U[j], uxm, uxp, Um[j], Up[j], uxpzm, W[j], wxm, wxp, Wm[j], Wp[j], wzpxm, n, ua[j], va[j]);
		}
		double F[C], G[C];
		load_run(fx, rz, x, F); load_run(fy, rz, x, G);
		const bool wn = z == ny - 1, ws = z == 0;
		double rx[C], ry[C];
		if (be) {
#pragma unroll
			for (int j = 0; j < C; j += 2) {
				*(double2*) (uo + i + j) = make_double2(U[j], U[j + 1]);
				*(double2*) (vo + i + j) = make_double2(W[j], W[j + 1]);
			}
#pragma unroll
			for (int j = 0; j < C; ++j) {
				rx[j] = U[j] + (dt / 2) * (1 / rho * F[j] - ua[j]);
				ry[j] = W[j] + (dt / 2) * (1 / rho * G[j] - va[j]);
			}
		} else {
			double O[C], Om[C], Op[C], Q[C], Qm[C], Qp[C];
			load_run(uo, rz, x, O); load_run(uo, rm, x, Om); load_run(uo, rp, x, Op);
			load_run(vo, rz, x, Q); load_run(vo, rm, x, Qm); load_run(vo, rp, x, Qp);
			const double oW = uo[rz + xm], oE = uo[rz + xc], qW = vo[rz + xm], qE = vo[rz + xc];
#pragma unroll
			for (int j = 0; j < C; ++j) {
				// wrap flags: x = 0 only for the run's first cell, x = nx - 1 only for its last
				const bool we = j == C - 1 && x + C == nx, ww = j == 0 && x == 0;
				const double hu = stencil5<double, true>(hoff, hdg, O[j], j < C - 1 ? O[j + 1] : oE,
				                                         j ? O[j - 1] : oW, Op[j], Om[j], we, ww, wn, ws);
				const double hv = stencil5<double, true>(hoff, hdg, Q[j], j < C - 1 ? Q[j + 1] : qE,
				                                         j ? Q[j - 1] : qW, Qp[j], Qm[j], we, ww, wn, ws);
				rx[j] = hu + dt * (1 / rho * F[j] - ua[j]);
				ry[j] = hv + dt * (1 / rho * G[j] - va[j]);
			}
		}
		if (STORE_B) {
#pragma unroll
			for (int j = 0; j < C; j += 2) {
				*(double2*) (bu + i + j) = make_double2(rx[j], rx[j + 1]);
				*(double2*) (bv + i + j) = make_double2(ry[j], ry[j + 1]);
			}
#pragma unroll
			for (int j = 0; j < C; ++j) {
				acc_u[1] += rx[j] * rx[j];
				acc_v[1] += ry[j] * ry[j];
			}
			continue;
		}
		double r_u[C], r_v[C];
#pragma unroll
		for (int j = 0; j < C; ++j) {
			const bool we = j == C - 1 && x + C == nx, ww = j == 0 && x == 0;
			const double Au = stencil5<double, true>(off, dg, U[j], j < C - 1 ? U[j + 1] : uE, j ? U[j - 1] : uW,
			                                         Up[j], Um[j], we, ww, wn, ws);
			const double Av = stencil5<double, true>(off, dg, W[j], j < C - 1 ? W[j + 1] : wE, j ? W[j - 1] : wW,
			                                         Wp[j], Wm[j], we, ww, wn, ws);
			r_u[j] = rx[j] - Au;
			r_v[j] = ry[j] - Av;
		}
#pragma unroll
		for (int j = 0; j < C; j += 2) {
			*(double2*) (a.r[0] + i + j) = make_double2(r_u[j], r_u[j + 1]);
			*(double2*) (a.r[1] + i + j) = make_double2(r_v[j], r_v[j + 1]);
		}
#pragma unroll
		for (int j = 0; j < C; ++j) {
			acc_u[0] += r_u[j] * r_u[j];
			acc_u[1] += rx[j] * rx[j];
			acc_v[0] += r_v[j] * r_v[j];
			acc_v[1] += ry[j] * ry[j];
		}
	}
	reduce_partial<2>(acc_u, a.pu[0]);
	reduce_partial<2>(acc_v, a.pu[1]);
}

// cells per thread of the right-hand-side launch: pairs (2048^2, both
// fields: 71.8 us; runs of 4: 82.8, of 8: 108.6, one cell per thread: 80.3;
// 512- and 1024-thread blocks of pairs: 73.2 and 77.6)
constexpr int ADV_RUN = 2;

// The CG forms' start for the velocity solves (no reference counterpart: the
// reference starts from the current velocity).  A = I - sL has its spectrum
// in [1, lmax], lmax = 1 + 8|s|n^2 (the periodic 5-point Laplacian's
// eigenvalues lie in [-8n^2, 0]).  The start is a fixed polynomial in A
// applied to b, evaluated in six applications of A:
//   two Neumann steps from x = b: r1 = b - A b, x2 = b + r1,
//   r2 = r1 - A r1 = (I - A)^2 b;
//   four Chebyshev steps on [1, lmax] from (x2, r2) (Saad, Iterative Methods
//   for Sparse Linear Systems, 2nd ed., Algorithm 12.1): d0 = r2 / theta,
//   x += d_k, r -= A d_k, d_{k+1} = ca_k d_k + cb_k r
// so that r0 = (I - A)^2 C4(A) b with |C4| <= 1 / T4((lmax + 1) / (lmax - 1))
// on the spectrum.  The Neumann factors remove the smooth part of the step's
// right-hand side (eigenvalues near 1), the Chebyshev factor the rough part
// the spread forces bring in.  At 2048^2 (8|s|n^2 = 0.34) the start's residual
// is about 1e-11 |b|, below the solves' tolerance: the CG takes no iteration
// (2-3 after the three-term Neumann start x0 = 3b - 3Ab + A^2 b, 4-5 from the
// current velocity).  Used only while 8|s|n^2 < 0.5 (ns_step).
// b is stored by the right-hand-side launch (which also sums b.b); one pass
// forms x0 and r from it and sums r.r.  Six 5-point passes would read about
// 200 B per cell and field; this kernel reads b and writes x0 and r
// (20-24 B).
// A block holds a tile of 128 columns x NW * RPW rows in registers (96 rows:
// 12 waves of 8): lane l of wave w holds columns 2l, 2l + 1 of rows
// w * RPW .. w * RPW + RPW - 1.  The k-th application's result is exact on
// the tile shrunk by k cells on every side, so the tile writes its
// 116 x (NW * RPW - 12) inner cells; tiles overlap by 12 cells in each
// direction.  East and west neighbours come from the adjacent lane (DPP),
// north and south from the strip itself or, for a wave's first and last
// rows, from the adjacent wave's boundary rows exchanged through LDS.  Every
// cell takes the same operations (A v = dg v + off ((E + W) + (N + S)), fused
// multiply-adds, no wrap-dependent order), so a cell's values are the same
// bits in every tile that forms them.  164 VGPRs: three waves per SIMD, one
// block per CU.  (Measured at 2048^2, both fields: 8-wave blocks 56.8 us,
// 6-wave 67.0 us, 12-wave 54.0 us; with the reference's CSR-order stencil
// and its wrap flags instead of the fused form, 75.9 us.)
// Tile columns start at 116 tx - 6 (even: two-column loads stay aligned):
// tile 0 writes columns 0 .. 115, the last tile stops at n - 1.
constexpr int ST_HALO = 6;               // applications of A per start
constexpr int ST_CHEB = 4;               // of which Chebyshev steps
constexpr int ST_OX = 128 - 2 * ST_HALO; // columns written per tile
#ifndef AFM_ST_NW
#define AFM_ST_NW 12
#endif
#ifndef AFM_ST_RPW
#define AFM_ST_RPW 8
#endif
constexpr int ST_NW = AFM_ST_NW, ST_RPW = AFM_ST_RPW;   // waves per block, rows per wave

template<int NW, int RPW>
struct StartTile {
	static constexpr int LY = NW * RPW;            // rows held
	static constexpr int OY = LY - 2 * ST_HALO;    // rows written
};

// The Chebyshev steps' coefficients for the spectrum [1, lmax] (host side,
// start_poly): theta = (lmax + 1) / 2, delta = (lmax - 1) / 2,
// rho_0 = delta / theta, rho_{k+1} = delta / (2 theta - delta rho_k),
// ca_k = rho_{k+1} rho_k, cb_k = 2 / (2 theta - delta rho_k) (the algorithm's
// 2 rho_{k+1} / delta, finite at delta = 0).
struct StartPoly {
	double inv_theta;
	double ca[ST_CHEB - 1], cb[ST_CHEB - 1];
};

// xb[buf][wave + 1][first / last row][lane]: slots 0 and NW + 1 are never
// written (the rows they stand for lie outside the tile's exact cells)
template<int NW>
using StartXch = double2[2][NW + 2][2][64];

template<int NW, int RPW>
__device__ inline void st_publish(StartXch<NW>& xb, int buf, const double (&v)[RPW][2], int w, int lane)
{
	xb[buf][w + 1][0][lane] = make_double2(v[0][0], v[0][1]);
	xb[buf][w + 1][1][lane] = make_double2(v[RPW - 1][0], v[RPW - 1][1]);
}

// The sum of v's four neighbours at (row i, column o) of the lane's strip:
// E / W from the pair or the adjacent lane, S / N from the strip or the
// adjacent waves' rows up / dn.  (E + W) + (N + S) in every cell: the start
// needs no parity with the reference's CSR rows (the CG's stop test decides
// what it returns), only the same operations wherever tiles overlap.
template<int RPW>
__device__ inline double st_nsum(const double (&v)[RPW][2], int i, int o, double2 up, double2 dn)
{
	double E, W;
	if (o == 0) { W = lane_shift<DPP_WAVE_SHR1>(v[i][1]); E = v[i][1]; }
	else { W = v[i][0]; E = lane_shift<DPP_WAVE_SHL1>(v[i][0]); }
	const double S = i > 0 ? v[i - 1][o] : (o ? up.y : up.x);
	const double N = i < RPW - 1 ? v[i + 1][o] : (o ? dn.y : dn.x);
	return (E + W) + (N + S);
}

// y - A v with A v = dg v + off s, s the neighbour sum of v
__device__ inline double st_sub_apply(double y, double v, double s, double off, double dg)
{
	return __builtin_fma(-off, s, __builtin_fma(-dg, v, y));
}

// Marks v as formed here (an empty asm statement that takes and returns it
// in registers): without it the compiler sinks the x += d additions to the
// stores and holds every step's d in registers until then
__device__ inline void st_pin(double& v) { __asm__ volatile("" : "+v"(v)); }

// Publishes v's first and last rows and waits for the block (one LDS round;
// the two buffers alternate, so a wave that runs ahead writes the other one)
template<int NW, int RPW>
__device__ inline void st_post(StartXch<NW>& xb, int buf, const double (&v)[RPW][2], int w, int lane)
{
	st_publish<NW, RPW>(xb, buf, v, w, lane);
	__syncthreads();
}

// One tile: columns c0 .. c0 + 127 and rows r0 .. r0 + LY - 1 (unwrapped;
// EDGE when some of them lie outside 0 .. n - 1, so that their loads and
// stores wrap).  Adds the written cells' r^2 to acc.
template<typename TR, int NW, int RPW, bool EDGE>
__device__ inline void st_tile(StartXch<NW>& xb, int& buf, const double* __restrict__ b, double* __restrict__ x,
                               TR* __restrict__ r, int n, int c0, int r0, double off, double dg, const StartPoly& P,
                               double& acc)
{
	constexpr int LY = StartTile<NW, RPW>::LY;
	// the wave index (and with it the rows) is the same in every lane: held
	// in scalar registers
	const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), msk = n - 1;
	const int cu = c0 + 2 * lane;                     // the pair's first column, unwrapped (even)
	const int gc = EDGE ? (cu & msk) : cu;
	int gy[RPW];
#pragma unroll
	for (int i = 0; i < RPW; ++i) gy[i] = EDGE ? ((r0 + RPW * w + i) & msk) : r0 + RPW * w + i;
	double xv[RPW][2], rv[RPW][2], dv[RPW][2];
	double2 up, dn;
#pragma unroll
	for (int i = 0; i < RPW; ++i) {
		const double2 t = *(const double2*) (b + (size_t) n * gy[i] + gc);
		xv[i][0] = t.x; xv[i][1] = t.y;
	}
	// r1 = b - A b (x = b)
	st_post<NW, RPW>(xb, buf, xv, w, lane);
	up = xb[buf][w][1][lane]; dn = xb[buf][w + 2][0][lane];
	buf ^= 1;
#pragma unroll
	for (int i = 0; i < RPW; ++i)
#pragma unroll
		for (int o = 0; o < 2; ++o)
			rv[i][o] = st_sub_apply(xv[i][o], xv[i][o], st_nsum<RPW>(xv, i, o, up, dn), off, dg);
	// x2 = b + r1, r2 = r1 - A r1
	st_post<NW, RPW>(xb, buf, rv, w, lane);
	up = xb[buf][w][1][lane]; dn = xb[buf][w + 2][0][lane];
	buf ^= 1;
#pragma unroll
	for (int i = 0; i < RPW; ++i)
#pragma unroll
		for (int o = 0; o < 2; ++o)
			dv[i][o] = st_sub_apply(rv[i][o], rv[i][o], st_nsum<RPW>(rv, i, o, up, dn), off, dg);
#pragma unroll
	for (int i = 0; i < RPW; ++i)
#pragma unroll
		for (int o = 0; o < 2; ++o) {
			xv[i][o] += rv[i][o];
			st_pin(xv[i][o]);
			rv[i][o] = dv[i][o];
		}
	// the Chebyshev steps: d_k, then x += d_k, r -= A d_k.  The strip's first
	// and last rows of d_k are formed before the LDS round, the others while
	// its reads are in flight.
#pragma unroll
	for (int k = 0; k < ST_CHEB; ++k) {
		const double ca = k ? P.ca[k - 1] : 0.0, cb = k ? P.cb[k - 1] : P.inv_theta;
#pragma unroll
		for (int i = 0; i < RPW; i += RPW - 1)
#pragma unroll
			for (int o = 0; o < 2; ++o) dv[i][o] = k ? __builtin_fma(ca, dv[i][o], cb * rv[i][o]) : cb * rv[i][o];
		st_post<NW, RPW>(xb, buf, dv, w, lane);
		up = xb[buf][w][1][lane]; dn = xb[buf][w + 2][0][lane];
		buf ^= 1;
#pragma unroll
		for (int i = 1; i < RPW - 1; ++i)
#pragma unroll
			for (int o = 0; o < 2; ++o) dv[i][o] = k ? __builtin_fma(ca, dv[i][o], cb * rv[i][o]) : cb * rv[i][o];
#pragma unroll
		for (int i = 0; i < RPW; ++i)
#pragma unroll
			for (int o = 0; o < 2; ++o) {
				rv[i][o] = st_sub_apply(rv[i][o], dv[i][o], st_nsum<RPW>(dv, i, o, up, dn), off, dg);
				xv[i][o] += dv[i][o];
				st_pin(xv[i][o]);
			}
	}
	// the written cells stored
	bool colok[2];
#pragma unroll
	for (int o = 0; o < 2; ++o) {
		const int lc = 2 * lane + o;
		colok[o] = lc >= ST_HALO && lc < 128 - ST_HALO && (!EDGE || cu + o <= n - 1);
	}
#pragma unroll
	for (int i = 0; i < RPW; ++i) {
		const int ly = RPW * w + i;
		if (!(ly >= ST_HALO && ly < LY - ST_HALO && (!EDGE || r0 + ly < n))) continue;
		const size_t row = (size_t) n * gy[i];
		if (colok[0] && colok[1]) {
			*(double2*) (x + row + gc) = make_double2(xv[i][0], xv[i][1]);
			if (sizeof(TR) == 4) *(float2*) (r + row + gc) = make_float2((float) rv[i][0], (float) rv[i][1]);
			else *(double2*) (r + row + gc) = make_double2(rv[i][0], rv[i][1]);
		} else {
#pragma unroll
			for (int o = 0; o < 2; ++o)
				if (colok[o]) {
					x[row + gc + o] = xv[i][o];
					r[row + gc + o] = (TR) rv[i][o];
				}
		}
#pragma unroll
		for (int o = 0; o < 2; ++o)
			if (colok[o]) acc += rv[i][o] * rv[i][o];
	}
}

// blockIdx.y = field; blocks walk tiles t = blockIdx.x, + gridDim.x, ...
// (gridDim.x at most the right-hand-side launch's block count: the r.r
// partials of blocks beyond it are the zeros that launch wrote)
template<typename TR, int NW, int RPW>
__global__ __launch_bounds__(64 * NW) void k_cg_start_fused(CGArgs a, const double* __restrict__ b0,
                                                            const double* __restrict__ b1, int n, double off,
                                                            double dg, StartPoly P, int tiles_x, int ntiles)
{
	using ST = StartTile<NW, RPW>;
	__shared__ StartXch<NW> xb;
	const int f = blockIdx.y;
	if (!a.x[f]) return;
	const double* __restrict__ b = f ? b1 : b0;
	double* __restrict__ x = a.x[f];
	TR* __restrict__ r = (TR*) a.r[f];
	int buf = 0;
	double acc[1] = {0.0};
	for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
		const int tx = t % tiles_x, ty = t / tiles_x;
		const int c0 = ST_OX * tx - ST_HALO, r0 = ST::OY * ty - ST_HALO;
		const bool edge = c0 < 0 || c0 + 127 > n - 1 || r0 < 0 || r0 + ST::LY - 1 > n - 1;
		if (edge) st_tile<TR, NW, RPW, true>(xb, buf, b, x, r, n, c0, r0, off, dg, P, acc[0]);
		else st_tile<TR, NW, RPW, false>(xb, buf, b, x, r, n, c0, r0, off, dg, P, acc[0]);
	}
	reduce_partial<1>(acc, a.pu[f]);
}

static inline unsigned rblocks(size_t N)
{
	size_t b = (N + 255) / 256;
	return (unsigned) (b < (size_t) MAX_RED_BLOCKS ? b : (size_t) MAX_RED_BLOCKS);
}

int launch_sum(Ctx* c, const double* x, size_t N, double* out_host)
{
	SolveCtl h{};
	AFM_HIP(hipMemcpyAsync(&c->ctl[3], &h, sizeof(SolveCtl), hipMemcpyHostToDevice, c->stream));
	hipLaunchKernelGGL(k_sum, dim3(rblocks(N)), dim3(256), 0, c->stream, x, N, c->partials);
	hipLaunchKernelGGL(k_sum_fin, dim3(1), dim3(256), 0, c->stream, rblocks(N), &c->ctl[3], c->partials);
	AFM_LAUNCH_CHECK();
	AFM_HIP(hipMemcpyAsync(&c->hctl[3], &c->ctl[3], sizeof(SolveCtl), hipMemcpyDeviceToHost, c->stream));
	AFM_HIP(hipStreamSynchronize(c->stream));
	*out_host = c->hctl[3].sum2;
	return 0;
}

// -------------------------------------------------------------- host ops --
static inline dim3 pgrid(size_t N) { return dim3((unsigned) ((N + 255) / 256)); }

int op_div(Ctx* c, const double* u, const double* v, double* d)
{
	hipLaunchKernelGGL(k_div, pgrid(c->N), dim3(256), 0, c->stream, u, v, d, (int) c->nx, (int) c->ny, c->n);
	AFM_LAUNCH_CHECK();
	return 0;
}

int op_grad_sub(Ctx* c, const double* p, double* u, double* v)
{
	hipLaunchKernelGGL(k_grad_sub, pgrid(c->N), dim3(256), 0, c->stream, p, u, v, (int) c->nx, (int) c->ny, c->n,
	                   (double*) nullptr, (double*) nullptr, 0);
	AFM_LAUNCH_CHECK();
	return 0;
}

int op_advect(Ctx* c, const double* u, const double* v, double* ua, double* va)
{
	hipLaunchKernelGGL(k_advect, pgrid(c->N), dim3(256), 0, c->stream, u, v, ua, va, (int) c->nx, (int) c->ny, c->n);
	AFM_LAUNCH_CHECK();
	return 0;
}

static void helm_coef(const Ctx* c, double scale, double& off, double& dg)
{
	off = scale * c->n * c->n;          // ns/helmholtz.h:68
	dg = 1. - 4. * off;                 // I + s L5 on the periodic plane
}

int op_helm_apply(Ctx* c, double scale, const double* x, double* y)
{
	double off, dg;
	helm_coef(c, scale, off, dg);
	hipLaunchKernelGGL(k_helm_apply, pgrid(c->N), dim3(256), 0, c->stream, x, y, (int) c->nx, (int) c->ny, off, dg);
	AFM_LAUNCH_CHECK();
	return 0;
}

// CG on up to two fields at once.  Scratch: r, z, p0, p1 per field.
struct CGWork {
	double *r[2] = {nullptr, nullptr}, *p[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
};

static int dalloc(std::vector<void*>& own, void** p, size_t bytes)
{
	AFM_HIP(hipMalloc(p, bytes));
	own.push_back(*p);
	return 0;
}

static int cg_alloc(Ctx* c, CGWork& w, std::vector<void*>& own)
{
	for (int f = 0; f < 2; ++f) {
		AFM_TRY(dalloc(own, (void**) &w.r[f], c->N * 8));
		AFM_TRY(dalloc(own, (void**) &w.p[f][0], c->N * 8));
		AFM_TRY(dalloc(own, (void**) &w.p[f][1], c->N * 8));
	}
	return 0;
}

// Solves (I + scale L) x_f = b_f for the non-null fields; iters[f] receives
// the iteration counts.
// The right-hand side of a velocity step, formed inside the CG start instead
// of being stored (ns_step).
struct AdvectInit {
	const double *fx, *fy;
	double *uo, *vo;
	double dt, rho, hoff, hdg;
	int be;
	int neumann;      // the CG forms' polynomial start (k_cg_start_fused)
};

// The start's Chebyshev coefficients (StartPoly) for the spectrum of
// A = I + (off / n^2) L, [1, 1 - 8 off] (off < 0; [1 - 8 off, 1] for off > 0)
static StartPoly start_poly(double off)
{
	StartPoly P{};
	const double theta = 1.0 - 4.0 * off, delta = 4.0 * std::fabs(off);
	P.inv_theta = 1.0 / theta;
	double rho = delta / theta;
	for (int k = 0; k < ST_CHEB - 1; ++k) {
		const double den = 2.0 * theta - delta * rho, rn = delta / den;
		P.ca[k] = rn * rho;
		P.cb[k] = 2.0 / den;
		rho = rn;
	}
	return P;
}

// spec (optional): enqueues work that needs the converged solution, right
// after the first read-back of the control blocks is enqueued; the host then
// waits for that read-back only, so the work runs while the host decides.
// *spec_ok tells whether the solve had converged at that first test (the
// work is then valid; otherwise the CG went on and the work must be redone).
int op_cg2(Ctx* c, CGWork& w, double scale, double* const x[2], const double* const b[2], double tol, int tol_mode,
           int max_iters, int iters[2], double res[2], const AdvectInit* adv = nullptr, bool p32 = false,
           int (*spec)(void*) = nullptr, void* spec_arg = nullptr, bool* spec_ok = nullptr)
{
	if (spec_ok) *spec_ok = false;
	bool first_read = true;
	double off, dg;
	helm_coef(c, scale, off, dg);
	CGArgs a{};
	for (int f = 0; f < 2; ++f) {
		a.x[f] = x[f];
		a.b[f] = b[f];
		a.r[f] = w.r[f];
		a.p[f][0] = w.p[f][0];
		a.p[f][1] = w.p[f][1];
		a.ctl[f] = &c->ctl[1 + f];
		a.pu[f] = c->partials + (size_t) f * 4 * MAX_RED_BLOCKS;
		a.pp[f] = a.pu[f] + 2 * MAX_RED_BLOCKS;
	}
	a.tol_mode = tol_mode;
	const unsigned rb = rblocks(c->N);
	// runs of ADV_RUN cells along x for the right-hand side (row lengths a
	// multiple of the run: every power of two from ADV_RUN)
	const bool pairs = c->nx % ADV_RUN == 0;
	// FP32 r only where the polynomial start writes it (k_cg_start_fused)
	const bool r32 = p32 && adv && adv->neumann;
	if (adv && adv->neumann) {
		// b goes to the first p buffers (the first p update reads r and
		// writes p[.][1], the second writes p[.][0])
		if (pairs)
			hipLaunchKernelGGL((k_advect_cg_init_runs<true, ADV_RUN>), dim3(rb), dim3(ADV_BS), 0, c->stream, a, adv->fx, adv->fy,
			                   adv->uo, adv->vo, (int) c->nx, (int) c->ny, c->n, adv->dt, adv->rho, adv->be,
			                   adv->hoff, adv->hdg, off, dg, tol, max_iters, w.p[0][0], w.p[1][0]);
		else
			hipLaunchKernelGGL(k_advect_cg_init<true>, dim3(rb), dim3(256), 0, c->stream, a, adv->fx, adv->fy, adv->uo,
			                   adv->vo, (int) c->nx, (int) c->ny, c->n, adv->dt, adv->rho, adv->be, adv->hoff,
			                   adv->hdg, off, dg, tol, max_iters, w.p[0][0], w.p[1][0]);
		using ST = StartTile<ST_NW, ST_RPW>;
		const int n = (int) c->nx;
		const int tiles_x = (n + ST_OX - 1) / ST_OX, ntiles = tiles_x * ((n + ST::OY - 1) / ST::OY);
		const unsigned sb = (unsigned) ntiles < rb ? (unsigned) ntiles : rb;
		const StartPoly P = start_poly(off);
		if (r32)
			hipLaunchKernelGGL((k_cg_start_fused<float, ST_NW, ST_RPW>), dim3(sb, 2), dim3(64 * ST_NW), 0, c->stream,
			                   a, w.p[0][0], w.p[1][0], n, off, dg, P, tiles_x, ntiles);
		else
			hipLaunchKernelGGL((k_cg_start_fused<double, ST_NW, ST_RPW>), dim3(sb, 2), dim3(64 * ST_NW), 0, c->stream,
			                   a, w.p[0][0], w.p[1][0], n, off, dg, P, tiles_x, ntiles);
	} else if (adv)
		if (pairs)
			hipLaunchKernelGGL((k_advect_cg_init_runs<false, ADV_RUN>), dim3(rb), dim3(ADV_BS), 0, c->stream, a, adv->fx, adv->fy,
			                   adv->uo, adv->vo, (int) c->nx, (int) c->ny, c->n, adv->dt, adv->rho, adv->be,
			                   adv->hoff, adv->hdg, off, dg, tol, max_iters, nullptr, nullptr);
		else
			hipLaunchKernelGGL(k_advect_cg_init<false>, dim3(rb), dim3(256), 0, c->stream, a, adv->fx, adv->fy, adv->uo,
			                   adv->vo, (int) c->nx, (int) c->ny, c->n, adv->dt, adv->rho, adv->be, adv->hoff,
			                   adv->hdg, off, dg, tol, max_iters, nullptr, nullptr);
	else
		hipLaunchKernelGGL(k_cg_init, dim3(rb, 2), dim3(256), 0, c->stream, a, (int) c->nx, (int) c->ny, off, dg, tol,
		                   max_iters);
	hipLaunchKernelGGL(k_cg_init_fin, dim3(2), dim3(256), 0, c->stream, a, rb);
	AFM_LAUNCH_CHECK();
	int enq = 0;
	// as many iterations as the previous solve took (none when its start met
	// the tolerance), 4 for a context's first solve
	int batch = c->last_cg_iters >= 0 ? c->last_cg_iters : 4;
	for (;;) {
		if (batch > max_iters - enq) batch = max_iters - enq;
		// two launches per iteration: each folds the finaliser of the one
		// before it; the batch's last update is finalised on its own
		for (int s = 0; s < batch; ++s) {
			const int pin = enq & 1;
			if (r32) {
				hipLaunchKernelGGL((k_cg_p<float, float>), dim3(rb, 2), dim3(256), 0, c->stream, a, (int) c->nx,
				                   (int) c->ny, off, dg, pin, enq, s > 0 ? 1 : 0, rb);
				hipLaunchKernelGGL((k_cg_update<float, float>), dim3(rb, 2), dim3(256), 0, c->stream, a, (int) c->nx,
				                   (int) c->ny, off, dg, pin ^ 1, enq, rb);
			} else if (p32) {
				hipLaunchKernelGGL((k_cg_p<float, double>), dim3(rb, 2), dim3(256), 0, c->stream, a, (int) c->nx,
				                   (int) c->ny, off, dg, pin, enq, s > 0 ? 1 : 0, rb);
				hipLaunchKernelGGL((k_cg_update<float, double>), dim3(rb, 2), dim3(256), 0, c->stream, a, (int) c->nx,
				                   (int) c->ny, off, dg, pin ^ 1, enq, rb);
			} else {
				hipLaunchKernelGGL((k_cg_p<double, double>), dim3(rb, 2), dim3(256), 0, c->stream, a, (int) c->nx,
				                   (int) c->ny, off, dg, pin, enq, s > 0 ? 1 : 0, rb);
				hipLaunchKernelGGL((k_cg_update<double, double>), dim3(rb, 2), dim3(256), 0, c->stream, a, (int) c->nx,
				                   (int) c->ny, off, dg, pin ^ 1, enq, rb);
			}
			AFM_LAUNCH_CHECK();
			++enq;
		}
		if (batch > 0) {
			hipLaunchKernelGGL(k_cg_update_fin, dim3(2), dim3(256), 0, c->stream, a, enq, rb);
			AFM_LAUNCH_CHECK();
		}
		AFM_HIP(hipMemcpyAsync(&c->hctl[1], &c->ctl[1], 2 * sizeof(SolveCtl), hipMemcpyDeviceToHost, c->stream));
		bool all;
		if (spec && first_read) {
			AFM_HIP(hipEventRecord(c->cg_read_ev, c->stream));
			AFM_TRY(spec(spec_arg));
			AFM_HIP(hipEventSynchronize(c->cg_read_ev));
			all = c->hctl[1].done && c->hctl[2].done;
			if (spec_ok) *spec_ok = all;
		} else {
			AFM_HIP(hipStreamSynchronize(c->stream));
			all = c->hctl[1].done && c->hctl[2].done;
		}
		first_read = false;
		if (all || enq >= max_iters) break;
		batch = 2;
	}
	int mx = 0;
	for (int f = 0; f < 2; ++f) {
		if (iters) iters[f] = x[f] ? c->hctl[1 + f].count : 0;
		if (res) res[f] = x[f] ? c->hctl[1 + f].res : 0.0;
		if (x[f] && c->hctl[1 + f].count > mx) mx = c->hctl[1 + f].count;
	}
	c->last_cg_iters = mx;
	// the stop flag also fires at the iteration limit: convergence is the test
	for (int f = 0; f < 2; ++f)
		if (x[f] && !(c->hctl[1 + f].res <= c->hctl[1 + f].tol)) {
			set_error("CG did not converge within max_iters");
			return -4;
		}
	return 0;
}

// ------------------------------------------------------------ NS solver ---
struct NS {
	Ctx* c = nullptr;
	double mu, rho, dt, tol;
	int tol_mode, precision;
	double *u, *v;                         // caller-owned, aliased (ns/solver.h:54-56)
	double *uo = nullptr, *vo = nullptr, *phi = nullptr, *fz = nullptr;
	// The CG forms (precision 1, 3) start a step's projection from the
	// potentials of the previous steps of the same kind (0: backward-Euler
	// half step, 1: Crank-Nicolson full step): pot[kind] is the next start,
	// the linear extrapolation 2 phi_n - phi_{n-1} of the last two potentials
	// of the kind (phi_n alone after the first), pot[kind + 2] the last
	// potential phi_n.  The potential of every projection is kept where get_p
	// reads it (phi_last).
	double* pot[4] = {nullptr, nullptr, nullptr, nullptr};
	bool pot_ok[4] = {false, false, false, false};
	const double* phi_last = nullptr;
	CGWork cg;
	bool half = false;
	int last_cycles = 0, last_cg[2] = {0, 0};
	double last_rms = 0.0, last_thr = 0.0;   // the last projection's final RMS residual and its bound
	int max_cycles = 200, max_cg = 10000;
	std::vector<void*> own;
	~NS() { for (void* p : own) (void) hipFree(p); }
};

// kind: the step kind whose start potential the solve uses and replaces
// (0, 1), or -1 for a projection outside the time steps
// The projection's divergence and its sum for the mean in one pass (even row
// lengths; returns 1 then), else the divergence alone (0)
static int ns_div(NS* s, const double* u, const double* v)
{
	Ctx* c = s->c;
	if (c->nx % 2 == 0) {
		hipLaunchKernelGGL(k_div_sum, dim3(rblocks(c->N)), dim3(256), 0, c->stream, u, v, s->phi, (int) c->nx,
		                   (int) c->ny, c->n, &c->ctl[0], c->partials, s->tol_mode == 1 ? 0.0 : s->tol, s->max_cycles);
		AFM_LAUNCH_CHECK();
		return 1;
	}
	AFM_TRY(op_div(c, u, v, s->phi));
	return 0;
}

// div_ready: ns_div(s, u, v) has been enqueued since u, v last changed
// (ns_step's velocity solve, which enqueues it while the host waits)
int ns_project(NS* s, double* u, double* v, int kind, bool div_ready)
{
	Ctx* c = s->c;
	bool sums = false;
	if (div_ready) sums = c->nx % 2 == 0;
	else {
		const int r = ns_div(s, u, v);
		if (r < 0) return r;
		sums = r == 1;
	}
	int cyc = 0;
	const double* sol = nullptr;
	const bool warm = kind >= 0 && s->pot[kind];
	const double* x0 = (warm && s->pot_ok[kind]) ? s->pot[kind] : nullptr;
	AFM_TRY(mg_solve(c, s->phi, s->tol, s->tol_mode, 2, 2, s->precision, s->max_cycles, &cyc, nullptr, &sol, x0,
	                 sums));
	s->last_cycles = cyc;
	s->last_rms = c->hctl[0].res;
	s->last_thr = c->hctl[0].tol;
	// the gradient is taken where the solve left the potential, and the
	// potential is stored (to phi, or as the kind's last potential, with the
	// kind's next start extrapolated from it and the one before) in the same
	// pass
	double* keep = warm ? s->pot[kind + 2] : s->phi;
	hipLaunchKernelGGL(k_grad_sub, pgrid(c->N), dim3(256), 0, c->stream, sol, u, v, (int) c->nx, (int) c->ny, c->n,
	                   keep, warm ? s->pot[kind] : nullptr, warm && s->pot_ok[kind + 2] ? 1 : 0);
	AFM_LAUNCH_CHECK();
	if (warm) s->pot_ok[kind] = s->pot_ok[kind + 2] = true;
	s->phi_last = keep;
	return 0;
}

int ns_create(Ctx* c, double mu, double rho, double dt, double tol, int tol_mode, int precision, double* u0,
              double* v0, NS** out)
{
	NS* s = new NS();
	s->c = c; s->mu = mu; s->rho = rho; s->dt = dt; s->tol = tol; s->tol_mode = tol_mode; s->precision = precision;
	s->u = u0; s->v = v0;
	int rc = 0;
	std::vector<void*>& o = s->own;
	if ((rc = dalloc(o, (void**) &s->uo, c->N * 8)) || (rc = dalloc(o, (void**) &s->vo, c->N * 8)) ||
	    (rc = dalloc(o, (void**) &s->phi, c->N * 8)) || (rc = cg_alloc(c, s->cg, o))) {
		delete s;
		return rc;
	}
	if (precision == 1 || precision == 3)
		for (int k = 0; k < 4; ++k)
			if ((rc = dalloc(o, (void**) &s->pot[k], c->N * 8))) { delete s; return rc; }
	s->phi_last = s->phi;
	AFM_HIP(hipMemsetAsync(s->uo, 0, c->N * 8, c->stream));
	AFM_HIP(hipMemsetAsync(s->vo, 0, c->N * 8, c->stream));
	rc = ns_project(s, s->u, s->v, -1, false);     // ns/solver.h:71
	if (rc) { delete s; return rc; }
	*out = s;
	return 0;
}

// ns/solver.h:90-140
int ns_step(NS* s, const double* fx, const double* fy)
{
	Ctx* c = s->c;
	const double sm = s->mu * s->dt / (2 * s->rho);      // phhm scale
	const double ss = -s->mu * s->dt / (2 * s->rho);     // phhs scale
	double hoff, hdg;
	helm_coef(c, sm, hoff, hdg);
	// advection and the step's right-hand side are formed inside the CG start
	// The polynomial start leaves r0 = (I - A)^2 C4(A) b; A = I - ss L has
	// its spectrum in [1, 1 + 8|ss|n^2], where the Neumann factor (I - A)^2
	// only damps while 8|ss|n^2 is well below 1.  Beyond that the CG forms
	// start from the current velocity, as the FP64 form always does.
	const bool neumann = (s->precision == 1 || s->precision == 3) && 8.0 * std::fabs(ss) * c->n * c->n < 0.5;
	const AdvectInit adv{fx, fy, s->uo, s->vo, s->dt, s->rho, hoff, hdg, s->half ? 0 : 1, neumann ? 1 : 0};
	double* xs[2] = {s->u, s->v};
	const double* bs[2] = {nullptr, nullptr};
	// the mixed form stores the CG's search directions in FP32; the
	// projection's divergence is enqueued behind the CG's first read-back
	// (valid when the solve had then converged: the start met the
	// tolerance, or the predicted iterations sufficed)
	auto spec = [](void* p) -> int {
		NS* ns = (NS*) p;
		const int r = ns_div(ns, ns->u, ns->v);
		return r < 0 ? r : 0;
	};
	bool div_ready = false;
	AFM_TRY(op_cg2(c, s->cg, ss, xs, bs, s->tol, s->tol_mode, s->max_cg, s->last_cg, nullptr, &adv, s->precision == 1,
	               spec, s, &div_ready));
	AFM_TRY(ns_project(s, s->u, s->v, s->half ? 1 : 0, div_ready));
	s->half = !s->half;
	return 0;
}

// ns/solver.h:86: p = phhs * phidt
int ns_get_p(NS* s, double* p)
{
	return op_helm_apply(s->c, -s->mu * s->dt / (2 * s->rho), s->phi_last, p);
}

// ------------------------------------------------------------ IB driver ---
// One dt (the reference has no driver; its intended coupling is the comment
// at test/ib_ns_test.cc:80-90):
//   F = -k (X - R);  f = S(X) F;              BE half step -> u^{n+1/2}
//   X_h = X + (dt/2) J(X) u^{n+1/2}
//   F = -k (X_h - R);  f = S(X_h) F;          CN full step -> u^{n+1}
//   X = X + dt (0.5 (J(X_h) u^n + J(X_h) u^{n+1}))
struct IB {
	NS* ns = nullptr;
	long m = 0;
	double *X = nullptr, *R = nullptr, *k = nullptr, *ds = nullptr, *Xh = nullptr, *fx = nullptr, *fy = nullptr;
	std::vector<void*> own;
	~IB() { for (void* p : own) (void) hipFree(p); }
};

int ib_create(NS* ns, long m, const double* X, const double* R, const double* k, const double* ds, IB** out)
{
	Ctx* c = ns->c;
	IB* b = new IB();
	b->ns = ns;
	b->m = m;
	size_t mb = (size_t) (m > 0 ? m : 1) * 8;
	int rc = 0;
	std::vector<void*>& o = b->own;
	if ((rc = dalloc(o, (void**) &b->X, 2 * mb)) || (rc = dalloc(o, (void**) &b->R, 2 * mb)) ||
	    (rc = dalloc(o, (void**) &b->k, mb)) || (rc = dalloc(o, (void**) &b->ds, mb)) ||
	    (rc = dalloc(o, (void**) &b->Xh, 2 * mb)) || (rc = dalloc(o, (void**) &b->fx, c->N * 8)) ||
	    (rc = dalloc(o, (void**) &b->fy, c->N * 8)) || (rc = ensure_markers(c, m))) {
		delete b;
		return rc;
	}
	if (m > 0) {
		AFM_HIP(hipMemcpyAsync(b->X, X, 2 * m * 8, hipMemcpyDeviceToDevice, c->stream));
		AFM_HIP(hipMemcpyAsync(b->R, R, 2 * m * 8, hipMemcpyDeviceToDevice, c->stream));
		AFM_HIP(hipMemcpyAsync(b->k, k, m * 8, hipMemcpyDeviceToDevice, c->stream));
		AFM_HIP(hipMemcpyAsync(b->ds, ds, m * 8, hipMemcpyDeviceToDevice, c->stream));
	}
	*out = b;
	return 0;
}

int ib_step(IB* b)
{
	NS* s = b->ns;
	Ctx* c = s->c;
	const double dt = s->dt;
	if (s->half) { set_error("the NS solver must be at a full step"); return -1; }
	AFM_TRY(launch_spread(c, 3, nullptr, nullptr, b->ds, b->X, b->R, b->k, b->m, b->fx, b->fy));
	AFM_TRY(ns_step(s, b->fx, b->fy));
	AFM_TRY(launch_ib_move_half(c, s->u, s->v, b->X, b->m, 0.5 * dt, b->Xh));
	AFM_TRY(launch_spread(c, 3, nullptr, nullptr, b->ds, b->Xh, b->R, b->k, b->m, b->fx, b->fy));
	AFM_TRY(ns_step(s, b->fx, b->fy));
	AFM_TRY(launch_ib_move_full(c, s->uo, s->vo, s->u, s->v, b->Xh, b->m, dt, b->X));
	return 0;
}

// ------------------------------------------------------------- adjoint --
// Reverse-mode gradient of J (kinetic energy + centroid displacement, see
// afm_adj.hip) through `steps` IB steps, with checkpoints of (u, v, X) every
// `every` steps and recomputation of each segment's tape in the reverse sweep.
struct StepTape { double *u = nullptr, *v = nullptr, *uh = nullptr, *vh = nullptr, *X = nullptr, *Xh = nullptr; };

static int d2d(Ctx* c, double* dst, const double* src, size_t count)
{
	AFM_HIP(hipMemcpyAsync(dst, src, count * 8, hipMemcpyDeviceToDevice, c->stream));
	return 0;
}

static int ib_step_taped(IB* b, StepTape* tp)
{
	NS* s = b->ns;
	Ctx* c = s->c;
	const double dt = s->dt;
	if (tp) {
		AFM_TRY(d2d(c, tp->u, s->u, c->N)); AFM_TRY(d2d(c, tp->v, s->v, c->N)); AFM_TRY(d2d(c, tp->X, b->X, 2 * b->m));
	}
	AFM_TRY(launch_spread(c, 3, nullptr, nullptr, b->ds, b->X, b->R, b->k, b->m, b->fx, b->fy));
	AFM_TRY(ns_step(s, b->fx, b->fy));
	if (tp) { AFM_TRY(d2d(c, tp->uh, s->u, c->N)); AFM_TRY(d2d(c, tp->vh, s->v, c->N)); }
	AFM_TRY(launch_ib_move_half(c, s->u, s->v, b->X, b->m, 0.5 * dt, b->Xh));
	if (tp) AFM_TRY(d2d(c, tp->Xh, b->Xh, 2 * b->m));
	AFM_TRY(launch_spread(c, 3, nullptr, nullptr, b->ds, b->Xh, b->R, b->k, b->m, b->fx, b->fy));
	AFM_TRY(ns_step(s, b->fx, b->fy));
	AFM_TRY(launch_ib_move_full(c, s->uo, s->vo, s->u, s->v, b->Xh, b->m, dt, b->X));
	return 0;
}

int ib_grad(IB* b, int steps, int every, long msize, double* J_out, double* gk, double* gR)
{
	NS* s = b->ns;
	Ctx* c = s->c;
	const size_t N = c->N;
	const long m = b->m;
	if (msize <= 0 || m % msize != 0) { set_error("msize must divide the marker count"); return -1; }
	if (steps < 0 || every < 1) { set_error("steps >= 0 and checkpoint_every >= 1 required"); return -1; }
	if (s->half) { set_error("the NS solver must be at a full step"); return -1; }
	const long nm = m / msize;
	const bool want_grad = gk || gR;
	std::vector<void*> own;
	bool oom = false;
	auto A = [&](size_t count) -> double* {
		void* p = nullptr;
		if (hipMalloc(&p, (count ? count : 1) * 8) != hipSuccess) { oom = true; return nullptr; }
		own.push_back(p);
		return (double*) p;
	};
	int rc = 0;
	auto fail = [&](int r) { (void) hipStreamSynchronize(c->stream); for (void* p : own) (void) hipFree(p); return r; };
	const int nseg = steps > 0 ? (steps + every - 1) / every : 0;
	double* c0 = A(2 * nm); double* c1 = A(2 * nm); double* part = A(nm > 0 ? nm : 1);
	double* Xb = A(2 * m);
	double* ub = A(N); double* vb = A(N);
	if (!c0 || !c1 || !part || !Xb || !ub || !vb) { set_error("out of device memory (adjoint)"); return fail(-3); }
	if ((rc = adj_centroids(c, b->X, nm, msize, nullptr, c0, nullptr, nullptr))) return fail(rc);
	// a checkpoint: u, v, X and, for the CG forms, the projections' start
	// and last potentials (so that a segment's recomputation repeats the
	// forward run bit for bit)
	const bool warm = s->pot[0] != nullptr;
	const int K = warm ? 7 : 3;
	std::vector<double*> ck(want_grad ? K * nseg : 0);
	std::vector<char> ck_ok(want_grad ? 4 * nseg : 0);
	for (size_t i = 0; i < ck.size(); ++i) {
		ck[i] = A(i % K == 2 ? 2 * m : N);
		if (!ck[i]) { set_error("out of device memory (checkpoints)"); return fail(-3); }
	}
	auto save_ck = [&](int sg) -> int {
		AFM_TRY(d2d(c, ck[K * sg], s->u, N));
		AFM_TRY(d2d(c, ck[K * sg + 1], s->v, N));
		AFM_TRY(d2d(c, ck[K * sg + 2], b->X, 2 * m));
		for (int k = 0; warm && k < 4; ++k) {
			AFM_TRY(d2d(c, ck[K * sg + 3 + k], s->pot[k], N));
			ck_ok[4 * sg + k] = s->pot_ok[k];
		}
		return 0;
	};
	auto load_ck = [&](int sg) -> int {
		AFM_TRY(d2d(c, s->u, ck[K * sg], N));
		AFM_TRY(d2d(c, s->v, ck[K * sg + 1], N));
		AFM_TRY(d2d(c, b->X, ck[K * sg + 2], 2 * m));
		for (int k = 0; warm && k < 4; ++k) {
			AFM_TRY(d2d(c, s->pot[k], ck[K * sg + 3 + k], N));
			s->pot_ok[k] = ck_ok[4 * sg + k] != 0;
		}
		return 0;
	};
	for (int st = 0; st < steps; ++st) {
		if (want_grad && st % every == 0 && (rc = save_ck(st / every))) return fail(rc);
		if ((rc = ib_step_taped(b, nullptr))) return fail(rc);
	}
	// objective and its adjoint at the final time
	double ke = 0.0;
	if ((rc = adj_kinetic(c, s->u, s->v, ub, vb, &ke))) return fail(rc);
	if ((rc = adj_centroids(c, b->X, nm, msize, c0, c1, Xb, part))) return fail(rc);
	std::vector<double> hp(nm > 0 ? nm : 1);
	if (nm > 0 && hipMemcpy(hp.data(), part, nm * 8, hipMemcpyDeviceToHost) != hipSuccess) { set_error("copy"); return fail(-3); }
	double J = ke;
	for (long i = 0; i < nm; ++i) J += hp[i];
	if (J_out) *J_out = J;
	if (!want_grad || steps == 0) {
		if (gk) AFM_HIP(hipMemsetAsync(gk, 0, m * 8, c->stream));
		if (gR) AFM_HIP(hipMemsetAsync(gR, 0, 2 * m * 8, c->stream));
		return fail(0);
	}
	// final state, kept to restore after the sweep
	double* fu = A(N); double* fv = A(N); double* fX = A(2 * m);
	double* fs[4] = {nullptr, nullptr, nullptr, nullptr};
	bool fs_ok[4];
	for (int k = 0; k < 4; ++k) {
		if (warm) fs[k] = A(N);
		fs_ok[k] = s->pot_ok[k];
	}
	std::vector<StepTape> tape(every);
	for (auto& t : tape) { t.u = A(N); t.v = A(N); t.uh = A(N); t.vh = A(N); t.X = A(2 * m); t.Xh = A(2 * m); }
	double* unb = A(N); double* vnb = A(N); double* uhb = A(N); double* vhb = A(N);
	double* us = A(N); double* vs = A(N); double* rbu = A(N); double* rbv = A(N);
	double* t1 = A(N); double* t2 = A(N);
	double* Xnb = A(2 * m); double* Xhb = A(2 * m);
	double* Ux = A(m); double* Uy = A(m); double* Jx = A(m); double* Jy = A(m);
	double* cx = A(m); double* cy = A(m); double* volv = A(m);
	double* kb = gk ? gk : A(m); double* Rb = gR ? gR : A(2 * m);
	if (!fu || !fv || !fX || !unb || !vnb || !uhb || !vhb || !us || !vs || !rbu || !rbv || !t1 || !t2 || !Xnb ||
	    !Xhb || !Ux || !Uy || !Jx || !Jy || !cx || !cy || !volv || !kb || !Rb ||
	    (warm && (!fs[0] || !fs[1] || !fs[2] || !fs[3])) || oom) {
		set_error("out of device memory (adjoint tape)");
		return fail(-3);
	}
	// the reverse sweep's projections overwrite phi, which get_p reads in the
	// FP64 and FP32-cycle forms (the CG forms read the full step's start)
	double* fphi = A(N);
	if (!fphi) { set_error("out of device memory (adjoint tape)"); return fail(-3); }
	if ((rc = d2d(c, fu, s->u, N)) || (rc = d2d(c, fv, s->v, N)) || (rc = d2d(c, fX, b->X, 2 * m)) ||
	    (rc = d2d(c, fphi, s->phi, N)))
		return fail(rc);
	for (int k = 0; warm && k < 4; ++k)
		if ((rc = d2d(c, fs[k], s->pot[k], N))) return fail(rc);
	const double vol = 1. / (c->n * c->n);
	{
		std::vector<double> hv(m, vol);
		AFM_HIP(hipMemcpy(volv, hv.data(), m * 8, hipMemcpyHostToDevice));
	}
	AFM_HIP(hipMemsetAsync(kb, 0, m * 8, c->stream));
	AFM_HIP(hipMemsetAsync(Rb, 0, 2 * m * 8, c->stream));
	const double dt = s->dt, rho = s->rho, half = 0.5 * dt;
	const double sm = s->mu * s->dt / (2 * s->rho), ss = -s->mu * s->dt / (2 * s->rho);
	auto solve_H = [&](double* xu, double* xv, const double* bu, const double* bv) -> int {
		AFM_TRY(d2d(c, xu, bu, N));
		AFM_TRY(d2d(c, xv, bv, N));
		double* xs[2] = {xu, xv};
		const double* bs[2] = {bu, bv};
		return op_cg2(c, s->cg, ss, xs, bs, s->tol, s->tol_mode, s->max_cg, nullptr, nullptr);
	};
	for (int sg = nseg - 1; sg >= 0; --sg) {
		const int len = std::min(every, steps - sg * every);
		if ((rc = load_ck(sg))) return fail(rc);
		s->half = false;
		for (int j = 0; j < len; ++j)
			if ((rc = ib_step_taped(b, &tape[j]))) return fail(rc);
		for (int j = len - 1; j >= 0; --j) {
			const StepTape& tp = tape[j];
			const double* unx = j == len - 1 ? (sg == nseg - 1 ? fu : ck[K * (sg + 1)]) : tape[j + 1].u;
			const double* vnx = j == len - 1 ? (sg == nseg - 1 ? fv : ck[K * (sg + 1) + 1]) : tape[j + 1].v;
			// step 7: X' = X + dt 0.5 (J(Xh) u^n + J(Xh) u^{n+1})
			AFM_TRY(d2d(c, Xnb, Xb, 2 * m));
			AFM_TRY(adj_split_scale(c, Xb, 0.5 * dt, m, Ux, Uy));
			AFM_HIP(hipMemsetAsync(unb, 0, N * 8, c->stream));
			AFM_HIP(hipMemsetAsync(vnb, 0, N * 8, c->stream));
			AFM_TRY(launch_spread(c, 3, Ux, Uy, volv, tp.Xh, nullptr, nullptr, m, ub, vb, 1));
			AFM_TRY(launch_spread(c, 3, Ux, Uy, volv, tp.Xh, nullptr, nullptr, m, unb, vnb, 1));
			AFM_HIP(hipMemsetAsync(Xhb, 0, 2 * m * 8, c->stream));
			AFM_TRY(adj_interp_dX(c, 0, tp.u, tp.Xh, Ux, m, Xhb));
			AFM_TRY(adj_interp_dX(c, 0, unx, tp.Xh, Ux, m, Xhb));
			AFM_TRY(adj_interp_dX(c, 1, tp.v, tp.Xh, Uy, m, Xhb));
			AFM_TRY(adj_interp_dX(c, 1, vnx, tp.Xh, Uy, m, Xhb));
			// step 6: Crank-Nicolson full step
			AFM_TRY(d2d(c, us, ub, N)); AFM_TRY(d2d(c, vs, vb, N));
			AFM_TRY(ns_project(s, us, vs, -1, false));
			AFM_TRY(solve_H(rbu, rbv, us, vs));
			AFM_TRY(op_helm_apply(c, sm, rbu, t1)); AFM_TRY(adj_axpy(c, unb, t1, 1.0, N));
			AFM_TRY(op_helm_apply(c, sm, rbv, t2)); AFM_TRY(adj_axpy(c, vnb, t2, 1.0, N));
			AFM_TRY(adj_scale_copy(c, us, rbu, dt * (1 / rho), N));     // f1 adjoint
			AFM_TRY(adj_scale_copy(c, vs, rbv, dt * (1 / rho), N));
			AFM_TRY(adj_scale_copy(c, t1, rbu, -dt, N));                // advection adjoint
			AFM_TRY(adj_scale_copy(c, t2, rbv, -dt, N));
			AFM_HIP(hipMemsetAsync(uhb, 0, N * 8, c->stream));
			AFM_HIP(hipMemsetAsync(vhb, 0, N * 8, c->stream));
			AFM_TRY(adj_advect_T(c, t1, t2, tp.uh, tp.vh, uhb, vhb));
			// step 5: f1 = S(Xh) F1, F1 = -k (Xh - R)
			AFM_TRY(launch_interp(c, 0, us, tp.Xh, m, Jx));
			AFM_TRY(launch_interp(c, 1, vs, tp.Xh, m, Jy));
			AFM_TRY(adj_spring(c, tp.Xh, b->R, b->k, b->ds, Jx, Jy, m, kb, Rb, Xhb, cx, cy));
			AFM_TRY(adj_interp_dX(c, 0, us, tp.Xh, cx, m, Xhb));
			AFM_TRY(adj_interp_dX(c, 1, vs, tp.Xh, cy, m, Xhb));
			// step 4: Xh = X + half J(X) u_h
			AFM_TRY(adj_axpy(c, Xnb, Xhb, 1.0, 2 * m));
			AFM_TRY(adj_split_scale(c, Xhb, half, m, Ux, Uy));
			AFM_TRY(launch_spread(c, 3, Ux, Uy, volv, tp.X, nullptr, nullptr, m, uhb, vhb, 1));
			AFM_TRY(adj_interp_dX(c, 0, tp.uh, tp.X, Ux, m, Xnb));
			AFM_TRY(adj_interp_dX(c, 1, tp.vh, tp.X, Uy, m, Xnb));
			// step 3: backward-Euler half step
			AFM_TRY(ns_project(s, uhb, vhb, -1, false));
			AFM_TRY(solve_H(rbu, rbv, uhb, vhb));
			AFM_TRY(adj_axpy(c, unb, rbu, 1.0, N));
			AFM_TRY(adj_axpy(c, vnb, rbv, 1.0, N));
			AFM_TRY(adj_scale_copy(c, us, rbu, (dt / 2) * (1 / rho), N));
			AFM_TRY(adj_scale_copy(c, vs, rbv, (dt / 2) * (1 / rho), N));
			AFM_TRY(adj_scale_copy(c, t1, rbu, -(dt / 2), N));
			AFM_TRY(adj_scale_copy(c, t2, rbv, -(dt / 2), N));
			AFM_TRY(adj_advect_T(c, t1, t2, tp.u, tp.v, unb, vnb));
			// steps 1-2: f0 = S(X) F0, F0 = -k (X - R)
			AFM_TRY(launch_interp(c, 0, us, tp.X, m, Jx));
			AFM_TRY(launch_interp(c, 1, vs, tp.X, m, Jy));
			AFM_TRY(adj_spring(c, tp.X, b->R, b->k, b->ds, Jx, Jy, m, kb, Rb, Xnb, cx, cy));
			AFM_TRY(adj_interp_dX(c, 0, us, tp.X, cx, m, Xnb));
			AFM_TRY(adj_interp_dX(c, 1, vs, tp.X, cy, m, Xnb));
			std::swap(ub, unb); std::swap(vb, vnb); std::swap(Xb, Xnb);
		}
	}
	// leave the solver at the final state, as a forward run would
	if ((rc = d2d(c, s->u, fu, N)) || (rc = d2d(c, s->v, fv, N)) || (rc = d2d(c, b->X, fX, 2 * m)) ||
	    (rc = d2d(c, s->phi, fphi, N)))
		return fail(rc);
	for (int k = 0; warm && k < 4; ++k) {
		if ((rc = d2d(c, s->pot[k], fs[k], N))) return fail(rc);
		s->pot_ok[k] = fs_ok[k];
	}
	s->phi_last = warm ? s->pot[3] : s->phi;     // the last forward projection: a full step's
	s->half = false;
	return fail(0);
}

int ib_markers(IB* b, double* X_out)
{
	Ctx* c = b->ns->c;
	if (b->m > 0) AFM_HIP(hipMemcpyAsync(X_out, b->X, 2 * b->m * 8, hipMemcpyDeviceToDevice, c->stream));
	return 0;
}

Ctx* ib_ctx(IB* b) { return b->ns->c; }
void ib_free(IB* b) { (void) hipStreamSynchronize(b->ns->c->stream); delete b; }
Ctx* ns_ctx(NS* s) { return s->c; }
void ns_free(NS* s) { (void) hipStreamSynchronize(s->c->stream); delete s; }

// the CG forms' potentials of a step kind (see NS): kind 0, 1 the next
// start, 2, 3 the last potential; copy out, or in
int ns_get_start(NS* s, int kind, double* phi, int* valid)
{
	if (kind < 0 || kind > 3) { set_error("kind must be 0 / 1 (next start of the half / full step) or 2 / 3 (their last potential)"); return -1; }
	*valid = s->pot[kind] && s->pot_ok[kind];
	if (*valid) AFM_HIP(hipMemcpyAsync(phi, s->pot[kind], s->c->N * 8, hipMemcpyDeviceToDevice, s->c->stream));
	return 0;
}

int ns_set_start(NS* s, int kind, const double* phi)
{
	if (kind < 0 || kind > 3) { set_error("kind must be 0 / 1 (next start of the half / full step) or 2 / 3 (their last potential)"); return -1; }
	if (!s->pot[kind]) {
		if (!phi) return 0;
		set_error("only the CG forms (precision 1, 3) start from a potential");
		return -1;
	}
	s->pot_ok[kind] = phi != nullptr;
	// a start set alone has no last potential to extrapolate from
	if (kind < 2) s->pot_ok[kind + 2] = false;
	if (phi) AFM_HIP(hipMemcpyAsync(s->pot[kind], phi, s->c->N * 8, hipMemcpyDeviceToDevice, s->c->stream));
	return 0;
}

int ns_stats(NS* s, int* cyc, int* cgu, int* cgv, int* half)
{
	if (cyc) *cyc = s->last_cycles;
	if (cgu) *cgu = s->last_cg[0];
	if (cgv) *cgv = s->last_cg[1];
	if (half) *half = s->half ? 1 : 0;
	return 0;
}

int ns_last_residual(NS* s, double* rms, double* thr)
{
	*rms = s->last_rms;
	*thr = s->last_thr;
	return 0;
}

int ns_old(NS* s, const double** uo, const double** vo)
{
	*uo = s->uo;
	*vo = s->vo;
	return 0;
}

// ------------------------------------------ allocation-free Helmholtz CG --
// A variant of op_cg_single for callers that cannot allocate device memory.
// The residual and the two search directions of its one field live in this
// file-scope array, of which every GPU holds its own copy until the process
// ends.  Host code never passes `scratch` itself (its name is a host-side
// stand-in, not a device address): cg_scratch_setup looks up the device
// address once per device, with that device current, and the solve uses
// that.  Every vector in it is written before it is read (r and the first p
// by the CG start, the second p by the first iteration), so nothing needs
// clearing between calls; calls that overlap on one GPU race on it.
constexpr size_t CG_SCRATCH_CELLS = (size_t) 2048 * 2048;   // 3 vectors: 96 MiB
constexpr int CG_SCRATCH_DEVICES = 64;
static __device__ __align__(256) double scratch[3 * CG_SCRATCH_CELLS];
static double* scratch_dev[CG_SCRATCH_DEVICES];
static std::mutex scratch_mu;

// once per device (from the context's creation), with `device` current
int cg_scratch_setup(int device)
{
	std::lock_guard<std::mutex> lock(scratch_mu);
	if (device < 0 || device >= CG_SCRATCH_DEVICES) { set_error("device index beyond the scratch table"); return -1; }
	if (scratch_dev[device]) return 0;
	void* p = nullptr;
	const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
	if (e != hipSuccess || !p) {
		set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
		return -3;
	}
	scratch_dev[device] = static_cast<double*>(p);
	return 0;
}

int op_cg_single_noalloc(Ctx* c, double scale, const double* b, double* x, double tol, int tol_mode, int max_iters,
                         int* iters, double* final_norm)
{
	if (c->N > CG_SCRATCH_CELLS) {
		set_error("the allocation-free CG takes grids of at most 2048 x 2048 cells");
		return -1;
	}
	double* s = nullptr;
	{
		std::lock_guard<std::mutex> lock(scratch_mu);
		s = (c->device >= 0 && c->device < CG_SCRATCH_DEVICES) ? scratch_dev[c->device] : nullptr;
	}
	if (!s) { set_error("the allocation-free CG's buffer is not set up on this device"); return -3; }
	CGWork w;     // the second field stays null: the CG kernels skip a field whose x is null
	w.r[0] = s;
	w.p[0][0] = s + c->N;
	w.p[0][1] = s + 2 * c->N;
	double* xs[2] = {x, nullptr};
	const double* bs[2] = {b, nullptr};
	int it[2];
	double res[2];
	const int rc = op_cg2(c, w, scale, xs, bs, tol, tol_mode, max_iters, it, res);
	if (rc == 0 || rc == -4) {
		if (iters) *iters = it[0];
		if (final_norm) *final_norm = res[0];
	}
	return rc;
}

// Single-field CG with scratch owned for the duration of the call.
int op_cg_single(Ctx* c, double scale, const double* b, double* x, double tol, int tol_mode, int max_iters,
                 int* iters, double* final_norm)
{
	std::vector<void*> own;
	CGWork w;
	int rc = cg_alloc(c, w, own);
	if (rc == 0) {
		double* xs[2] = {x, nullptr};
		const double* bs[2] = {b, nullptr};
		int it[2];
		double res[2];
		rc = op_cg2(c, w, scale, xs, bs, tol, tol_mode, max_iters, it, res);
		if (iters) *iters = it[0];
		if (final_norm) *final_norm = res[0];
	}
	// the buffers are freed only once the solve's kernels are done with them
	const hipError_t se = hipStreamSynchronize(c->stream);
	for (void* p : own) (void) hipFree(p);
	if (rc == 0 && se != hipSuccess) {
		set_error(std::string("hipStreamSynchronize: ") + hipGetErrorString(se));
		return -3;
	}
	return rc;
}

}  // namespace afm
