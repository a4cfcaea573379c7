// afm_adj.hip -- discrete adjoint of the IB time step on gfx950, with
// checkpoint-and-recompute (no reference counterpart).
//
// Objective at the final time:  J = 1/2 sum (u^2 + v^2) h^2
//                                  + sum_m |c_m(T) - c_m(0)|^2,
// c_m the centroid of membrane m (consecutive blocks of msize markers).
// Gradients with respect to the per-marker stiffness k and rest positions R.
//
// Adjoint building blocks (each the transpose of a forward kernel):
//   interpolation^T (field)  = spreading of the marker adjoint with ds = h^2
//                              (k_spread_tile, accumulate mode);
//   spreading^T (force)      = interpolation of the field adjoint, scaled ds/h^2;
//   d/dX of both             = k_interp_dX, a gather with delta' in place of
//                              delta along one axis;
//   advection^T              = k_advect_T (gather form, no atomics);
//   Helmholtz and projection = the same CG and multigrid solves (both
//                              operators are symmetric on the periodic grid).
// Solves are differentiated as exact linear solves.
#include "afm_device.h"

namespace afm {

// d delta/dr on the window (|n r| <= 2): -(0.25 n)(pi n/2) sin(theta)
__device__ inline double ddelta(double r, unsigned n)
{
	double nr = n * r;
	if (fabs(nr) >= 3.0) return 0.0;
	double a = delta_arg(r, n);
	return -(0.25 * n) * (3.141592653589793 * n / 2) * sin(a);
}

__global__ void k_axpy(double* __restrict__ y, const double* __restrict__ x, double a, size_t N)
{
	size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i < N) y[i] += a * x[i];
}

__global__ void k_scale_copy(double* __restrict__ y, const double* __restrict__ x, double a, size_t N)
{
	size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i < N) y[i] = a * x[i];
}

// (ub, vb) += A'(u, v)^T (ab, bb): gather form of the transpose of adv_at.
__global__ __launch_bounds__(256) void k_advect_T(const double* __restrict__ ab, const double* __restrict__ bb,
                                                  const double* __restrict__ u, const double* __restrict__ v,
                                                  double* __restrict__ ub, double* __restrict__ vb, int nx, int ny,
                                                  unsigned n)
{
	const size_t N = (size_t) nx * ny;
	size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i >= N) return;
	int x, z;
	cell_xy(i, nx, x, z);
	auto id = [&](int a, int b) -> size_t {
		a = (a + nx) % nx; b = (b + ny) % ny;
		return (size_t) a + (size_t) nx * b;
	};
	const double nn = (double) n;
	// adjoints of the intermediate products at cell (a, b)
	auto uub = [&](int a, int b) { return nn * (ab[id(a - 1, b)] - ab[id(a, b)]); };
	auto wwb = [&](int a, int b) { return nn * (bb[id(a, b - 1)] - bb[id(a, b)]); };
	auto uwb = [&](int a, int b) { return nn * (ab[id(a, b - 1)] - ab[id(a, b)]) + nn * (bb[id(a - 1, b)] - bb[id(a, b)]); };
	auto quu = [&](int a, int b) { return 0.5 * (u[id(a, b)] + u[id(a - 1, b)]); };
	auto qww = [&](int a, int b) { return 0.5 * (v[id(a, b)] + v[id(a, b - 1)]); };
	auto quw = [&](int a, int b) { return 0.5 * (u[id(a, b - 1)] + u[id(a, b)]); };
	auto ruw = [&](int a, int b) { return 0.5 * (v[id(a - 1, b)] + v[id(a, b)]); };
	// u[i] enters uu(x, z), uu(x+1, z), uw(x, z), uw(x, z+1)
	double gu = quu(x, z) * uub(x, z) + quu(x + 1, z) * uub(x + 1, z)
	          + 0.5 * ruw(x, z) * uwb(x, z) + 0.5 * ruw(x, z + 1) * uwb(x, z + 1);
	// v[i] enters ww(x, z), ww(x, z+1), uw(x, z), uw(x+1, z)
	double gv = qww(x, z) * wwb(x, z) + qww(x, z + 1) * wwb(x, z + 1)
	          + 0.5 * quw(x, z) * uwb(x, z) + 0.5 * quw(x + 1, z) * uwb(x + 1, z);
	ub[i] += gu;
	vb[i] += gv;
}

// Xb[k] += coef_k * (d/dX_k, d/dY_k) (J f)_k, J the interpolation of comp.
__global__ void k_interp_dX(int comp, const double* __restrict__ f, const double* __restrict__ X,
                            const double* __restrict__ coef, long m, unsigned n, unsigned W, unsigned D,
                            unsigned nx, unsigned ny, double* __restrict__ Xb)
{
	long k = blockIdx.x * (long) blockDim.x + threadIdx.x;
	if (k >= m) return;
	const double cx = X[2 * k], cy = X[2 * k + 1];
	int ix0, iy0;
	double dx[4], dy[4];
	ib_window(comp, cx, cy, n, W, D, nx, ny, ix0, iy0, dx, dy);
	// recompute the window abscissae for the derivative factors
	double q0 = cx, q1 = cy;
	if (comp == 0) q1 -= 0.5 / n; else q0 -= 0.5 / n;
	q0 = fmod(fmod(q0, (double) W) + W, (double) W);
	q1 = fmod(fmod(q1, (double) D) + D, (double) D);
	double fx = floor(q0 * n), fy = floor(q1 * n);
	if (!(fx >= 0.0 && fx <= (double) nx)) fx = 0.0;
	if (!(fy >= 0.0 && fy <= (double) ny)) fy = 0.0;
	double ddx[4], ddy[4];
	for (int d = 0; d < 4; ++d) {
		double x = (fx + d - 1.0) / n, y = (fy + d - 1.0) / n;
		if (comp == 0) y += 0.5 / n; else x += 0.5 / n;
		ddx[d] = ddelta(x - cx, n);
		ddy[d] = ddelta(y - cy, n);
	}
	const double vol = 1. / (n * n);
	double gx = 0.0, gy = 0.0;
	for (int b = 0; b < 4; ++b) {
		int iy = iy0 + b; if (iy >= (int) ny) iy -= ny;
		for (int a = 0; a < 4; ++a) {
			int ix = ix0 + a; if (ix >= (int) nx) ix -= nx;
			double fv = vol * f[ix + (size_t) nx * iy];
			gx += -ddx[a] * dy[b] * fv;
			gy += -dx[a] * ddy[b] * fv;
		}
	}
	Xb[2 * k] += coef[k] * gx;
	Xb[2 * k + 1] += coef[k] * gy;
}

// From X'b: Ub_c = scale * Xb[., c]  (the marker adjoint of a position update)
__global__ void k_split_scale(const double* __restrict__ Xb, double scale, long m, double* __restrict__ Ux,
                              double* __restrict__ Uy)
{
	long k = blockIdx.x * (long) blockDim.x + threadIdx.x;
	if (k >= m) return;
	Ux[k] = scale * Xb[2 * k];
	Uy[k] = scale * Xb[2 * k + 1];
}

// Spring and spreading adjoint at positions P (X or X_h):
//   Fb = (ds/h^2) J(P) fb;  kb += -(P - R).Fb;  Rb += k Fb;  Pb += -k Fb;
//   cx, cy = ds F / h^2 (coefficients of the position derivative of S).
__global__ void k_spring_adj(const double* __restrict__ P, const double* __restrict__ R, const double* __restrict__ kk,
                             const double* __restrict__ ds, const double* __restrict__ Jx, const double* __restrict__ Jy,
                             double vol, long m, double* __restrict__ kb, double* __restrict__ Rb,
                             double* __restrict__ Pb, double* __restrict__ cx, double* __restrict__ cy)
{
	long i = blockIdx.x * (long) blockDim.x + threadIdx.x;
	if (i >= m) return;
	const double Fbx = ds[i] / vol * Jx[i], Fby = ds[i] / vol * Jy[i];
	const double ex = P[2 * i] - R[2 * i], ey = P[2 * i + 1] - R[2 * i + 1];
	kb[i] += -ex * Fbx - ey * Fby;
	Rb[2 * i] += kk[i] * Fbx;
	Rb[2 * i + 1] += kk[i] * Fby;
	Pb[2 * i] += -kk[i] * Fbx;
	Pb[2 * i + 1] += -kk[i] * Fby;
	cx[i] = ds[i] * (-kk[i] * ex) / vol;
	cy[i] = ds[i] * (-kk[i] * ey) / vol;
}

// Centroid per membrane (one block each, fixed-order sum); with c0 given,
// also the objective's marker adjoint 2 (c - c0) / msize and the partial
// objective sum |c - c0|^2 per membrane.
__global__ __launch_bounds__(256) void k_centroids(const double* __restrict__ X, long msize, const double* __restrict__ c0,
                                                   double* __restrict__ c, double* __restrict__ Xb,
                                                   double* __restrict__ part)
{
	const long mm = blockIdx.x;
	double v[2] = {0.0, 0.0};
	for (long j = threadIdx.x; j < msize; j += 256) {
		v[0] += X[2 * (mm * msize + j)];
		v[1] += X[2 * (mm * msize + j) + 1];
	}
	block_sum<2>(v);
	__shared__ double cc[2];
	if (threadIdx.x == 0) {
		cc[0] = v[0] / msize;
		cc[1] = v[1] / msize;
		c[2 * mm] = cc[0];
		c[2 * mm + 1] = cc[1];
		if (c0) {
			double ex = cc[0] - c0[2 * mm], ey = cc[1] - c0[2 * mm + 1];
			part[mm] = ex * ex + ey * ey;
		}
	}
	__syncthreads();
	if (c0 && Xb)
		for (long j = threadIdx.x; j < msize; j += 256) {
			Xb[2 * (mm * msize + j)] = 2 * (cc[0] - c0[2 * mm]) / msize;
			Xb[2 * (mm * msize + j) + 1] = 2 * (cc[1] - c0[2 * mm + 1]) / msize;
		}
}

// sum (u^2 + v^2) h^2 / 2 -> ctl->sum2 ; ub = u h^2, vb = v h^2
__global__ __launch_bounds__(256) void k_kinetic(const double* __restrict__ u, const double* __restrict__ v, size_t N,
                                                 double h2, double* __restrict__ ub, double* __restrict__ vb,
                                                 SolveCtl* ctl, double* partials)
{
	double acc[1] = {0.0};
	for (size_t i = blockIdx.x * 256 + threadIdx.x; i < N; i += (size_t) gridDim.x * 256) {
		acc[0] += 0.5 * (u[i] * u[i] + v[i] * v[i]) * h2;
		if (ub) { ub[i] = u[i] * h2; vb[i] = v[i] * h2; }
	}
	reduce_partial<1>(acc, partials);
}

__global__ __launch_bounds__(256) void k_kinetic_fin(unsigned nblk, SolveCtl* ctl, const double* partials)
{
	double tot[1];
	red_total<1>(partials, nblk, tot);
	if (threadIdx.x == 0) ctl->sum2 = tot[0];
}

// ------------------------------------------------------------- launchers --
static inline dim3 g1(size_t N) { return dim3((unsigned) ((N + 255) / 256)); }

int adj_axpy(Ctx* c, double* y, const double* x, double a, size_t N)
{
	hipLaunchKernelGGL(k_axpy, g1(N), dim3(256), 0, c->stream, y, x, a, N);
	AFM_LAUNCH_CHECK();
	return 0;
}

int adj_scale_copy(Ctx* c, double* y, const double* x, double a, size_t N)
{
	hipLaunchKernelGGL(k_scale_copy, g1(N), dim3(256), 0, c->stream, y, x, a, N);
	AFM_LAUNCH_CHECK();
	return 0;
}

int adj_advect_T(Ctx* c, const double* ab, const double* bb, const double* u, const double* v, double* ub, double* vb)
{
	hipLaunchKernelGGL(k_advect_T, g1(c->N), dim3(256), 0, c->stream, ab, bb, u, v, ub, vb, (int) c->nx, (int) c->ny, c->n);
	AFM_LAUNCH_CHECK();
	return 0;
}

int adj_interp_dX(Ctx* c, int comp, const double* f, const double* X, const double* coef, long m, double* Xb)
{
	if (m <= 0) return 0;
	hipLaunchKernelGGL(k_interp_dX, g1(m), dim3(256), 0, c->stream, comp, f, X, coef, m, c->n, c->W, c->D, c->nx,
	                   c->ny, Xb);
	AFM_LAUNCH_CHECK();
	return 0;
}

int adj_split_scale(Ctx* c, const double* Xb, double scale, long m, double* Ux, double* Uy)
{
	if (m <= 0) return 0;
	hipLaunchKernelGGL(k_split_scale, g1(m), dim3(256), 0, c->stream, Xb, scale, m, Ux, Uy);
	AFM_LAUNCH_CHECK();
	return 0;
}

int adj_spring(Ctx* c, const double* P, const double* R, const double* k, const double* ds, const double* Jx,
               const double* Jy, long m, double* kb, double* Rb, double* Pb, double* cx, double* cy)
{
	if (m <= 0) return 0;
	const double vol = 1. / (c->n * c->n);
	hipLaunchKernelGGL(k_spring_adj, g1(m), dim3(256), 0, c->stream, P, R, k, ds, Jx, Jy, vol, m, kb, Rb, Pb, cx, cy);
	AFM_LAUNCH_CHECK();
	return 0;
}

int adj_centroids(Ctx* c, const double* X, long nm, long msize, const double* c0, double* cout, double* Xb,
                  double* part)
{
	if (nm <= 0) return 0;
	hipLaunchKernelGGL(k_centroids, dim3((unsigned) nm), dim3(256), 0, c->stream, X, msize, c0, cout, Xb, part);
	AFM_LAUNCH_CHECK();
	return 0;
}

int adj_kinetic(Ctx* c, const double* u, const double* v, double* ub, double* vb, double* out_host)
{
	SolveCtl h{};
	AFM_HIP(hipMemcpyAsync(&c->ctl[3], &h, sizeof(SolveCtl), hipMemcpyHostToDevice, c->stream));
	size_t b = (c->N + 255) / 256;
	unsigned nb = (unsigned) (b < (size_t) MAX_RED_BLOCKS ? b : (size_t) MAX_RED_BLOCKS);
	const double h2 = 1.0 / ((double) c->n * c->n);
	hipLaunchKernelGGL(k_kinetic, dim3(nb), dim3(256), 0, c->stream, u, v, c->N, h2, ub, vb, &c->ctl[3], c->partials);
	hipLaunchKernelGGL(k_kinetic_fin, dim3(1), dim3(256), 0, c->stream, nb, &c->ctl[3], c->partials);
	AFM_LAUNCH_CHECK();
	AFM_HIP(hipMemcpyAsync(&c->hctl[3], &c->ctl[3], sizeof(SolveCtl), hipMemcpyDeviceToHost, c->stream));
	AFM_HIP(hipStreamSynchronize(c->stream));
	*out_host = c->hctl[3].sum2;
	return 0;
}

}  // namespace afm
