// afm_ns3.hip -- the reference's 3-D wall-bounded Navier-Stokes step:
// ns::solver<geometry<L,1,1,1>> (ns/solver.h:37-236) on n x n x n cells,
// periodic in x and z, walls in y, with its projection (ns/projection.h) and
// Helmholtz operators (ns/helmholtz.h):
//   u, w   on the n^3 x- and z-faces, Helmholtz helmholtz_periodic (kind 1:
//          second-order wall closure in the first and last rows);
//   v      on the n*(n-1)*n interior y-faces, helmholtz_dirichlet (kind 3);
//   walls  n*n vectors ut, ub, wt, wb (index x + n*z), updated in place by
//          projection::boundary as the reference updates the caller's.
// One thread per row, global memory: this is the parity path for the
// reference's own geometry, not the 2-D benchmark path.  Every row uses the
// reference's arithmetic in its CSR column order; the CG reductions are
// fixed-order trees (the reference sums sequentially), so solves agree to
// rounding and the operators exactly.
#include "afm_device.h"

namespace afm {

// One row of ns::helmholtz_periodic (kind 1) or ns::helmholtz_dirichlet
// (kind 3, ny = the interior y-face rows) in ascending column order.
struct HRow {
	size_t col[7];
	double a[7];
	int cnt;
};

__device__ inline void helm3_row(unsigned x, unsigned y, unsigned z, unsigned nx, unsigned ny, unsigned nz, unsigned n,
                                 int kind, double scale, HRow& r)
{
	const size_t i = x + (size_t) nx * (y + (size_t) ny * z);
	const double off = scale * n * n;
	double diagonal = kind == 3 ? 1. - scale * 6. * (double) (n * n) : 1. - 6. * off;
	double left = off, right = off, top = off, bottom = off, front = off, back = off;
	if (nx == 1) { diagonal += left + right; left = right = 0; }
	else if (nx == 2) { if (x == 0) { right += left; left = 0; } else { left += right; right = 0; } }
	if (kind == 1) {
		if (ny == 1) { diagonal -= 3 * (top + bottom); top = bottom = 0; }
		else if (y == 0) { diagonal -= 2 * bottom; top += off / 3; bottom = 0; }
		else if (y == ny - 1) { diagonal -= 2 * top; bottom += off / 3; top = 0; }
	} else {
		if (y == 0) bottom = 0;
		if (y == ny - 1) top = 0;
	}
	if (nz == 1) { diagonal += back + front; back = front = 0; }
	else if (nz == 2) { if (z == 0) { front += back; back = 0; } else { back += front; front = 0; } }
	r.cnt = 0;
	auto put = [&](double v, size_t c) { r.a[r.cnt] = v; r.col[r.cnt] = c; r.cnt++; };
	if (front != 0 && z == nz - 1) put(front, x + (size_t) nx * y);
	if (back != 0 && z > 0) put(back, i - (size_t) nx * ny);
	if (bottom != 0 && y > 0) put(bottom, i - nx);
	if (right != 0 && x == nx - 1) put(right, (size_t) nx * (y + (size_t) ny * z));
	if (left != 0 && x > 0) put(left, i - 1);
	if (diagonal != 0) put(diagonal, i);
	if (right != 0 && x < nx - 1) put(right, i + 1);
	if (left != 0 && x == 0) put(left, nx - 1 + (size_t) nx * (y + (size_t) ny * z));
	if (top != 0 && y < ny - 1) put(top, i + nx);
	if (front != 0 && z < nz - 1) put(front, i + (size_t) nx * ny);
	if (back != 0 && z == 0) put(back, x + (size_t) nx * (y + (size_t) ny * (nz - 1)));
}

__device__ inline void cell_xyz(size_t i, unsigned nx, unsigned ny, unsigned& x, unsigned& y, unsigned& z)
{
	x = (unsigned) (i % nx);
	y = (unsigned) ((i / nx) % ny);
	z = (unsigned) (i / ((size_t) nx * ny));
}

// dt/matrix.h:81-94 row sum: starts at 0, terms in column order
__device__ inline double h3_rowsum(const HRow& r, const double* __restrict__ x)
{
	double v = 0.0;
	for (int k = 0; k < r.cnt; ++k) v += r.a[k] * x[r.col[k]];
	return v;
}

__global__ void k_h3_apply(int kind, double scale, const double* __restrict__ x, double* __restrict__ y, unsigned nx,
                           unsigned ny, unsigned nz, unsigned n)
{
	const size_t N = (size_t) nx * ny * nz;
	const size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i >= N) return;
	unsigned a, b, c;
	cell_xyz(i, nx, ny, a, b, c);
	HRow r;
	helm3_row(a, b, c, nx, ny, nz, n, kind, scale, r);
	y[i] = h3_rowsum(r, x);
}

// ---------------------------------------------------------------- CG ----
// alg/cg.h:32-56.  r = b - A x, p = r; partial sums of r.r
__global__ __launch_bounds__(256) void k_h3_cg_init(int kind, double scale, const double* __restrict__ b,
                                                    const double* __restrict__ x, double* __restrict__ r,
                                                    double* __restrict__ p, unsigned nx, unsigned ny, unsigned nz,
                                                    unsigned n, double* partials)
{
	const size_t N = (size_t) nx * ny * nz;
	double acc[1] = {0.0};
	for (size_t i = blockIdx.x * 256 + threadIdx.x; i < N; i += (size_t) gridDim.x * 256) {
		unsigned a, bb, c;
		cell_xyz(i, nx, ny, a, bb, c);
		HRow row;
		helm3_row(a, bb, c, nx, ny, nz, n, kind, scale, row);
		const double ri = b[i] - h3_rowsum(row, x);
		r[i] = ri;
		p[i] = ri;
		acc[0] += ri * ri;
	}
	reduce_partial<1>(acc, partials);
}

// z = A p; partial sums of p.z
__global__ __launch_bounds__(256) void k_h3_cg_ap(int kind, double scale, const double* __restrict__ p,
                                                  double* __restrict__ z, unsigned nx, unsigned ny, unsigned nz,
                                                  unsigned n, double* partials)
{
	const size_t N = (size_t) nx * ny * nz;
	double acc[1] = {0.0};
	for (size_t i = blockIdx.x * 256 + threadIdx.x; i < N; i += (size_t) gridDim.x * 256) {
		unsigned a, b, c;
		cell_xyz(i, nx, ny, a, b, c);
		HRow row;
		helm3_row(a, b, c, nx, ny, nz, n, kind, scale, row);
		const double zi = h3_rowsum(row, p);
		z[i] = zi;
		acc[0] += p[i] * zi;
	}
	reduce_partial<1>(acc, partials);
}

// nu = n0 / (p.z); x += nu p; z *= nu; r -= z; partial sums of r.r
__global__ __launch_bounds__(256) void k_h3_cg_upd(const double* __restrict__ n0, const double* __restrict__ pz,
                                                   double* __restrict__ x, const double* __restrict__ p,
                                                   double* __restrict__ z, double* __restrict__ r, size_t N,
                                                   double* partials)
{
	const double nu = *n0 / *pz;
	double acc[1] = {0.0};
	for (size_t i = blockIdx.x * 256 + threadIdx.x; i < N; i += (size_t) gridDim.x * 256) {
		x[i] += nu * p[i];
		const double zi = z[i] * nu;
		z[i] = zi;
		const double ri = r[i] - zi;
		r[i] = ri;
		acc[0] += ri * ri;
	}
	reduce_partial<1>(acc, partials);
}

// mu = n1 / n0; p *= mu; p += r
__global__ void k_h3_cg_p(const double* __restrict__ n1, const double* __restrict__ n0, double* __restrict__ p,
                          const double* __restrict__ r, size_t N)
{
	const size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i >= N) return;
	const double mu = *n1 / *n0;
	const double t = p[i] * mu;
	p[i] = t + r[i];
}

__global__ __launch_bounds__(256) void k_h3_fin(unsigned nblk, const double* __restrict__ partials,
                                                double* __restrict__ out)
{
	double tot[1];
	red_total<1>(partials, nblk, tot);
	if (threadIdx.x == 0) *out = tot[0];
}

// ---------------------------------------------------- div, grad, advection
// alg/div.h:29-58 on nx x ny x nz cells: v has nx*(ny-1)*nz entries,
// top/bottom nx*nz (index x + nx*z)
__global__ void k_div3(const double* __restrict__ u, const double* __restrict__ v, const double* __restrict__ w,
                       const double* __restrict__ top, const double* __restrict__ bottom, double* __restrict__ d,
                       unsigned nx, unsigned ny, unsigned nz, unsigned n)
{
	const size_t N = (size_t) nx * ny * nz;
	const size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i >= N) return;
	unsigned x, y, z;
	cell_xyz(i, nx, ny, x, y, z);
	const size_t ixz = x + (size_t) nx * z;
	double dv;
	if (ny > 1) {
		if (y == 0) dv = v[i - (size_t) nx * z] - bottom[ixz];
		else if (y == ny - 1) dv = top[ixz] - v[i - (size_t) nx * (z + 1)];
		else dv = v[i - (size_t) nx * z] - v[i - (size_t) nx * (z + 1)];
	} else {
		dv = top[ixz] - bottom[ixz];
	}
	const size_t ixp = (x + 1) % nx + (size_t) nx * (y + (size_t) ny * z);
	const size_t izp = x + (size_t) nx * (y + (size_t) ny * ((z + 1) % nz));
	d[i] = n * ((u[ixp] - u[i]) + dv + (w[izp] - w[i]));
}

// alg/grad.h:32-43 on the n^3 cube
__global__ void k_grad3(const double* __restrict__ p, double* __restrict__ px, double* __restrict__ py,
                        double* __restrict__ pz, unsigned n)
{
	const size_t N = (size_t) n * n * n;
	const size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i >= N) return;
	unsigned x, y, z;
	cell_xyz(i, n, n, x, y, z);
	const size_t ixm = (x + n - 1) % n + (size_t) n * (y + (size_t) n * z);
	const size_t izm = x + (size_t) n * (y + (size_t) n * ((z + n - 1) % n));
	px[i] = n * (p[i] - p[ixm]);
	pz[i] = n * (p[i] - p[izm]);
	if (y > 0) py[i - (size_t) n * (z + 1)] = n * (p[i] - p[i - n]);
}

// ns/solver.h:156-198: the centred products (squares at cells, products at
// edges, v^2 on the walls into vvt / vvb)
__global__ void k_adv3_prod(const double* __restrict__ u, const double* __restrict__ v, const double* __restrict__ w,
                            double* __restrict__ uu, double* __restrict__ vv, double* __restrict__ ww,
                            double* __restrict__ uv, double* __restrict__ vw, double* __restrict__ uw,
                            double* __restrict__ vvt, double* __restrict__ vvb, unsigned n)
{
	const unsigned nx = n, ny = n, nz = n;
	const size_t N = (size_t) n * n * n;
	const size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i >= N) return;
	unsigned x, y, z;
	cell_xyz(i, nx, ny, x, y, z);
	const size_t ixz = x + (size_t) nx * z, iv = i - (size_t) nx * z;
	const size_t ixm = (x + nx - 1) % nx + (size_t) nx * (y + (size_t) ny * z);
	const size_t izm = x + (size_t) nx * (y + (size_t) ny * ((z + nz - 1) % nz));
	double q = 0.5 * (u[i] + u[ixm]);
	uu[i] = q * q;
	if (y == 0) {
		q = 0.5 * (v[iv]);
		vvb[ixz] = q * q;
	} else if (y == ny - 1) {
		q = 0.5 * (v[i - (size_t) nx * (z + 1)]);
		vvt[ixz] = q * q;
	} else {
		q = 0.5 * (v[iv] + v[i - (size_t) nx * (z + 1)]);
		vv[i - (size_t) nx * (2 * z + 1)] = q * q;
	}
	q = 0.5 * (w[i] + w[izm]);
	ww[i] = q * q;
	if (y < ny - 1) {
		const size_t ivxm = (x + nx - 1) % nx + (size_t) nx * (y + (size_t) (ny - 1) * z);
		const size_t ivzm = x + (size_t) nx * (y + (size_t) (ny - 1) * ((z + nz - 1) % nz));
		q = 0.5 * (u[i + nx] + u[i]);
		double r = 0.5 * (v[ivxm] + v[iv]);
		uv[iv] = q * r;
		q = 0.5 * (v[ivzm] + v[iv]);
		r = 0.5 * (w[i + nx] + w[i]);
		vw[iv] = q * r;
	}
	q = 0.5 * (u[izm] + u[i]);
	const double r = 0.5 * (w[ixm] + w[i]);
	uw[i] = q * r;
}

// projection::boundary (ns/projection.h:75-106), as written: nx*ny wall
// points, and the z-component's second-row terms read phidtx
__global__ void k_proj_bound3(const double* __restrict__ phx, const double* __restrict__ phz, double* __restrict__ ut,
                              double* __restrict__ ub, double* __restrict__ wt, double* __restrict__ wb, unsigned n)
{
	const unsigned nx = n, ny = n;
	const size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i >= (size_t) nx * ny) return;
	const unsigned x = (unsigned) (i % nx), z = (unsigned) (i / nx);
	const size_t ixzb = x + (size_t) nx * ny * z, ixzt = x + (size_t) nx * ((size_t) ny * (z + 1) - 1);
	const double pxb1 = phx[ixzb], pxb2 = phx[ixzb + nx], pxt1 = phx[ixzt], pxt2 = phx[ixzt - nx],
	             pzb1 = phz[ixzb], pzb2 = phx[ixzb + nx], pzt1 = phz[ixzt], pzt2 = phx[ixzt - nx];
	ut[i] += (9 * pxt1 - pxt2) / 8;
	ub[i] += (9 * pxb1 - pxb2) / 8;
	wt[i] += (9 * pzt1 - pzt2) / 8;
	wb[i] += (9 * pzb1 - pzb2) / 8;
}

// ns/solver.h:105-130, u or w (cell rows, kind 1).  Backward-Euler half step:
//   r = old + (-1 * boundary_s) + (dt/2) (f / rho - adv)
// Crank-Nicolson full step:
//   r = H_m old + (boundary_m - boundary_s) + dt (f / rho - adv)
// boundary_k: helmholtz_periodic::boundary (ns/helmholtz.h:177-196), nonzero
// only in the wall rows.
__global__ void k_rhs3_cell(int cn, const double* __restrict__ old, const double* __restrict__ adv,
                            const double* __restrict__ f, const double* __restrict__ top,
                            const double* __restrict__ bot, double* __restrict__ out, unsigned n, double sm,
                            double ss, double dt, double rho)
{
	const size_t N = (size_t) n * n * n;
	const size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i >= N) return;
	unsigned x, y, z;
	cell_xyz(i, n, n, x, y, z);
	const size_t ixz = x + (size_t) n * z;
	const double t = f[i] * (1 / rho) - adv[i];
	if (!cn) {
		double b = 0.0;
		if (y == 0) b = 8. / 3. * ss * n * n * bot[ixz];
		else if (y == n - 1) b = 8. / 3. * ss * n * n * top[ixz];
		b *= -1;
		out[i] = old[i] + b + t * (dt / 2);
	} else {
		double bm = 0.0, bs = 0.0;
		if (y == 0) { bm = 8. / 3. * sm * n * n * bot[ixz]; bs = 8. / 3. * ss * n * n * bot[ixz]; }
		else if (y == n - 1) { bm = 8. / 3. * sm * n * n * top[ixz]; bs = 8. / 3. * ss * n * n * top[ixz]; }
		HRow row;
		helm3_row(x, y, z, n, n, n, n, 1, sm, row);
		out[i] = h3_rowsum(row, old) + (bm - bs) + t * dt;
	}
}

// the same for v (interior y-face rows, kind 3; the reference adds no wall term)
__global__ void k_rhs3_face(int cn, const double* __restrict__ old, const double* __restrict__ adv,
                            const double* __restrict__ f, double* __restrict__ out, unsigned n, double sm, double dt,
                            double rho)
{
	const size_t Nv = (size_t) n * (n - 1) * n;
	const size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i >= Nv) return;
	const double t = f[i] * (1 / rho) - adv[i];
	if (!cn) {
		out[i] = old[i] + t * (dt / 2);
	} else {
		unsigned x, y, z;
		cell_xyz(i, n, n - 1, x, y, z);
		HRow row;
		helm3_row(x, y, z, n, n - 1, n, n, 3, sm, row);
		out[i] = h3_rowsum(row, old) + t * dt;
	}
}

// helmholtz_periodic::boundary (kind 1, ns/helmholtz.h:177-196: n^3 rows,
// 8/3 s n^2 times the wall value in the first and last row) and
// helmholtz_dirichlet::boundary (kind 3, ns/helmholtz.h:321-337: n (n-1) n
// rows, s n^2 times it); every other row 0.  One thread per column (x, z),
// bottom row written before the top one, as the reference loop does.
__global__ void k_hbound3(int kind, double scale, const double* __restrict__ top, const double* __restrict__ bot,
                          double* __restrict__ out, unsigned n)
{
	const size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i >= (size_t) n * n) return;
	const unsigned x = (unsigned) (i % n), z = (unsigned) (i / n);
	const unsigned ny = kind == 1 ? n : n - 1;
	for (unsigned y = 0; y < ny; ++y) out[x + (size_t) n * (y + (size_t) ny * z)] = 0.0;
	if (kind == 1) {
		out[x + (size_t) n * ny * z] = 8. / 3. * scale * n * n * bot[i];
		out[x + (size_t) n * (ny - 1 + (size_t) ny * z)] = 8. / 3. * scale * n * n * top[i];
	} else {
		out[x + (size_t) n * ny * z] = scale * n * n * bot[i];
		out[x + (size_t) n * (ny - 1 + (size_t) ny * z)] = scale * n * n * top[i];
	}
}

__global__ void k_sub3v(double* __restrict__ x, const double* __restrict__ y, size_t N)
{
	const size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i < N) x[i] -= y[i];
}

// ------------------------------------------------------------- host side --
static inline dim3 grid_of(size_t N) { return dim3((unsigned) ((N + 255) / 256)); }
static inline unsigned red_blocks(size_t N)
{
	const size_t b = (N + 255) / 256;
	return (unsigned) (b < (size_t) MAX_RED_BLOCKS ? b : (size_t) MAX_RED_BLOCKS);
}

static int read_dev_scalar(Ctx* c, const double* d, double* h)
{
	AFM_HIP(hipMemcpyAsync(h, d, 8, hipMemcpyDeviceToHost, c->stream));
	AFM_HIP(hipStreamSynchronize(c->stream));
	return 0;
}

// CG workspace: r, p, z of the largest system and 3 device scalars
struct CG3 {
	double *r = nullptr, *p = nullptr, *z = nullptr, *s = nullptr;
};

// alg::conjugate_gradient (alg/cg.h:32-56) on one Helmholtz system: the
// reference's loop and stop test sqrt(r.r) > tol (absolute); the scalars
// stay on the device (n0, n1 alternate between s[0] and s[1], p.z in s[2])
// and the host reads r.r once per iteration for the test.
static int h3_cg(Ctx* c, const CG3& w, int kind, double scale, const double* b, double* x, unsigned n, double tol,
                 int max_iters, int* iters, double* final_norm)
{
	const unsigned nx = n, ny = kind == 3 ? n - 1 : n, nz = n;
	const size_t N = (size_t) nx * ny * nz;
	const unsigned nb = red_blocks(N);
	hipLaunchKernelGGL(k_h3_cg_init, dim3(nb), dim3(256), 0, c->stream, kind, scale, b, x, w.r, w.p, nx, ny, nz, n,
	                   c->partials);
	hipLaunchKernelGGL(k_h3_fin, dim3(1), dim3(256), 0, c->stream, nb, c->partials, w.s);
	AFM_LAUNCH_CHECK();
	double n0 = 0.0;
	AFM_TRY(read_dev_scalar(c, w.s, &n0));
	int k = 0;
	while (sqrt(n0) > tol && k < max_iters) {
		double* s0 = w.s + (k & 1);
		double* s1 = w.s + ((k + 1) & 1);
		double* spz = w.s + 2;
		hipLaunchKernelGGL(k_h3_cg_ap, dim3(nb), dim3(256), 0, c->stream, kind, scale, w.p, w.z, nx, ny, nz, n,
		                   c->partials);
		hipLaunchKernelGGL(k_h3_fin, dim3(1), dim3(256), 0, c->stream, nb, c->partials, spz);
		hipLaunchKernelGGL(k_h3_cg_upd, dim3(nb), dim3(256), 0, c->stream, s0, spz, x, w.p, w.z, w.r, N, c->partials);
		hipLaunchKernelGGL(k_h3_fin, dim3(1), dim3(256), 0, c->stream, nb, c->partials, s1);
		hipLaunchKernelGGL(k_h3_cg_p, grid_of(N), dim3(256), 0, c->stream, s1, s0, w.p, w.r, N);
		AFM_LAUNCH_CHECK();
		AFM_TRY(read_dev_scalar(c, s1, &n0));
		++k;
	}
	if (iters) *iters = k;
	if (final_norm) *final_norm = sqrt(n0);
	if (sqrt(n0) > tol) {
		set_error("3-D Helmholtz CG did not converge within max_iters");
		return -4;
	}
	return 0;
}

struct NS3 {
	Ctx* c = nullptr;
	unsigned L = 0, n = 0;
	size_t N = 0, Nv = 0, Nxz = 0;
	double mu = 0, rho = 0, dt = 0, tol = 0, sm = 0, ss = 0;
	double *u = nullptr, *v = nullptr, *w = nullptr;    // the caller's, updated in place
	double *u_old = nullptr, *v_old = nullptr, *w_old = nullptr, *uadv = nullptr, *vadv = nullptr, *wadv = nullptr;
	double *phx = nullptr, *phy = nullptr, *phz = nullptr, *phi = nullptr;
	double *uu = nullptr, *ww = nullptr, *uw = nullptr, *uv = nullptr, *vw = nullptr, *vv = nullptr;
	double *vvt = nullptr, *vvb = nullptr, *zero = nullptr, *rx = nullptr, *ry = nullptr, *rz = nullptr;
	CG3 cg;
	Mg3* mg = nullptr;
	bool half = false;
	int last_cycles = 0, last_cg = 0;
	std::vector<void*> own;
};

static int ns3_alloc(NS3* s, double** p, size_t count)
{
	if (hipMalloc((void**) p, (count ? count : 1) * 8) != hipSuccess) {
		set_error("out of device memory (3-D Navier-Stokes)");
		return -3;
	}
	s->own.push_back(*p);
	return 0;
}

// projection::project (ns/projection.h:60-73): div with zero walls, the
// multigrid solve (absolute tolerance, V(2,2)), grad, subtract
static int ns3_project(NS3* s)
{
	Ctx* c = s->c;
	const unsigned n = s->n;
	hipLaunchKernelGGL(k_div3, grid_of(s->N), dim3(256), 0, c->stream, s->u, s->v, s->w, s->zero, s->zero, s->phi, n,
	                   n, n, n);
	AFM_LAUNCH_CHECK();
	AFM_TRY(mg3_solve_with(s->mg, s->phi, s->tol, 0, 1000, &s->last_cycles, nullptr));
	hipLaunchKernelGGL(k_grad3, grid_of(s->N), dim3(256), 0, c->stream, s->phi, s->phx, s->phy, s->phz, n);
	hipLaunchKernelGGL(k_sub3v, grid_of(s->N), dim3(256), 0, c->stream, s->u, s->phx, s->N);
	hipLaunchKernelGGL(k_sub3v, grid_of(s->Nv), dim3(256), 0, c->stream, s->v, s->phy, s->Nv);
	hipLaunchKernelGGL(k_sub3v, grid_of(s->N), dim3(256), 0, c->stream, s->w, s->phz, s->N);
	AFM_LAUNCH_CHECK();
	return 0;
}

static int ns3_advect(NS3* s)
{
	Ctx* c = s->c;
	const unsigned n = s->n;
	hipLaunchKernelGGL(k_adv3_prod, grid_of(s->N), dim3(256), 0, c->stream, s->u, s->v, s->w, s->uu, s->vv, s->ww,
	                   s->uv, s->vw, s->uw, s->vvt, s->vvb, n);
	hipLaunchKernelGGL(k_div3, grid_of(s->N), dim3(256), 0, c->stream, s->uu, s->uv, s->uw, s->zero, s->zero,
	                   s->uadv, n, n, n, n);
	hipLaunchKernelGGL(k_div3, grid_of(s->Nv), dim3(256), 0, c->stream, s->uv, s->vv, s->vw, s->vvt, s->vvb,
	                   s->vadv, n, n - 1, n, n);
	hipLaunchKernelGGL(k_div3, grid_of(s->N), dim3(256), 0, c->stream, s->uw, s->vw, s->ww, s->zero, s->zero,
	                   s->wadv, n, n, n, n);
	AFM_LAUNCH_CHECK();
	return 0;
}

void ns3_destroy(NS3* s)
{
	if (!s) return;
	mg3_destroy(s->mg);
	(void) hipStreamSynchronize(s->c->stream);
	for (void* p : s->own) (void) hipFree(p);
	delete s;
}

int ns3_create(Ctx* c, unsigned L, double mu, double rho, double dt, double tol, double* u, double* v, double* w,
               NS3** out)
{
	NS3* s = new NS3;
	s->c = c; s->L = L; s->n = 1u << L;
	const unsigned n = s->n;
	s->N = (size_t) n * n * n; s->Nv = (size_t) n * (n - 1) * n; s->Nxz = (size_t) n * n;
	s->mu = mu; s->rho = rho; s->dt = dt; s->tol = tol;
	s->sm = mu * dt / (2 * rho);
	s->ss = -mu * dt / (2 * rho);
	s->u = u; s->v = v; s->w = w;
	int rc = 0;
	const size_t N = s->N, Nv = s->Nv, Nxz = s->Nxz, Nvv = (size_t) n * (n - 2) * n;
	struct { double** p; size_t count; } bufs[] = {
		{&s->u_old, N}, {&s->v_old, Nv}, {&s->w_old, N}, {&s->uadv, N}, {&s->vadv, Nv}, {&s->wadv, N},
		{&s->phx, N}, {&s->phy, Nv}, {&s->phz, N}, {&s->phi, N}, {&s->uu, N}, {&s->ww, N}, {&s->uw, N},
		{&s->uv, Nv}, {&s->vw, Nv}, {&s->vv, Nvv}, {&s->vvt, Nxz}, {&s->vvb, Nxz}, {&s->zero, Nxz},
		{&s->rx, N}, {&s->ry, Nv}, {&s->rz, N}, {&s->cg.r, N}, {&s->cg.p, N}, {&s->cg.z, N}, {&s->cg.s, 4}};
	for (auto& b : bufs)
		if ((rc = ns3_alloc(s, b.p, b.count))) break;
	if (!rc) rc = mg3_create(c, L, 1, 2, 2, &s->mg);
	if (!rc && hipMemsetAsync(s->zero, 0, Nxz * 8, c->stream) != hipSuccess) { set_error("hipMemsetAsync failed"); rc = -2; }
	if (!rc && hipMemsetAsync(s->phx, 0, N * 8, c->stream) != hipSuccess) { set_error("hipMemsetAsync failed"); rc = -2; }
	if (!rc && hipMemsetAsync(s->phz, 0, N * 8, c->stream) != hipSuccess) { set_error("hipMemsetAsync failed"); rc = -2; }
	// ns/solver.h:71: the constructor projects the initial velocity
	if (!rc) rc = ns3_project(s);
	if (rc) { ns3_destroy(s); return rc; }
	*out = s;
	return 0;
}

// ns::solver::step (ns/solver.h:90-140)
int ns3_step(NS3* s, double* ut, double* ub, double* wt, double* wb, const double* fx, const double* fy,
             const double* fz)
{
	Ctx* c = s->c;
	const unsigned n = s->n;
	AFM_TRY(ns3_advect(s));
	if (!s->half) {
		AFM_HIP(hipMemcpyAsync(s->u_old, s->u, s->N * 8, hipMemcpyDeviceToDevice, c->stream));
		AFM_HIP(hipMemcpyAsync(s->v_old, s->v, s->Nv * 8, hipMemcpyDeviceToDevice, c->stream));
		AFM_HIP(hipMemcpyAsync(s->w_old, s->w, s->N * 8, hipMemcpyDeviceToDevice, c->stream));
	}
	const int cn = s->half ? 1 : 0;
	hipLaunchKernelGGL(k_proj_bound3, grid_of(s->Nxz), dim3(256), 0, c->stream, s->phx, s->phz, ut, ub, wt, wb, n);
	hipLaunchKernelGGL(k_rhs3_cell, grid_of(s->N), dim3(256), 0, c->stream, cn, s->u_old, s->uadv, fx, ut, ub, s->rx,
	                   n, s->sm, s->ss, s->dt, s->rho);
	hipLaunchKernelGGL(k_rhs3_face, grid_of(s->Nv), dim3(256), 0, c->stream, cn, s->v_old, s->vadv, fy, s->ry, n,
	                   s->sm, s->dt, s->rho);
	hipLaunchKernelGGL(k_rhs3_cell, grid_of(s->N), dim3(256), 0, c->stream, cn, s->w_old, s->wadv, fz, wt, wb, s->rz,
	                   n, s->sm, s->ss, s->dt, s->rho);
	AFM_LAUNCH_CHECK();
	int a = 0, b = 0, d = 0;
	AFM_TRY(h3_cg(c, s->cg, 1, s->ss, s->rx, s->u, n, s->tol, 100000, &a, nullptr));
	AFM_TRY(h3_cg(c, s->cg, 3, s->ss, s->ry, s->v, n, s->tol, 100000, &b, nullptr));
	AFM_TRY(h3_cg(c, s->cg, 1, s->ss, s->rz, s->w, n, s->tol, 100000, &d, nullptr));
	s->last_cg = a + b + d;
	AFM_TRY(ns3_project(s));
	s->half = !s->half;
	return 0;
}

// ns/solver.h:83: p = phhs * phidt
int ns3_get_p(NS3* s, double* p)
{
	const unsigned n = s->n;
	hipLaunchKernelGGL(k_h3_apply, grid_of(s->N), dim3(256), 0, s->c->stream, 1, s->ss, s->phi, p, n, n, n, n);
	AFM_LAUNCH_CHECK();
	return 0;
}

void ns3_stats(const NS3* s, int* cycles, int* cg, int* half)
{
	if (cycles) *cycles = s->last_cycles;
	if (cg) *cg = s->last_cg;
	if (half) *half = s->half ? 1 : 0;
}

// ------------------------------------------------ operator entry points --
int ns3_div(Ctx* c, unsigned L, const double* u, const double* v, const double* w, const double* top,
            const double* bottom, double* d)
{
	const unsigned n = 1u << L;
	hipLaunchKernelGGL(k_div3, grid_of((size_t) n * n * n), dim3(256), 0, c->stream, u, v, w, top, bottom, d, n, n, n,
	                   n);
	AFM_LAUNCH_CHECK();
	return 0;
}

int ns3_grad(Ctx* c, unsigned L, const double* p, double* px, double* py, double* pz)
{
	const unsigned n = 1u << L;
	hipLaunchKernelGGL(k_grad3, grid_of((size_t) n * n * n), dim3(256), 0, c->stream, p, px, py, pz, n);
	AFM_LAUNCH_CHECK();
	return 0;
}

int ns3_advect(Ctx* c, unsigned L, const double* u, const double* v, const double* w, double* ua, double* va,
               double* wa)
{
	NS3 t;
	t.c = c; t.n = 1u << L;
	const unsigned n = t.n;
	t.N = (size_t) n * n * n; t.Nv = (size_t) n * (n - 1) * n; t.Nxz = (size_t) n * n;
	t.u = const_cast<double*>(u); t.v = const_cast<double*>(v); t.w = const_cast<double*>(w);
	t.uadv = ua; t.vadv = va; t.wadv = wa;
	int rc = 0;
	struct { double** p; size_t count; } bufs[] = {
		{&t.uu, t.N}, {&t.ww, t.N}, {&t.uw, t.N}, {&t.uv, t.Nv}, {&t.vw, t.Nv}, {&t.vv, (size_t) n * (n - 2) * n},
		{&t.vvt, t.Nxz}, {&t.vvb, t.Nxz}, {&t.zero, t.Nxz}};
	for (auto& b : bufs)
		if ((rc = ns3_alloc(&t, b.p, b.count))) break;
	if (!rc && hipMemsetAsync(t.zero, 0, t.Nxz * 8, c->stream) != hipSuccess) { set_error("hipMemsetAsync failed"); rc = -2; }
	if (!rc) rc = ns3_advect(&t);
	(void) hipStreamSynchronize(c->stream);
	for (void* p : t.own) (void) hipFree(p);
	return rc;
}

int ns3_helmholtz_apply(Ctx* c, unsigned L, int kind, double scale, const double* x, double* y)
{
	const unsigned n = 1u << L, ny = kind == 3 ? n - 1 : n;
	hipLaunchKernelGGL(k_h3_apply, grid_of((size_t) n * ny * n), dim3(256), 0, c->stream, kind, scale, x, y, n, ny,
	                   n, n);
	AFM_LAUNCH_CHECK();
	return 0;
}

int ns3_helmholtz_cg(Ctx* c, unsigned L, int kind, double scale, const double* b, double* x, double tol,
                     int max_iters, int* iters, double* final_norm)
{
	const unsigned n = 1u << L;
	const size_t N = (size_t) n * n * n;
	NS3 t;
	t.c = c;
	int rc = 0;
	for (double** p : {&t.cg.r, &t.cg.p, &t.cg.z})
		if ((rc = ns3_alloc(&t, p, N))) break;
	if (!rc) rc = ns3_alloc(&t, &t.cg.s, 4);
	if (!rc) rc = h3_cg(c, t.cg, kind, scale, b, x, n, tol, max_iters, iters, final_norm);
	(void) hipStreamSynchronize(c->stream);
	for (void* p : t.own) (void) hipFree(p);
	return rc;
}

// ns::helmholtz_*::boundary(top, bottom) into out (n^3 or n (n-1) n values)
int ns3_helmholtz_boundary(Ctx* c, unsigned L, int kind, double scale, const double* top, const double* bottom,
                           double* out)
{
	const unsigned n = 1u << L;
	hipLaunchKernelGGL(k_hbound3, grid_of((size_t) n * n), dim3(256), 0, c->stream, kind, scale, top, bottom, out, n);
	AFM_LAUNCH_CHECK();
	return 0;
}

// projection::boundary (ns/projection.h:75-106) with the potential's
// gradients phx, phz (n^3 each): the wall vectors are corrected in place
int ns3_projection_boundary(Ctx* c, unsigned L, const double* phx, const double* phz, double* ut, double* ub,
                            double* wt, double* wb)
{
	const unsigned n = 1u << L;
	hipLaunchKernelGGL(k_proj_bound3, grid_of((size_t) n * n), dim3(256), 0, c->stream, phx, phz, ut, ub, wt, wb, n);
	AFM_LAUNCH_CHECK();
	return 0;
}

}  // namespace afm
