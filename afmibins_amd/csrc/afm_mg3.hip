// afm_mg3.hip -- the reference's native 3-D mode of the multigrid Poisson
// solve: mg::solver<G> (mg/solver.h:35-271) on a cube of s x s x s cells,
// s = c 2^L, periodic in x and z, walls (homogeneous Neumann) in y --
// geometry<L,c,c,c> of dt/geometry.h (c a power of two): <L,1,1,1> is the
// geometry of the reference's test/mg_test3.cc, <1,8,8,8> that of
// test/mg_test2.cc.  Level l has s_l = c 2^l cells per side and laplacian
// scale n_l = 2^l.
//
// One thread per cell, global memory: this is the parity path for the
// reference's own geometry, not the 2-D benchmark path, and it is not tiled.
// Every row uses the reference's arithmetic in its CSR column order:
//   laplacian    mg/laplacian.h:46-153 (kind 0: walls fold into the diagonal,
//                nx or nz of 1 or 2 fold their periodic entries);
//   smoother     alg/smooth.h:35-51 row update (sum of the off-diagonal terms
//                in column order, then (b - sum) / diagonal), red-black order
//                (rows of one colour are independent, so the GPU result
//                equals the same sweep done row by row);
//   restriction  mg/solver.h:153-181, 64 taps with y clamped, tap order;
//   prolongation mg/solver.h:71-151, trilinear on interior rows, bilinear on
//                the wall rows, written with the reference's own index
//                expression for those rows (it names ny where nx belongs;
//                equal on cubes, which is why only cubes are accepted);
//   level 0      mg/solver.h:264-271: mean removal and CG (alg/cg.h:32-56,
//                the right-hand side as the start) on the c^3 grid with
//                n = 1; for c = 1 the laplacian has no stored entry, so the
//                result is 0.
// The solve (mg/solver.h:218-237): nested start, mean removal, V-cycles on
// the residual until RMS(r) <= tol, with fixed-order GPU reductions (the
// oracle sums sequentially, so solves agree to rounding, V-cycles exactly).
#include "afm_device.h"

namespace afm {

// One row of mg::laplacian (kind 0) on an nx x ny x nz grid: up to 7 entries
// in ascending column order.
struct Row7 {
	size_t col[7];
	double a[7];
	int cnt;
};

__device__ inline void lap3_row(unsigned x, unsigned y, unsigned z, unsigned nx, unsigned ny, unsigned nz, double n,
                                Row7& r)
{
	const size_t i = x + (size_t) nx * (y + (size_t) ny * z);
	const double off = n * n;
	double diagonal = -6.0 * off;
	double left = off, right = off, top = off, bottom = off, front = off, back = off;
	if (nx == 1) { diagonal += left + right; left = right = 0; }
	else if (nx == 2) { if (x == 0) { right += left; left = 0; } else { left += right; right = 0; } }
	if (y == 0) { diagonal += bottom; bottom = 0; }
	if (y == ny - 1) { diagonal += top; top = 0; }
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

__device__ inline void cell3(size_t i, unsigned nx, unsigned ny, unsigned& x, unsigned& y, unsigned& z)
{
	x = (unsigned) (i % nx);
	y = (unsigned) ((i / nx) % ny);
	z = (unsigned) (i / ((size_t) nx * ny));
}

// y = L x (dt/matrix.h:81-94: the row sum starts at 0, terms in column order)
// on s^3 cells with laplacian scale n
__global__ void k_lap3(const double* __restrict__ x, double* __restrict__ y, unsigned s, unsigned n)
{
	const size_t N = (size_t) s * s * s;
	const size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i >= N) return;
	unsigned a, b, c;
	cell3(i, s, s, a, b, c);
	Row7 r;
	lap3_row(a, b, c, s, s, s, (double) n, r);
	double v = 0.0;
	for (int k = 0; k < r.cnt; ++k) v += r.a[k] * x[r.col[k]];
	y[i] = v;
}

// r = b - L x
__global__ void k_resid3(const double* __restrict__ x, const double* __restrict__ b, double* __restrict__ r,
                         unsigned s, unsigned n)
{
	const size_t N = (size_t) s * s * s;
	const size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i >= N) return;
	unsigned a, bb, c;
	cell3(i, s, s, a, bb, c);
	Row7 row;
	lap3_row(a, bb, c, s, s, s, (double) n, row);
	double v = 0.0;
	for (int k = 0; k < row.cnt; ++k) v += row.a[k] * x[row.col[k]];
	r[i] = b[i] - v;
}

// rows of one colour ((x + y + z) & 1 == colour), alg/smooth.h:39-50
__global__ void k_rb3(double* __restrict__ x, const double* __restrict__ b, unsigned s, unsigned n, unsigned colour)
{
	const size_t N = (size_t) s * s * s;
	const size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i >= N) return;
	unsigned a, bb, c;
	cell3(i, s, s, a, bb, c);
	if (((a + bb + c) & 1u) != colour) return;
	Row7 row;
	lap3_row(a, bb, c, s, s, s, (double) n, row);
	double value = 0.0, diagonal = 0.0;
	for (int k = 0; k < row.cnt; ++k) {
		if (row.col[k] == i) { diagonal = row.a[k]; continue; }
		value += row.a[k] * x[row.col[k]];
	}
	x[i] = (b[i] - value) / diagonal;
}

// coarse cell of a (2n)^3 fine grid: 64 taps, y clamped, weights / 512
__global__ void k_restrict3(const double* __restrict__ in, double* __restrict__ out, unsigned nf)
{
	const unsigned nc = nf / 2;
	const size_t N = (size_t) nc * nc * nc;
	const size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i >= N) return;
	unsigned xcc, ycc, zcc;
	cell3(i, nc, nc, xcc, ycc, zcc);
	const unsigned xcf = 2 * xcc + 1, ycf = 2 * ycc + 1, zcf = 2 * zcc + 1;
	double o = 0.0;
	for (int t = 0; t < 64; ++t) {
		const int dx = (t % 4) - 2, dy = (t % 16) / 4 - 2, dz = t / 16 - 2;
		const unsigned x = (xcf + dx + nf) % nf;
		const int yl = -((int) ycf);
		unsigned y = ycf + (dy > yl ? dy : yl);
		if (y > nf - 1) y = nf - 1;
		const unsigned z = (zcf + dz + nf) % nf;
		const int w = (4 - abs(2 * dx + 1)) * (4 - abs(2 * dy + 1)) * (4 - abs(2 * dz + 1));
		o += w / 512. * in[x + (size_t) nf * (y + (size_t) nf * z)];
	}
	out[i] = o;
}

// prolongation, interior rows: one coarse cell writes its 8 fine cells
__global__ void k_prolong3_interior(const double* __restrict__ in, double* __restrict__ out, unsigned nf)
{
	const unsigned h = nf / 2;
	if (h < 2) return;
	const size_t N = (size_t) h * (h - 1) * h;          // ycc < h - 1
	const size_t t = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (t >= N) return;
	const unsigned xcc = (unsigned) (t % h), ycc = (unsigned) ((t / h) % (h - 1)), zcc = (unsigned) (t / ((size_t) h * (h - 1)));
	const unsigned xnc = (xcc + 1) % h, ync = ycc + 1, znc = (zcc + 1) % h;
	const unsigned xcf = 2 * xcc + 1, ycf = 2 * ycc + 1, zcf = 2 * zcc + 1;
	const unsigned xnf = (xcf + 1) % nf, ynf = ycf + 1, znf = (zcf + 1) % nf;
	auto C = [&](unsigned a, unsigned b, unsigned c) { return in[a + (size_t) h * (b + (size_t) h * c)]; };
	auto F = [&](unsigned a, unsigned b, unsigned c) -> double& { return out[a + (size_t) nf * (b + (size_t) nf * c)]; };
	const double blbc = C(xcc, ycc, zcc), brbc = C(xnc, ycc, zcc), ulbc = C(xcc, ync, zcc), urbc = C(xnc, ync, zcc),
	             blfc = C(xcc, ycc, znc), brfc = C(xnc, ycc, znc), ulfc = C(xcc, ync, znc), urfc = C(xnc, ync, znc);
	F(xcf, ycf, zcf) = 27. / 64. * blbc + 9. / 64. * (brbc + ulbc + blfc) + 3. / 64. * (urbc + ulfc + brfc) + 1. / 64. * urfc;
	F(xnf, ycf, zcf) = 27. / 64. * brbc + 9. / 64. * (urbc + blbc + brfc) + 3. / 64. * (ulbc + blfc + urfc) + 1. / 64. * ulfc;
	F(xcf, ynf, zcf) = 27. / 64. * ulbc + 9. / 64. * (urbc + blbc + ulfc) + 3. / 64. * (brbc + blfc + urfc) + 1. / 64. * brfc;
	F(xnf, ynf, zcf) = 27. / 64. * urbc + 9. / 64. * (brbc + ulbc + urfc) + 3. / 64. * (blbc + ulfc + brfc) + 1. / 64. * blfc;
	F(xcf, ycf, znf) = 27. / 64. * blfc + 9. / 64. * (brfc + ulfc + blbc) + 3. / 64. * (urfc + ulbc + brbc) + 1. / 64. * urbc;
	F(xnf, ycf, znf) = 27. / 64. * brfc + 9. / 64. * (urfc + blfc + brbc) + 3. / 64. * (ulfc + blbc + urbc) + 1. / 64. * ulbc;
	F(xcf, ynf, znf) = 27. / 64. * ulfc + 9. / 64. * (urfc + blfc + ulbc) + 3. / 64. * (brfc + blbc + urbc) + 1. / 64. * brbc;
	F(xnf, ynf, znf) = 27. / 64. * urfc + 9. / 64. * (brfc + ulfc + urbc) + 3. / 64. * (blfc + ulbc + brbc) + 1. / 64. * blbc;
}

// prolongation, the two wall rows (bilinear in x and z); the index of the
// top row as the reference writes it (mg/solver.h:141-148)
__global__ void k_prolong3_walls(const double* __restrict__ in, double* __restrict__ out, unsigned nf)
{
	const unsigned h = nf / 2;
	const size_t N = (size_t) h * h;
	const size_t t = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (t >= N) return;
	const unsigned xcc = (unsigned) (t % h), zcc = (unsigned) (t / h);
	const unsigned nx = nf, ny = nf;
	const unsigned hx = h, hy = h;
	const unsigned xnc = (xcc + 1) % hx, znc = (zcc + 1) % h;
	const unsigned xcf = 2 * xcc + 1, zcf = 2 * zcc + 1;
	const unsigned xnf = (xcf + 1) % nx, znf = (zcf + 1) % nf;
	const double blbc = in[xcc + (size_t) hx * hy * zcc], brbc = in[xnc + (size_t) hx * hy * zcc],
	             blfc = in[xcc + (size_t) hx * hy * znc], brfc = in[xnc + (size_t) hx * hy * znc],
	             ulbc = in[xcc + (size_t) hx * (hy - 1 + (size_t) hy * zcc)],
	             urbc = in[xnc + (size_t) hx * (hy - 1 + (size_t) hy * zcc)],
	             ulfc = in[xcc + (size_t) hx * (hy - 1 + (size_t) hy * znc)],
	             urfc = in[xnc + (size_t) hx * (hy - 1 + (size_t) hy * znc)];
	out[xcf + (size_t) nx * ny * zcf] = 9. / 16. * blbc + 3. / 16. * (brbc + blfc) + 1. / 16. * brfc;
	out[xnf + (size_t) nx * ny * zcf] = 9. / 16. * brbc + 3. / 16. * (blbc + brfc) + 1. / 16. * blfc;
	out[xcf + (size_t) nx * ny * znf] = 9. / 16. * blfc + 3. / 16. * (brfc + blbc) + 1. / 16. * brbc;
	out[xnf + (size_t) nx * ny * znf] = 9. / 16. * brfc + 3. / 16. * (blfc + brbc) + 1. / 16. * blbc;
	out[xcf + (size_t) ny * (ny - 1 + (size_t) ny * zcf)] = 9. / 16. * ulbc + 3. / 16. * (urbc + ulfc) + 1. / 16. * urfc;
	out[xnf + (size_t) ny * (ny - 1 + (size_t) ny * zcf)] = 9. / 16. * urbc + 3. / 16. * (ulbc + urfc) + 1. / 16. * ulfc;
	out[xcf + (size_t) ny * (ny - 1 + (size_t) ny * znf)] = 9. / 16. * ulfc + 3. / 16. * (urfc + ulbc) + 1. / 16. * urbc;
	out[xnf + (size_t) ny * (ny - 1 + (size_t) ny * znf)] = 9. / 16. * urfc + 3. / 16. * (ulfc + urbc) + 1. / 16. * ulbc;
}

__global__ void k_axpy3(double* __restrict__ x, const double* __restrict__ e, size_t N)
{
	const size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i < N) x[i] += e[i];
}

__global__ void k_fill3(double* __restrict__ x, double v, size_t N)
{
	const size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i < N) x[i] = v;
}

__global__ void k_copy3(double* __restrict__ dst, const double* __restrict__ src, size_t N)
{
	const size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i < N) dst[i] = src[i];
}

// partial sums of x (sq = 0) or x^2 (sq = 1), MAX_RED_BLOCKS at most
__global__ __launch_bounds__(256) void k_sum3(const double* __restrict__ x, size_t N, int sq, double* partials)
{
	double acc[1] = {0.0};
	for (size_t i = blockIdx.x * 256 + threadIdx.x; i < N; i += (size_t) gridDim.x * 256)
		acc[0] += sq ? x[i] * x[i] : x[i] * 1.0;
	reduce_partial<1>(acc, partials);
}

__global__ __launch_bounds__(256) void k_sum3_fin(unsigned nblk, const double* __restrict__ partials,
                                                  double* __restrict__ out)
{
	double tot[1];
	red_total<1>(partials, nblk, tot);
	if (threadIdx.x == 0) *out = tot[0];
}

// x -= c (the mean, read on the device)
__global__ void k_sub3(double* __restrict__ x, const double* __restrict__ c, double scale, size_t N)
{
	const size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i < N) x[i] -= (*c * scale) * 1.0;
}

// partial sums of a[i] * b[i]
__global__ __launch_bounds__(256) void k_dot3(const double* __restrict__ a, const double* __restrict__ b, size_t N,
                                              double* partials)
{
	double acc[1] = {0.0};
	for (size_t i = blockIdx.x * 256 + threadIdx.x; i < N; i += (size_t) gridDim.x * 256) acc[0] += a[i] * b[i];
	reduce_partial<1>(acc, partials);
}

// alg/cg.h:45-47: x += nu p; z *= nu; r -= z
__global__ void k_cg3_step(double* __restrict__ x, double* __restrict__ r, const double* __restrict__ p,
                           double* __restrict__ z, double nu, size_t N)
{
	const size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i >= N) return;
	x[i] += nu * p[i];
	const double zi = z[i] * nu;
	z[i] = zi;
	r[i] -= zi;
}

// alg/cg.h:51-52: p *= mu; p += r
__global__ void k_cg3_dir(double* __restrict__ p, const double* __restrict__ r, double mu, size_t N)
{
	const size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i >= N) return;
	p[i] = p[i] * mu;
	p[i] = p[i] + r[i];
}

// ------------------------------------------------------------- host side --
static inline dim3 g3(size_t N) { return dim3((unsigned) ((N + 255) / 256)); }

static unsigned red3(size_t N)
{
	size_t b = (N + 255) / 256;
	return (unsigned) (b < (size_t) MAX_RED_BLOCKS ? b : (size_t) MAX_RED_BLOCKS);
}

// Level buffers of one call, [0..L]: X the level's iterate, B the level's
// right-hand side (written by the restriction from the level above), R the
// level's residual / prolongated correction (recurse) or right-hand-side
// copy (nested).  A level's rhs passed to recurse3 must not be its R.
struct Mg3Ctx {
	Ctx* c;
	unsigned L, W;           // levels, cube width c (cells per side at level 0)
	int nu1, nu2;
	double tol0 = 0.0;       // level 0's CG tolerance (the solve's absolute one)
	std::vector<double*> X, B, R;
	double* Z0 = nullptr;    // level 0's CG direction product
	double* scal;
	unsigned side(unsigned l) const { return W << l; }
	size_t cells(unsigned l) const { const size_t s = side(l); return s * s * s; }
};

// The workspace of mg3_solve_with: the levels, plus the iterate, the
// residual and a scratch copy of the right-hand side at level L, apart from
// the level buffers that recurse3 / nested3 use.
struct Mg3 {
	Mg3Ctx m;
	double *xs = nullptr, *res = nullptr, *copyb = nullptr;
	std::vector<void*> own;
};

// sum (sq = 0) or sum of squares (sq = 1) of a device vector into *out (device)
static int sum3(Ctx* c, const double* x, size_t N, int sq, double* out)
{
	hipLaunchKernelGGL(k_sum3, dim3(red3(N)), dim3(256), 0, c->stream, x, N, sq, c->partials);
	hipLaunchKernelGGL(k_sum3_fin, dim3(1), dim3(256), 0, c->stream, red3(N), c->partials, out);
	AFM_LAUNCH_CHECK();
	return 0;
}

// one red-black sweep of L x = b on s^3 cells with scale n
static int smooth3(Ctx* c, double* x, const double* b, unsigned s, unsigned n)
{
	const size_t N = (size_t) s * s * s;
	for (unsigned colour = 0; colour < 2; ++colour)
		hipLaunchKernelGGL(k_rb3, g3(N), dim3(256), 0, c->stream, x, b, s, n, colour);
	AFM_LAUNCH_CHECK();
	return 0;
}

static int read_scalar(Ctx* c, const double* d, double* h)
{
	AFM_HIP(hipMemcpyAsync(h, d, 8, hipMemcpyDeviceToHost, c->stream));
	AFM_HIP(hipStreamSynchronize(c->stream));
	return 0;
}

static int dot3(Ctx* c, const double* a, const double* b, size_t N, double* dev, double* host)
{
	hipLaunchKernelGGL(k_dot3, dim3(red3(N)), dim3(256), 0, c->stream, a, b, N, c->partials);
	hipLaunchKernelGGL(k_sum3_fin, dim3(1), dim3(256), 0, c->stream, red3(N), c->partials, dev);
	AFM_LAUNCH_CHECK();
	return read_scalar(c, dev, host);
}

// alg::conjugate_gradient (alg/cg.h:32-56) on the laplacian of s^3 cells with
// scale n: x holds the start and receives the result; r, p, z are scratch of
// s^3 doubles; stop when sqrt(r.r) <= tol.  Returns -4 past max_iters.
static int cg3_lap(Ctx* c, unsigned s, unsigned n, double* x, const double* b, double* r, double* p, double* z,
                   double* scal, double tol, int max_iters, int* iters, double* final_norm)
{
	const size_t N = (size_t) s * s * s;
	hipLaunchKernelGGL(k_resid3, g3(N), dim3(256), 0, c->stream, x, b, r, s, n);
	hipLaunchKernelGGL(k_copy3, g3(N), dim3(256), 0, c->stream, p, r, N);
	AFM_LAUNCH_CHECK();
	double n0 = 0.0, pz = 0.0;
	AFM_TRY(dot3(c, r, r, N, scal, &n0));
	int count = 0;
	while (sqrt(n0) > tol) {
		if (count >= max_iters) {
			if (final_norm) *final_norm = sqrt(n0);
			if (iters) *iters = count;
			set_error("3-D CG did not converge within max_iters");
			return -4;
		}
		hipLaunchKernelGGL(k_lap3, g3(N), dim3(256), 0, c->stream, p, z, s, n);
		AFM_LAUNCH_CHECK();
		AFM_TRY(dot3(c, p, z, N, scal, &pz));
		const double nu = n0 / pz;
		hipLaunchKernelGGL(k_cg3_step, g3(N), dim3(256), 0, c->stream, x, r, p, z, nu, N);
		AFM_LAUNCH_CHECK();
		double n1 = 0.0;
		AFM_TRY(dot3(c, r, r, N, scal, &n1));
		const double mu = n1 / n0;
		n0 = n1;
		hipLaunchKernelGGL(k_cg3_dir, g3(N), dim3(256), 0, c->stream, p, r, mu, N);
		AFM_LAUNCH_CHECK();
		++count;
	}
	if (iters) *iters = count;
	if (final_norm) *final_norm = sqrt(n0);
	return 0;
}

// recurse(l, rhs) (mg/solver.h:199-216): rhs replaced by the level's result
static int recurse3(Mg3Ctx& m, unsigned l, double* rhs)
{
	Ctx* c = m.c;
	const size_t N = m.cells(l);
	if (l == 0) {
		if (m.W == 1) {
			// mean removal of the 1-cell grid leaves 0; CG stops at once (no entry)
			hipLaunchKernelGGL(k_fill3, dim3(1), dim3(64), 0, c->stream, rhs, 0.0, (size_t) 1);
			AFM_LAUNCH_CHECK();
			return 0;
		}
		// mg/solver.h:264-271: rhs -= mean; CG with rhs as start and right-hand side
		AFM_TRY(sum3(c, rhs, N, 0, m.scal));
		hipLaunchKernelGGL(k_sub3, g3(N), dim3(256), 0, c->stream, rhs, m.scal, 1.0 / (double) N, N);
		hipLaunchKernelGGL(k_copy3, g3(N), dim3(256), 0, c->stream, m.B[0], rhs, N);
		AFM_LAUNCH_CHECK();
		return cg3_lap(c, m.W, 1, rhs, m.B[0], m.R[0], m.X[0], m.Z0, m.scal + 2, m.tol0, 100000, nullptr, nullptr);
	}
	double* x = m.X[l];
	double* rf = m.R[l];
	hipLaunchKernelGGL(k_fill3, g3(N), dim3(256), 0, c->stream, x, 0.0, N);
	for (int j = 0; j < m.nu1; ++j) AFM_TRY(smooth3(c, x, rhs, m.side(l), 1u << l));
	hipLaunchKernelGGL(k_resid3, g3(N), dim3(256), 0, c->stream, x, rhs, rf, m.side(l), 1u << l);
	hipLaunchKernelGGL(k_restrict3, g3(m.cells(l - 1)), dim3(256), 0, c->stream, rf, m.B[l - 1], m.side(l));
	AFM_LAUNCH_CHECK();
	AFM_TRY(recurse3(m, l - 1, m.B[l - 1]));
	const unsigned nf = m.side(l), h = nf / 2;
	if (h >= 2)
		hipLaunchKernelGGL(k_prolong3_interior, g3((size_t) h * (h - 1) * h), dim3(256), 0, c->stream, m.B[l - 1], rf, nf);
	hipLaunchKernelGGL(k_prolong3_walls, g3((size_t) h * h), dim3(256), 0, c->stream, m.B[l - 1], rf, nf);
	hipLaunchKernelGGL(k_axpy3, g3(N), dim3(256), 0, c->stream, x, rf, N);
	AFM_LAUNCH_CHECK();
	for (int j = 0; j < m.nu2; ++j) AFM_TRY(smooth3(c, x, rhs, m.side(l), 1u << l));
	hipLaunchKernelGGL(k_copy3, g3(N), dim3(256), 0, c->stream, rhs, x, N);
	AFM_LAUNCH_CHECK();
	return 0;
}

// nested(l, input) (mg/solver.h:183-197)
static int nested3(Mg3Ctx& m, unsigned l, double* input)
{
	if (l == 0) return 0;
	Ctx* c = m.c;
	const size_t N = m.cells(l);
	double* copy = m.R[l];
	hipLaunchKernelGGL(k_copy3, g3(N), dim3(256), 0, c->stream, copy, input, N);
	hipLaunchKernelGGL(k_restrict3, g3(m.cells(l - 1)), dim3(256), 0, c->stream, input, m.B[l - 1], m.side(l));
	AFM_LAUNCH_CHECK();
	AFM_TRY(nested3(m, l - 1, m.B[l - 1]));
	AFM_TRY(recurse3(m, l - 1, m.B[l - 1]));
	const unsigned nf = m.side(l), h = nf / 2;
	if (h >= 2)
		hipLaunchKernelGGL(k_prolong3_interior, g3((size_t) h * (h - 1) * h), dim3(256), 0, c->stream, m.B[l - 1], input, nf);
	hipLaunchKernelGGL(k_prolong3_walls, g3((size_t) h * h), dim3(256), 0, c->stream, m.B[l - 1], input, nf);
	AFM_LAUNCH_CHECK();
	for (int j = 0; j < m.nu2; ++j) AFM_TRY(smooth3(c, input, copy, m.side(l), 1u << l));
	return 0;
}

static int mg3_alloc(Ctx* c, unsigned L, unsigned W, int nu1, int nu2, Mg3Ctx& m, std::vector<void*>& own)
{
	m.c = c; m.L = L; m.W = W; m.nu1 = nu1; m.nu2 = nu2;
	m.X.assign(L + 1, nullptr); m.B.assign(L + 1, nullptr); m.R.assign(L + 1, nullptr);
	auto get = [&](double** p, size_t count) -> int {
		if (hipMalloc((void**) p, count * 8) != hipSuccess) { set_error("out of device memory (3-D multigrid)"); return -3; }
		own.push_back(*p);
		return 0;
	};
	for (unsigned l = 0; l <= L; ++l)
		for (double** p : {&m.X[l], &m.B[l], &m.R[l]}) AFM_TRY(get(p, m.cells(l)));
	AFM_TRY(get(&m.Z0, m.cells(0)));
	AFM_TRY(get(&m.scal, 8));
	return 0;
}

static void mg3_free(Ctx* c, std::vector<void*>& own)
{
	(void) hipStreamSynchronize(c->stream);
	for (void* p : own) (void) hipFree(p);
}

int mg3_laplacian(Ctx* c, unsigned L, unsigned W, const double* x, double* y)
{
	const unsigned s = W << L;
	hipLaunchKernelGGL(k_lap3, g3((size_t) s * s * s), dim3(256), 0, c->stream, x, y, s, 1u << L);
	AFM_LAUNCH_CHECK();
	return 0;
}

// one recurse(L, rhs) (mg/solver.h:199-216); tol is level 0's CG tolerance
int mg3_vcycle(Ctx* c, unsigned L, unsigned W, double* rhs, int nu1, int nu2, double tol)
{
	Mg3Ctx m;
	std::vector<void*> own;
	int rc = mg3_alloc(c, L, W, nu1, nu2, m, own);
	m.tol0 = tol;
	if (!rc) rc = recurse3(m, L, rhs);
	mg3_free(c, own);
	return rc;
}

// mg::solver<G>::restriction (mg/solver.h:153-181): fine level L >= 1 to L-1
int mg3_restrict(Ctx* c, unsigned L, unsigned W, const double* fine, double* coarse)
{
	const unsigned s = W << L, h = s / 2;
	hipLaunchKernelGGL(k_restrict3, g3((size_t) h * h * h), dim3(256), 0, c->stream, fine, coarse, s);
	AFM_LAUNCH_CHECK();
	return 0;
}

// mg::solver<G>::interpolation (mg/solver.h:71-151): level L-1 to L (every
// fine row is written)
int mg3_prolong(Ctx* c, unsigned L, unsigned W, const double* coarse, double* fine)
{
	const unsigned nf = W << L, h = nf / 2;
	if (h >= 2)
		hipLaunchKernelGGL(k_prolong3_interior, g3((size_t) h * (h - 1) * h), dim3(256), 0, c->stream, coarse, fine, nf);
	hipLaunchKernelGGL(k_prolong3_walls, g3((size_t) h * h), dim3(256), 0, c->stream, coarse, fine, nf);
	AFM_LAUNCH_CHECK();
	return 0;
}

// red-black Gauss-Seidel sweeps of L x = b at level L (alg/smooth.h:35-51 per row)
int mg3_smooth(Ctx* c, unsigned L, unsigned W, double* x, const double* b, int sweeps)
{
	for (int k = 0; k < sweeps; ++k) AFM_TRY(smooth3(c, x, b, W << L, 1u << L));
	return 0;
}

// alg::conjugate_gradient on the level-L laplacian: x is the start and the result
int mg3_cg(Ctx* c, unsigned L, unsigned W, const double* b, double* x, double tol, int max_iters, int* iters,
           double* final_norm)
{
	const unsigned s = W << L;
	const size_t N = (size_t) s * s * s;
	std::vector<void*> own;
	double *r = nullptr, *p = nullptr, *z = nullptr, *scal = nullptr;
	int rc = 0;
	for (double** q : {&r, &p, &z, &scal}) {
		if (hipMalloc((void**) q, (q == &scal ? 8 : N) * 8) != hipSuccess) { set_error("out of device memory (3-D CG)"); rc = -3; break; }
		own.push_back(*q);
	}
	if (!rc) rc = cg3_lap(c, s, 1u << L, x, b, r, p, z, scal, tol, max_iters, iters, final_norm);
	mg3_free(c, own);
	return rc;
}

int mg3_create(Ctx* c, unsigned L, unsigned W, int nu1, int nu2, Mg3** out)
{
	Mg3* g = new Mg3;
	int rc = mg3_alloc(c, L, W, nu1, nu2, g->m, g->own);
	const size_t N = g->m.cells(L);
	for (double** p : {&g->xs, &g->res, &g->copyb}) {
		if (rc) break;
		if (hipMalloc((void**) p, N * 8) != hipSuccess) { set_error("out of device memory (3-D multigrid)"); rc = -3; break; }
		g->own.push_back(*p);
	}
	if (rc) { mg3_destroy(g); return rc; }
	*out = g;
	return 0;
}

void mg3_destroy(Mg3* g)
{
	if (!g) return;
	mg3_free(g->m.c, g->own);
	delete g;
}

// mg::solver<G>::solve (mg/solver.h:218-237): rhs replaced by the solution
int mg3_solve(Ctx* c, unsigned L, unsigned W, double* rhs, double tol, int tol_mode, int nu1, int nu2,
              int max_cycles, int* cycles, double* final_rms)
{
	Mg3* g = nullptr;
	int rc = mg3_create(c, L, W, nu1, nu2, &g);
	if (rc) return rc;
	rc = mg3_solve_with(g, rhs, tol, tol_mode, max_cycles, cycles, final_rms);
	mg3_destroy(g);
	return rc;
}

int mg3_solve_with(Mg3* g, double* rhs, double tol, int tol_mode, int max_cycles, int* cycles, double* final_rms)
{
	Mg3Ctx& m = g->m;
	Ctx* c = m.c;
	const unsigned L = m.L, s = m.side(L), n = 1u << L;
	const size_t N = m.cells(L);
	double* sc = m.scal;
	double* xs = g->xs;
	double* res = g->res;
	double* copyb = g->copyb;
	double h_sq = 0.0;
	// the relative tolerance: RMS of rhs - mean
	AFM_TRY(sum3(c, rhs, N, 0, sc));
	hipLaunchKernelGGL(k_copy3, g3(N), dim3(256), 0, c->stream, copyb, rhs, N);
	hipLaunchKernelGGL(k_sub3, g3(N), dim3(256), 0, c->stream, copyb, sc, 1.0 / (double) N, N);
	AFM_TRY(sum3(c, copyb, N, 1, sc + 1));
	AFM_TRY(read_scalar(c, sc + 1, &h_sq));
	const double t = tol_mode == 1 ? tol * (sqrt(h_sq) / sqrt((double) N)) : tol;
	m.tol0 = t;      // every level's solver holds the same tolerance (mg/solver.h:60-62)
	// x = rhs; nested(L, x)
	hipLaunchKernelGGL(k_copy3, g3(N), dim3(256), 0, c->stream, xs, rhs, N);
	AFM_TRY(nested3(m, L, xs));
	// rhs -= mean (c = dot(rhs, ones) / N, as the reference computes it)
	AFM_TRY(sum3(c, rhs, N, 0, sc));
	hipLaunchKernelGGL(k_sub3, g3(N), dim3(256), 0, c->stream, rhs, sc, 1.0 / (double) N, N);
	hipLaunchKernelGGL(k_resid3, g3(N), dim3(256), 0, c->stream, xs, rhs, res, s, n);
	AFM_TRY(sum3(c, res, N, 1, sc + 1));
	AFM_TRY(read_scalar(c, sc + 1, &h_sq));
	double r = sqrt(h_sq) / sqrt((double) N);
	int count = 0;
	while (r > t && count < max_cycles) {
		AFM_TRY(recurse3(m, L, res));
		hipLaunchKernelGGL(k_axpy3, g3(N), dim3(256), 0, c->stream, xs, res, N);
		hipLaunchKernelGGL(k_resid3, g3(N), dim3(256), 0, c->stream, xs, rhs, res, s, n);
		AFM_TRY(sum3(c, res, N, 1, sc + 1));
		AFM_TRY(read_scalar(c, sc + 1, &h_sq));
		r = sqrt(h_sq) / sqrt((double) N);
		++count;
	}
	hipLaunchKernelGGL(k_copy3, g3(N), dim3(256), 0, c->stream, rhs, xs, N);
	AFM_LAUNCH_CHECK();
	if (cycles) *cycles = count;
	if (final_rms) *final_rms = r;
	if (r <= t) return 0;
	set_error("3-D multigrid did not converge within max_cycles");
	return -4;
}

}  // namespace afm
