// afm_csr.hip -- dt::matrix (dt/matrix.h:35-94) held as CSR on the device,
// for operators the library has no stencil kernel for (a matrix a program
// fills itself), and device vector updates for the drop-in classes.
//
//   spmv    y = A x, each row summed from 0 in column order, as
//           dt::matrix::operator* (dt/matrix.h:81-94) sums it;
//   cg      alg::conjugate_gradient (alg/cg.h:32-56) on such a matrix: the
//           reference's updates, with fixed-order tree dot products;
//   smooth  alg::smooth (alg/smooth.h:35-51): one lexicographic Gauss-Seidel
//           sweep, row after row (each row reads the rows updated before
//           it), so one lane walks the rows -- the reference's order;
//   axpy    y += a x.
// Row starts and columns are 64-bit (the reference's std::size_t).
#include <vector>

#include "afm_device.h"

namespace afm {

__global__ void k_csr_spmv(unsigned long rows, const unsigned long* __restrict__ rs,
                           const unsigned long* __restrict__ col, const double* __restrict__ val,
                           const double* __restrict__ x, double* __restrict__ y)
{
	const unsigned long r = blockIdx.x * (unsigned long) blockDim.x + threadIdx.x;
	if (r >= rows) return;
	double v = 0.0;
	for (unsigned long k = rs[r]; k < rs[r + 1]; ++k) v += val[k] * x[col[k]];
	y[r] = v;
}

__global__ void k_csr_smooth(unsigned long rows, const unsigned long* __restrict__ rs,
                             const unsigned long* __restrict__ col, const double* __restrict__ val,
                             double* __restrict__ x, const double* __restrict__ b)
{
	if (blockIdx.x != 0 || threadIdx.x != 0) return;
	for (unsigned long r = 0; r < rows; ++r) {
		double value = 0.0, diagonal = 0.0;
		for (unsigned long k = rs[r]; k < rs[r + 1]; ++k) {
			if (col[k] == r) { diagonal = val[k]; continue; }
			value += val[k] * x[col[k]];
		}
		x[r] = (b[r] - value) / diagonal;
	}
}

__global__ __launch_bounds__(256) void k_vdot(const double* __restrict__ a, const double* __restrict__ b,
                                              unsigned long N, double* partials)
{
	double acc[1] = {0.0};
	for (unsigned long i = blockIdx.x * 256ul + threadIdx.x; i < N; i += (unsigned long) gridDim.x * 256)
		acc[0] += a[i] * b[i];
	reduce_partial<1>(acc, partials);
}

__global__ __launch_bounds__(256) void k_vdot_fin(unsigned nblk, const double* __restrict__ partials,
                                                  double* __restrict__ out)
{
	double tot[1];
	red_total<1>(partials, nblk, tot);
	if (threadIdx.x == 0) *out = tot[0];
}

// r = b - A x with A x from k_csr_spmv; p = r
__global__ void k_vres(const double* __restrict__ b, const double* __restrict__ ax, double* __restrict__ r,
                       double* __restrict__ p, unsigned long N)
{
	const unsigned long i = blockIdx.x * (unsigned long) blockDim.x + threadIdx.x;
	if (i >= N) return;
	const double v = b[i] - ax[i];
	r[i] = v;
	p[i] = v;
}

// alg/cg.h:45-47: x += nu p; z *= nu; r -= z
__global__ void k_vcg_step(double* __restrict__ x, double* __restrict__ r, const double* __restrict__ p,
                           double* __restrict__ z, double nu, unsigned long N)
{
	const unsigned long i = blockIdx.x * (unsigned long) blockDim.x + threadIdx.x;
	if (i >= N) return;
	x[i] += nu * p[i];
	const double zi = z[i] * nu;
	z[i] = zi;
	r[i] -= zi;
}

// alg/cg.h:51-52: p *= mu; p += r
__global__ void k_vcg_dir(double* __restrict__ p, const double* __restrict__ r, double mu, unsigned long N)
{
	const unsigned long i = blockIdx.x * (unsigned long) blockDim.x + threadIdx.x;
	if (i >= N) return;
	p[i] = p[i] * mu;
	p[i] = p[i] + r[i];
}

__global__ void k_vaxpy(double a, const double* __restrict__ x, double* __restrict__ y, unsigned long N)
{
	const unsigned long i = blockIdx.x * (unsigned long) blockDim.x + threadIdx.x;
	if (i < N) y[i] += a * x[i];
}

static inline dim3 gv(unsigned long N) { return dim3((unsigned) ((N + 255) / 256)); }
static inline unsigned vblocks(unsigned long N)
{
	const unsigned long b = (N + 255) / 256;
	return (unsigned) (b < (unsigned long) MAX_RED_BLOCKS ? b : (unsigned long) MAX_RED_BLOCKS);
}

int csr_spmv(Ctx* c, unsigned long rows, const unsigned long* rs, const unsigned long* col, const double* val,
             const double* x, double* y)
{
	if (rows == 0) return 0;
	hipLaunchKernelGGL(k_csr_spmv, gv(rows), dim3(256), 0, c->stream, rows, rs, col, val, x, y);
	AFM_LAUNCH_CHECK();
	return 0;
}

int csr_smooth(Ctx* c, unsigned long rows, const unsigned long* rs, const unsigned long* col, const double* val,
               double* x, const double* b, int sweeps)
{
	for (int s = 0; s < sweeps && rows > 0; ++s)
		hipLaunchKernelGGL(k_csr_smooth, dim3(1), dim3(64), 0, c->stream, rows, rs, col, val, x, b);
	AFM_LAUNCH_CHECK();
	return 0;
}

int vec_axpy(Ctx* c, unsigned long N, double a, const double* x, double* y)
{
	if (N == 0) return 0;
	hipLaunchKernelGGL(k_vaxpy, gv(N), dim3(256), 0, c->stream, a, x, y, N);
	AFM_LAUNCH_CHECK();
	return 0;
}

static int vdot(Ctx* c, const double* a, const double* b, unsigned long N, double* dev, double* host)
{
	hipLaunchKernelGGL(k_vdot, dim3(vblocks(N)), dim3(256), 0, c->stream, a, b, N, c->partials);
	hipLaunchKernelGGL(k_vdot_fin, dim3(1), dim3(256), 0, c->stream, vblocks(N), c->partials, dev);
	AFM_LAUNCH_CHECK();
	AFM_HIP(hipMemcpyAsync(host, dev, 8, hipMemcpyDeviceToHost, c->stream));
	AFM_HIP(hipStreamSynchronize(c->stream));
	return 0;
}

// alg::conjugate_gradient on a device CSR matrix: x is the start and the
// result, stop when sqrt(r.r) <= tol; scratch is allocated for the call
int csr_cg(Ctx* c, unsigned long rows, const unsigned long* rs, const unsigned long* col, const double* val,
           const double* b, double* x, double tol, int max_iters, int* iters, double* final_norm)
{
	if (rows == 0) {
		if (iters) *iters = 0;
		if (final_norm) *final_norm = 0.0;
		return 0;
	}
	std::vector<void*> own;
	auto release = [&](int rc) {
		(void) hipStreamSynchronize(c->stream);
		for (void* p : own) (void) hipFree(p);
		return rc;
	};
	double *r = nullptr, *p = nullptr, *z = nullptr, *scal = nullptr;
	for (double** q : {&r, &p, &z, &scal}) {
		if (hipMalloc((void**) q, (q == &scal ? 1 : (rows ? rows : 1)) * 8) != hipSuccess) {
			set_error("out of device memory (CSR CG)");
			return release(-3);
		}
		own.push_back(*q);
	}
	int rc = 0;
	if ((rc = csr_spmv(c, rows, rs, col, val, x, z))) return release(rc);
	if (rows) hipLaunchKernelGGL(k_vres, gv(rows), dim3(256), 0, c->stream, b, z, r, p, rows);
	double n0 = 0.0, pz = 0.0, n1 = 0.0;
	if ((rc = vdot(c, r, r, rows, scal, &n0))) return release(rc);
	int count = 0;
	while (sqrt(n0) > tol) {
		if (count >= max_iters) {
			if (iters) *iters = count;
			if (final_norm) *final_norm = sqrt(n0);
			set_error("CG did not converge within max_iters");
			return release(-4);
		}
		if ((rc = csr_spmv(c, rows, rs, col, val, p, z))) return release(rc);
		if ((rc = vdot(c, p, z, rows, scal, &pz))) return release(rc);
		const double nu = n0 / pz;
		hipLaunchKernelGGL(k_vcg_step, gv(rows), dim3(256), 0, c->stream, x, r, p, z, nu, rows);
		if ((rc = vdot(c, r, r, rows, scal, &n1))) return release(rc);
		const double mu = n1 / n0;
		n0 = n1;
		hipLaunchKernelGGL(k_vcg_dir, gv(rows), dim3(256), 0, c->stream, p, r, mu, rows);
		++count;
	}
	if (iters) *iters = count;
	if (final_norm) *final_norm = sqrt(n0);
	return release(hipGetLastError() == hipSuccess ? 0 : -3);
}

}  // namespace afm
