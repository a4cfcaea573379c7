// afm_ib3.hip -- the IB transfer on the reference's 3-D MAC grid
// (geometry<L,1,1,1>: n^3 cells, periodic x and z, walls in y):
// ib::grid::interpolate (ib/grid.h:68-84) between markers and one MAC
// component, with the windows of ib/eulerian_grid.h:71-89 and
// ib/mac_grid.h:60-73 (u: comp 0, w: comp 2; n^3 faces) and of the v grid
// mac_grid<y> (ib/mac_grid.h:114-132; comp 1, the n*(n-1)*n interior
// y-faces), or the node grid eulerian_grid itself (comp 3: the window of
// ib/eulerian_grid.h:64-93 unshifted, nodes at (i, j, k)/n, n^3 values).  Taps the grid does not store -- the v window's wall faces and
// the u / w windows' ghost row above the top wall, which the reference
// indexes out of range -- read nothing and receive nothing.
//
// Interpolation: one thread per marker, the 64 taps in the reference's
// order, value += (w * vol) * f.  Spreading, deterministic without float
// atomics: every (marker, tap) pair emits its target index and its term
// (w * ds) * F; a stable radix sort by target keeps, within one target, the
// pairs in marker-then-tap order, and one thread per target sums its run in
// that order from 0.0 -- the accumulation order of the marker loop.
#include <hipcub/hipcub.hpp>

#include "afm_delta.h"
#include "afm_device.h"

namespace afm {

constexpr unsigned IB3_SKIP = 0xffffffffu;

// The tap t (0..63) of the window of marker c for component comp: its
// target index (IB3_SKIP when the tap is dropped) and weight
// delta(p0 - c0) * delta(p1 - c1) * delta(p2 - c2).
struct Win3 {
	double fx, fy, fz;
};

__device__ inline Win3 win3_base(int comp, const double c[3], unsigned n)
{
	double q[3] = {c[0], c[1], c[2]};
	if (comp == 0 || comp == 2) {
		q[(comp + 1) % 3] -= 0.5 / n;
		q[(comp + 2) % 3] -= 0.5 / n;
	}
	q[0] = fmod(fmod(q[0], 1.0) + 1.0, 1.0);
	q[2] = fmod(fmod(q[2], 1.0) + 1.0, 1.0);
	Win3 b;
	if (comp == 1) { b.fx = floor(q[0] * n - 0.5); b.fy = floor(q[1] * n - 1.0); b.fz = floor(q[2] * n - 0.5); }
	else { b.fx = floor(q[0] * n); b.fy = floor(q[1] * n); b.fz = floor(q[2] * n); }
	return b;
}

__device__ inline unsigned win3_tap(int comp, const Win3& b, const double c[3], unsigned n, int t, double& w)
{
	const int dix = t % 4, diy = (t / 4) % 4, diz = (t / 16) % 4;
	const long long nx = n, ny = n, nz = n;
	double p[3];
	p[0] = (b.fx + dix - 1.0) / n;
	p[1] = comp == 1 ? (b.fy + diy) / n : (b.fy + diy - 1.0) / n;
	p[2] = (b.fz + diz - 1.0) / n;
	if (p[1] < 0 || p[1] > 1.0) return IB3_SKIP;
	const long long iy = (long long) b.fy + diy;
	if (comp == 1 ? (iy == 0 || iy == ny) : iy > ny) return IB3_SKIP;
	// fx, fz >= -1 after the wrap, so the sums below are non-negative
	const long long ix = ((long long) b.fx + dix + nx - 1) % nx, iz = ((long long) b.fz + diz + nz - 1) % nz;
	long long idx;
	if (comp == 1) {
		p[0] = p[0] + 0.5 / n;
		p[2] = p[2] + 0.5 / n;
		idx = ix + nx * (iy - 1 + (ny - 1) * iz);
	} else {
		idx = ix + nx * (iy - 1 + ny * iz);
		if (comp != 3) {
			p[(comp + 1) % 3] += 0.5 / n;
			p[(comp + 2) % 3] += 0.5 / n;
		}
	}
	w = delta(p[0] - c[0], n) * delta(p[1] - c[1], n) * delta(p[2] - c[2], n);
	return (unsigned) idx;
}

__global__ void k_interp3(int comp, const double* __restrict__ f, const double* __restrict__ X, long m, unsigned n,
                          double* __restrict__ U)
{
	const long k = blockIdx.x * (long) blockDim.x + threadIdx.x;
	if (k >= m) return;
	const double c[3] = {X[3 * k], X[3 * k + 1], X[3 * k + 2]};
	const Win3 b = win3_base(comp, c, n);
	const double vol = 1. / (n * n * n);
	double value = 0.0;
	for (int t = 0; t < 64; ++t) {
		double w;
		const unsigned idx = win3_tap(comp, b, c, n, t, w);
		if (idx == IB3_SKIP) continue;
		value += w * vol * f[idx];
	}
	U[k] = value;
}

__global__ void k_spread3_terms(int comp, const double* __restrict__ F, const double* __restrict__ ds,
                                const double* __restrict__ X, long m, unsigned n, unsigned* __restrict__ keys,
                                unsigned* __restrict__ ids, double* __restrict__ terms)
{
	const long k = blockIdx.x * (long) blockDim.x + threadIdx.x;
	if (k >= m) return;
	const double c[3] = {X[3 * k], X[3 * k + 1], X[3 * k + 2]};
	const Win3 b = win3_base(comp, c, n);
	const double g = ds[k], Fk = F[k];
	for (int t = 0; t < 64; ++t) {
		double w = 0.0;
		const unsigned idx = win3_tap(comp, b, c, n, t, w);
		const size_t j = (size_t) k * 64 + t;
		keys[j] = idx;
		ids[j] = (unsigned) j;
		terms[j] = idx == IB3_SKIP ? 0.0 : w * g * Fk;
	}
}

// one thread per sorted pair; the first pair of each target sums its run
__global__ void k_spread3_sum(const unsigned* __restrict__ keys, const unsigned* __restrict__ ids,
                              const double* __restrict__ terms, size_t P, double* __restrict__ f)
{
	const size_t j = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (j >= P) return;
	const unsigned key = keys[j];
	if (key == IB3_SKIP || (j > 0 && keys[j - 1] == key)) return;
	double s = 0.0;
	for (size_t q = j; q < P && keys[q] == key; ++q) s += terms[ids[q]];
	f[key] = s;
}

int ib3_interp(Ctx* c, unsigned L, int comp, const double* field, const double* X, long m, double* U)
{
	if (m <= 0) return 0;
	hipLaunchKernelGGL(k_interp3, dim3((unsigned) ((m + 255) / 256)), dim3(256), 0, c->stream, comp, field, X, m,
	                   1u << L, U);
	AFM_LAUNCH_CHECK();
	return 0;
}

int ib3_spread(Ctx* c, unsigned L, int comp, const double* F, const double* ds, const double* X, long m, double* f)
{
	const unsigned n = 1u << L;
	const size_t N = (size_t) n * (comp == 1 ? n - 1 : n) * n;
	AFM_HIP(hipMemsetAsync(f, 0, N * 8, c->stream));
	if (m <= 0) return 0;
	const size_t P = (size_t) m * 64;
	std::vector<void*> own;
	auto release = [&](int rc) {
		(void) hipStreamSynchronize(c->stream);
		for (void* p : own) (void) hipFree(p);
		return rc;
	};
	unsigned *keys = nullptr, *keys_s = nullptr, *ids = nullptr, *ids_s = nullptr;
	double* terms = nullptr;
	for (void** p : {(void**) &keys, (void**) &keys_s, (void**) &ids, (void**) &ids_s}) {
		if (hipMalloc(p, P * 4) != hipSuccess) { set_error("out of device memory (3-D spreading)"); return release(-3); }
		own.push_back(*p);
	}
	if (hipMalloc((void**) &terms, P * 8) != hipSuccess) { set_error("out of device memory (3-D spreading)"); return release(-3); }
	own.push_back(terms);
	hipLaunchKernelGGL(k_spread3_terms, dim3((unsigned) ((m + 255) / 256)), dim3(256), 0, c->stream, comp, F, ds, X,
	                   m, n, keys, ids, terms);
	if (hipGetLastError() != hipSuccess) { set_error("3-D spreading launch failed"); return release(-2); }
	size_t tmp_bytes = 0;
	if (hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys, keys_s, ids, ids_s, (int) P, 0, 32, c->stream) !=
	    hipSuccess) {
		set_error("radix sort sizing failed");
		return release(-2);
	}
	void* tmp = nullptr;
	if (hipMalloc(&tmp, tmp_bytes ? tmp_bytes : 1) != hipSuccess) { set_error("out of device memory (3-D spreading)"); return release(-3); }
	own.push_back(tmp);
	if (hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys, keys_s, ids, ids_s, (int) P, 0, 32, c->stream) !=
	    hipSuccess) {
		set_error("radix sort failed");
		return release(-2);
	}
	hipLaunchKernelGGL(k_spread3_sum, dim3((unsigned) ((P + 255) / 256)), dim3(256), 0, c->stream, keys_s, ids_s,
	                   terms, P, f);
	if (hipGetLastError() != hipSuccess) { set_error("3-D spreading launch failed"); return release(-2); }
	return release(0);
}

}  // namespace afm
