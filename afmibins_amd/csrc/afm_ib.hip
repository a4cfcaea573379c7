// afm_ib.hip -- Lagrangian <-> Eulerian transfer on gfx950.
//
// Interpolation (ib/grid.h:68-84, marker targets): one lane per marker, the
// 4 + 4 delta factors evaluated once, 16 taps gathered in the reference's
// order.  Same arithmetic as the reference, so the result is bit-exact
// against the CPU restatement.
//
// Spreading (ib/grid.h:68-84, grid targets): the reference gathers, for
// every grid point, the markers found by a kd-tree box query.  Here markers
// are binned by the 16x16 tile of their window's base cell (integer
// counting sort), and one workgroup per tile collects the candidates of the
// 3x3 neighbouring bins, sorts them by marker index in LDS and lets each lane
// (one grid cell) sum its terms (w * ds) * F in ascending marker order.
// No float atomics: the result is identical run to run and equals a
// sequential scatter over markers in index order.
#include "afm_device.h"

namespace afm {

constexpr int TS = 16;            // spread tile edge (cells)
constexpr int KMAX = 4096;        // candidate capacity per tile

__global__ void k_delta(const double* __restrict__ r, long count, unsigned n, double* __restrict__ out)
{
	long i = blockIdx.x * (long) blockDim.x + threadIdx.x;
	if (i < count) out[i] = delta(r[i], n);
}

__global__ void k_interp(int comp, const double* __restrict__ f, const double* __restrict__ X, long m,
                         unsigned n, unsigned W, unsigned D, unsigned nx, unsigned ny, double* __restrict__ U)
{
	long k = blockIdx.x * (long) blockDim.x + threadIdx.x;
	if (k >= m) return;
	int ix0, iy0;
	double dx[4], dy[4];
	ib_window(comp, X[2 * k], X[2 * k + 1], n, W, D, nx, ny, ix0, iy0, dx, dy);
	const double vol = 1. / (n * n);
	U[k] = ib_gather(f, nx, ny, ix0, iy0, dx, dy, vol);
}

// X_h = X + half * (J_u(X) u, J_v(X) v)
__global__ void k_ib_move_half(const double* __restrict__ u, const double* __restrict__ v, const double* __restrict__ X,
                               long m, unsigned n, unsigned W, unsigned D, unsigned nx, unsigned ny, double half,
                               double* __restrict__ Xh)
{
	long k = blockIdx.x * (long) blockDim.x + threadIdx.x;
	if (k >= m) return;
	const double cx = X[2 * k], cy = X[2 * k + 1];
	const double vol = 1. / (n * n);
	int ix0, iy0;
	double dx[4], dy[4];
	ib_window(0, cx, cy, n, W, D, nx, ny, ix0, iy0, dx, dy);
	double Ux = ib_gather(u, nx, ny, ix0, iy0, dx, dy, vol);
	ib_window(1, cx, cy, n, W, D, nx, ny, ix0, iy0, dx, dy);
	double Uy = ib_gather(v, nx, ny, ix0, iy0, dx, dy, vol);
	Xh[2 * k] = cx + half * Ux;
	Xh[2 * k + 1] = cy + half * Uy;
}

// X = X + dt * (0.5 * (J(X_h) u_old + J(X_h) u))
__global__ void k_ib_move_full(const double* __restrict__ uo, const double* __restrict__ vo,
                               const double* __restrict__ u, const double* __restrict__ v,
                               const double* __restrict__ Xh, long m, unsigned n, unsigned W, unsigned D,
                               unsigned nx, unsigned ny, double dt, double* __restrict__ X)
{
	long k = blockIdx.x * (long) blockDim.x + threadIdx.x;
	if (k >= m) return;
	const double cx = Xh[2 * k], cy = Xh[2 * k + 1];
	const double vol = 1. / (n * n);
	int ix0, iy0;
	double dx[4], dy[4];
	ib_window(0, cx, cy, n, W, D, nx, ny, ix0, iy0, dx, dy);
	double U0x = ib_gather(uo, nx, ny, ix0, iy0, dx, dy, vol);
	double U1x = ib_gather(u, nx, ny, ix0, iy0, dx, dy, vol);
	ib_window(1, cx, cy, n, W, D, nx, ny, ix0, iy0, dx, dy);
	double U0y = ib_gather(vo, nx, ny, ix0, iy0, dx, dy, vol);
	double U1y = ib_gather(v, nx, ny, ix0, iy0, dx, dy, vol);
	X[2 * k] = X[2 * k] + dt * (0.5 * (U0x + U1x));
	X[2 * k + 1] = X[2 * k + 1] + dt * (0.5 * (U0y + U1y));
}

// Per (marker, component): window, source value and bin.  With R/k given,
// the source is the tether force F = -k (X - R); otherwise F[comp][k].
__global__ void k_spread_prep(int comp_mask, const double* __restrict__ Fx, const double* __restrict__ Fy,
                              const double* __restrict__ ds, const double* __restrict__ X,
                              const double* __restrict__ R, const double* __restrict__ kk, long m,
                              unsigned n, unsigned W, unsigned D, unsigned nx, unsigned ny, int ntx, int nty,
                              MarkerRec* __restrict__ rec, int* __restrict__ counts)
{
	long k = blockIdx.x * (long) blockDim.x + threadIdx.x;
	int comp = blockIdx.y;
	if (k >= m || !((comp_mask >> comp) & 1)) return;
	const double cx = X[2 * k], cy = X[2 * k + 1];
	MarkerRec r;
	ib_window(comp, cx, cy, n, W, D, nx, ny, r.ix0, r.iy0, r.dx, r.dy);
	double F;
	if (R) F = comp == 0 ? -kk[k] * (cx - R[2 * k]) : -kk[k] * (cy - R[2 * k + 1]);
	else F = comp == 0 ? Fx[k] : Fy[k];
	r.ds = ds[k];
	r.F = F;
	int hx = r.ix0 + 1; if (hx >= (int) nx) hx -= nx;
	int hy = r.iy0 + 1; if (hy >= (int) ny) hy -= ny;
	r.tile = hx / TS + ntx * (hy / TS);
	r.pad = 0;
	rec[comp * m + k] = r;
	atomicAdd(&counts[comp * ntx * nty + r.tile], 1);
}

// Exclusive scan of the bin counts of both components (one block of 1024):
// per-thread serial chunk sums, a Hillis-Steele scan of the chunk sums, then
// the per-thread chunk scan.
__global__ __launch_bounds__(1024) void k_spread_scan(const int* __restrict__ counts, int* __restrict__ offs,
                                                      int* __restrict__ cursor, int nbins)
{
	__shared__ int part[1024];
	const int tid = threadIdx.x, nt = blockDim.x;
	const int per = (nbins + nt - 1) / nt;
	const int lo = tid * per, hi = min(nbins, lo + per);
	int s = 0;
	for (int i = lo; i < hi; ++i) s += counts[i];
	part[tid] = s;
	__syncthreads();
	for (int d = 1; d < nt; d <<= 1) {
		int v = tid >= d ? part[tid - d] : 0;
		__syncthreads();
		part[tid] += v;
		__syncthreads();
	}
	int acc = part[tid] - s;                  // exclusive prefix of this chunk
	for (int i = lo; i < hi; ++i) { offs[i] = acc; cursor[i] = 0; acc += counts[i]; }
	if (tid == nt - 1) offs[nbins] = part[nt - 1];
}

__global__ void k_spread_fill(int comp_mask, const MarkerRec* __restrict__ rec, long m, int nb1,
                              const int* __restrict__ offs, int* __restrict__ cursor, int* __restrict__ list)
{
	long k = blockIdx.x * (long) blockDim.x + threadIdx.x;
	int comp = blockIdx.y;
	if (k >= m || !((comp_mask >> comp) & 1)) return;
	int b = comp * nb1 + rec[comp * m + k].tile;
	int slot = atomicAdd(&cursor[b], 1);
	list[offs[b] + slot] = (int) k;
}

// One workgroup (TS x TS lanes) per (tile, component).  The distinct
// neighbouring bins and their sizes are computed redundantly by every lane,
// so empty tiles (the common case) write their zeros without a barrier.
__global__ __launch_bounds__(256) void k_spread_tile(int comp_mask, const MarkerRec* __restrict__ rec, long m,
                                                     const int* __restrict__ offs, const int* __restrict__ list,
                                                     unsigned nx, unsigned ny, int ntx, int nty,
                                                     double* __restrict__ fx, double* __restrict__ fy,
                                                     int* __restrict__ err, int accumulate)
{
	__shared__ int cand[KMAX];
	const int comp = blockIdx.y;
	if (!((comp_mask >> comp) & 1)) return;
	const int nb1 = ntx * nty;
	const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
	const int tx = tile % ntx, ty = tile / ntx;
	const int tid = threadIdx.x;
	const int gx = tx * TS + tid % TS, gy = ty * TS + tid / TS;
	double* f = comp == 0 ? fx : fy;
	// distinct neighbouring bins (periodic; fewer than 3 tiles per axis
	// would otherwise visit a bin twice)
	int bxs[3], bys[3], nbx = 0, nby = 0;
	for (int d = -1; d <= 1; ++d) {
		int b = (tx + d + ntx) % ntx;
		bool dup = false;
		for (int j = 0; j < nbx; ++j) dup |= bxs[j] == b;
		if (!dup) bxs[nbx++] = b;
		b = (ty + d + nty) % nty;
		dup = false;
		for (int j = 0; j < nby; ++j) dup |= bys[j] == b;
		if (!dup) bys[nby++] = b;
	}
	int o0[9], cnt[9], nb = 0, K = 0;
	for (int j = 0; j < nby; ++j)
		for (int i = 0; i < nbx; ++i) {
			const int b = comp * nb1 + bxs[i] + ntx * bys[j];
			o0[nb] = offs[b];
			cnt[nb] = offs[b + 1] - o0[nb];
			K += cnt[nb];
			++nb;
		}
	if (K == 0) {
		if (!accumulate && gx < (int) nx && gy < (int) ny) f[gx + (size_t) nx * gy] = 0.0;
		return;
	}
	if (K > KMAX) {
		if (tid == 0) atomicOr(err, 1);
		K = KMAX;
	}
	for (int t = tid; t < K; t += blockDim.x) {
		int q = t, j = 0;
		while (q >= cnt[j]) { q -= cnt[j]; ++j; }
		cand[t] = list[o0[j] + q];
	}
	// bitonic sort of the candidate marker indices (ascending)
	int P = 1;
	while (P < K) P <<= 1;
	for (int t = K + tid; t < P; t += blockDim.x) cand[t] = 0x7fffffff;
	__syncthreads();
	for (int size = 2; size <= P; size <<= 1)
		for (int stride = size >> 1; stride > 0; stride >>= 1) {
			for (int t = tid; t < P / 2; t += blockDim.x) {
				int lo = 2 * stride * (t / stride) + (t % stride);
				int hi = lo + stride;
				bool up = ((lo & size) == 0);
				int a = cand[lo], b = cand[hi];
				if ((a > b) == up) { cand[lo] = b; cand[hi] = a; }
			}
			__syncthreads();
		}
	if (gx >= (int) nx || gy >= (int) ny) return;
	double acc = 0.0;
	const MarkerRec* rc = rec + comp * m;
	for (int j = 0; j < K; ++j) {
		const MarkerRec& r = rc[cand[j]];
		int a = gx - r.ix0; if (a < 0) a += nx;
		int b = gy - r.iy0; if (b < 0) b += ny;
		if (a < 4 && b < 4) {
			double w = r.dx[a] * r.dy[b];
			acc += w * r.ds * r.F;
		}
	}
	const size_t o = gx + (size_t) nx * gy;
	f[o] = accumulate ? f[o] + acc : acc;
}

int launch_delta(Ctx* c, const double* r, long count, unsigned n, double* out)
{
	if (count <= 0) return 0;
	hipLaunchKernelGGL(k_delta, dim3((count + 255) / 256), dim3(256), 0, c->stream, r, count, n, out);
	AFM_LAUNCH_CHECK();
	return 0;
}

int launch_interp(Ctx* c, int comp, const double* field, const double* X, long m, double* U)
{
	if (m <= 0) return 0;
	hipLaunchKernelGGL(k_interp, dim3((m + 255) / 256), dim3(256), 0, c->stream, comp, field, X, m, c->n, c->W,
	                   c->D, c->nx, c->ny, U);
	AFM_LAUNCH_CHECK();
	return 0;
}

int launch_ib_move_half(Ctx* c, const double* u, const double* v, const double* X, long m, double half, double* Xh)
{
	if (m <= 0) return 0;
	hipLaunchKernelGGL(k_ib_move_half, dim3((m + 255) / 256), dim3(256), 0, c->stream, u, v, X, m, c->n, c->W, c->D,
	                   c->nx, c->ny, half, Xh);
	AFM_LAUNCH_CHECK();
	return 0;
}

int launch_ib_move_full(Ctx* c, const double* uo, const double* vo, const double* u, const double* v,
                        const double* Xh, long m, double dt, double* X)
{
	if (m <= 0) return 0;
	hipLaunchKernelGGL(k_ib_move_full, dim3((m + 255) / 256), dim3(256), 0, c->stream, uo, vo, u, v, Xh, m, c->n,
	                   c->W, c->D, c->nx, c->ny, dt, X);
	AFM_LAUNCH_CHECK();
	return 0;
}

int ensure_markers(Ctx* c, long m)
{
	if (m <= c->mcap) return 0;
	long cap = 1024;
	while (cap < m) cap *= 2;
	if (c->rec) { (void) hipFree(c->rec); c->rec = nullptr; }
	if (c->bin_list) { (void) hipFree(c->bin_list); c->bin_list = nullptr; }
	AFM_HIP(hipMalloc(&c->rec, sizeof(MarkerRec) * 2 * cap));
	AFM_HIP(hipMalloc(&c->bin_list, sizeof(int) * 2 * cap));
	c->mcap = cap;
	return 0;
}

int launch_spread(Ctx* c, int comp_mask, const double* Fx, const double* Fy, const double* ds, const double* X,
                  const double* R, const double* k, long m, double* fx, double* fy, int accumulate)
{
	const int nb1 = c->ntx * c->nty, nb = 2 * nb1;
	int* counts = c->bins;
	int* offs = c->bins + nb;          // nb + 1 entries
	int* cursor = c->bins + 2 * nb + 1;
	if (m > 0) {
		AFM_TRY(ensure_markers(c, m));
		AFM_HIP(hipMemsetAsync(counts, 0, sizeof(int) * nb, c->stream));
		dim3 g((m + 255) / 256, 2);
		hipLaunchKernelGGL(k_spread_prep, g, dim3(256), 0, c->stream, comp_mask, Fx, Fy, ds, X, R, k, m, c->n, c->W,
		                   c->D, c->nx, c->ny, c->ntx, c->nty, c->rec, counts);
		AFM_LAUNCH_CHECK();
		hipLaunchKernelGGL(k_spread_scan, dim3(1), dim3(1024), 0, c->stream, counts, offs, cursor, nb);
		AFM_LAUNCH_CHECK();
		hipLaunchKernelGGL(k_spread_fill, g, dim3(256), 0, c->stream, comp_mask, c->rec, m, nb1, offs, cursor,
		                   c->bin_list);
		AFM_LAUNCH_CHECK();
	} else {
		AFM_HIP(hipMemsetAsync(offs, 0, sizeof(int) * (nb + 1), c->stream));
	}
	hipLaunchKernelGGL(k_spread_tile, dim3(nb1, 2), dim3(256), 0, c->stream, comp_mask, c->rec, m, offs,
	                   c->bin_list, c->nx, c->ny, c->ntx, c->nty, fx, fy, &c->ctl[3].err, accumulate);
	AFM_LAUNCH_CHECK();
	return 0;
}

}  // namespace afm
