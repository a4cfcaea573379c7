// afm_device.h -- device helpers: deterministic reductions, the periodic
// 5-point stencil in the reference's CSR summation order, and the 4x4 IB
// window.  Included by the .hip translation units only.
#pragma once
#include "afm_internal.h"

namespace afm {

// ------------------------------------------------------------ reduce ----
// Fixed-tree block sum: wave shuffle tree, then wave partials summed in
// order by thread 0.  The result is valid in thread 0 only.
template<int NV>
__device__ inline void block_sum(double (&v)[NV])
{
	__shared__ double wsum[NV][16];
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int nw = (blockDim.x * blockDim.y + 63) >> 6;
	const int tid = threadIdx.x + threadIdx.y * blockDim.x;
	const int lane2 = tid & 63, wave2 = tid >> 6;
	(void) lane; (void) wave;
#pragma unroll
	for (int j = 0; j < NV; ++j) {
		double s = v[j];
		for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
		if (lane2 == 0) wsum[j][wave2] = s;
	}
	__syncthreads();
	if (tid == 0) {
#pragma unroll
		for (int j = 0; j < NV; ++j) {
			double s = 0.0;
			for (int w = 0; w < nw; ++w) s += wsum[j][w];
			v[j] = s;
		}
	}
	__syncthreads();
}

// Two-phase deterministic reduction.  Phase 1 (in the producing kernel):
// each block writes its NV partial sums to partials[j * MAX_RED_BLOCKS +
// blockIdx.x] -- no fence, no atomics (a device-scope fence per block writes
// back the XCD's L2 and costs more than the kernel's own traffic).  Phase 2
// (a one-block kernel launched next on the same stream): red_total sums the
// partials in block order, so the result does not depend on scheduling.
template<int NV>
__device__ inline void reduce_partial(double (&v)[NV], double* partials)
{
	block_sum<NV>(v);
	if (threadIdx.x + threadIdx.y * blockDim.x == 0) {
#pragma unroll
		for (int j = 0; j < NV; ++j) partials[(size_t) j * MAX_RED_BLOCKS + blockIdx.x] = v[j];
	}
}

// p[t] + p[t + step] + p[t + 2 step] + ... over the indices below nblk, added
// in that order from 0.  The loads of four terms are issued before their
// adds (index clamped, term skipped): one memory round trip per four terms
// instead of one per term, the same bits.
__device__ inline double strided_sum(const double* __restrict__ p, unsigned t, unsigned step, unsigned nblk)
{
	double s = 0.0;
	for (unsigned b = t; b < nblk; b += 4 * step) {
		double v[4];
#pragma unroll
		for (int k = 0; k < 4; ++k) {
			const unsigned q = b + k * step;
			v[k] = p[q < nblk ? q : nblk - 1];
		}
#pragma unroll
		for (int k = 0; k < 4; ++k)
			if (b + k * step < nblk) s += v[k];
	}
	return s;
}

// Phase 2: totals in thread 0 (call from every thread of a one-block kernel).
template<int NV>
__device__ inline void red_total(const double* partials, unsigned nblk, double (&tot)[NV])
{
	const int tid = threadIdx.x;
#pragma unroll
	for (int j = 0; j < NV; ++j) tot[j] = strided_sum(partials + (size_t) j * MAX_RED_BLOCKS, tid, blockDim.x, nblk);
	block_sum<NV>(tot);
}

// ------------------------------------------------------------ stencil ---
// Periodic 5-point stencil (nx, ny >= 3), terms in ascending column order as
// stored by mg/laplacian.h:97-150 (and ns/helmholtz.h:119-172) with ny == 1:
//   N(wrap), S, E(wrap), W, [diag], E, W(wrap), N, S(wrap)
// sum starts from 0 and adds a*x per stored entry (dt/matrix.h:87-91).
template<typename T, bool DIAG>
__device__ inline T stencil5(T off, T dg, T c, T e, T w, T nn, T s, bool we, bool ww, bool wn, bool ws)
{
	T v = T(0);
	if (wn) v += off * nn;
	if (!ws) v += off * s;
	if (we) v += off * e;
	if (!ww) v += off * w;
	if (DIAG) v += dg * c;
	if (!we) v += off * e;
	if (ww) v += off * w;
	if (!wn) v += off * nn;
	if (ws) v += off * s;
	return v;
}

// stencil5 for power-of-two coefficients (the multigrid laplacian's n^2 and
// -4 n^2): off * x is then exact, so fma(off, x, v) rounds exactly as
// v + off * x does -- the same bits in one instruction instead of two
__device__ inline float fma_p2(float a, float x, float v) { return __builtin_fmaf(a, x, v); }
__device__ inline double fma_p2(double a, double x, double v) { return __builtin_fma(a, x, v); }

template<typename T, bool DIAG>
__device__ inline T stencil5p2(T off, T dg, T c, T e, T w, T nn, T s, bool we, bool ww, bool wn, bool ws)
{
	T v = T(0);
	if (wn) v = fma_p2(off, nn, v);
	if (!ws) v = fma_p2(off, s, v);
	if (we) v = fma_p2(off, e, v);
	if (!ww) v = fma_p2(off, w, v);
	if (DIAG) v = fma_p2(dg, c, v);
	if (!we) v = fma_p2(off, e, v);
	if (ww) v = fma_p2(off, w, v);
	if (!wn) v = fma_p2(off, nn, v);
	if (ws) v = fma_p2(off, s, v);
	return v;
}

// General row of the 2-D laplacian (any nx, ny >= 1), mg/laplacian.h:50-150
// with ny(reference) == 1 and the reference's nx/nz == 1, 2 folding.
struct LapRow {
	int col[5];
	double a[5];
	int cnt;
	double diag;
};

__device__ inline void lap_row(int x, int y, int nx, int ny, double off, LapRow& r)
{
	double diagonal = -6.0 * off, left = off, right = off, front = off, back = off;
	if (nx == 1) { diagonal += left + right; left = right = 0; }
	else if (nx == 2) { if (x == 0) { right += left; left = 0; } else { left += right; right = 0; } }
	diagonal += off;   // the reference's y axis folded (ny == 1: bottom and top)
	diagonal += off;
	if (ny == 1) { diagonal += back + front; back = front = 0; }
	else if (ny == 2) { if (y == 0) { front += back; back = 0; } else { back += front; front = 0; } }
	r.cnt = 0;
	r.diag = diagonal;
	const int i = x + nx * y;
	auto put = [&](double v, int c) { r.a[r.cnt] = v; r.col[r.cnt] = c; r.cnt++; };
	if (front != 0 && y == ny - 1) put(front, x);
	if (back != 0 && y > 0) put(back, i - nx);
	if (right != 0 && x == nx - 1) put(right, nx * y);
	if (left != 0 && x > 0) put(left, i - 1);
	if (diagonal != 0) put(diagonal, i);
	if (right != 0 && x < nx - 1) put(right, i + 1);
	if (left != 0 && x == 0) put(left, nx - 1 + nx * y);
	if (front != 0 && y < ny - 1) put(front, i + nx);
	if (back != 0 && y == 0) put(back, x + nx * (ny - 1));
}

// fmod(x, p) for p a power of two (the unit-square width and depth):
// x - p * trunc(x / p).  x / p and p * t are exact scalings and, when t != 0,
// the difference is exact (Sterbenz), so the value is fmod's up to the sign
// of a zero result -- which the window's fmod(fmod(q, W) + W, W) then adds to
// W -- without fmod's remainder loop.
__device__ inline double fmod_pow2(double x, unsigned p)
{
	const double pd = (double) p;
	return x - pd * trunc(x * (1.0 / pd));
}

// ------------------------------------------------------------- IB window --
// dx[d] = delta(rx[d]), dy[d] = delta(ry[d]): the fast paths of all eight
// factors first (no branch between them, so their chains interleave), then
// the exact path for the factors whose fast path could not decide (a wave
// takes that branch when any of its lanes needs it).
__device__ inline void window_deltas(const double rx[4], const double ry[4], unsigned n, double dx[4],
                                     double dy[4])
{
	double r4[4], f4[4];
#pragma unroll
	for (int d = 0; d < 4; ++d) r4[d] = rx[d];
	unsigned bad = delta_fast_n<4>(r4, n, f4);
#pragma unroll
	for (int d = 0; d < 4; ++d) { dx[d] = f4[d]; r4[d] = ry[d]; }
	bad |= delta_fast_n<4>(r4, n, f4) << 4;
#pragma unroll
	for (int d = 0; d < 4; ++d) dy[d] = f4[d];
	// one inlined copy of the exact path, in a loop over the factors that
	// need it (register selects, not indexed arrays: nothing goes to scratch)
	while (bad) {
		const int i = __builtin_ctz(bad);
		bad &= bad - 1;
		double r = rx[0];
#pragma unroll
		for (int d = 1; d < 8; ++d) r = i == d ? (d < 4 ? rx[d] : ry[d - 4]) : r;
		const double v = delta_exact_inl(r, n);
#pragma unroll
		for (int d = 0; d < 8; ++d) {
			if (d < 4) dx[d] = i == d ? v : dx[d];
			else dy[d - 4] = i == d ? v : dy[d - 4];
		}
	}
}

// The 4x4 window of one MAC component around marker (cx, cy):
// ib/eulerian_grid.h:64-93 through ib/mac_grid.h:49-73, restricted to the
// periodic (x, z) plane.  comp 0 (u, x-faces) is shifted by -h/2 in y before
// the search, comp 1 (v, y-faces) in x.  Tap (a, b) is the cell
// ((ix0 + a) % nx, (iy0 + b) % ny) with weight dx[a] * dy[b].
__device__ inline void ib_window(int comp, double cx, double cy, unsigned n, unsigned W, unsigned D,
                                 unsigned nx, unsigned ny, int& ix0, int& iy0, double dx[4], double dy[4])
{
	const double h = 1.0 / n;      // n is a power of two: x / n == x * h exactly
	double q0 = cx, q1 = cy;
	if (comp == 0) q1 -= 0.5 * h; else q0 -= 0.5 * h;
	q0 = fmod_pow2(fmod_pow2(q0, W) + W, W);
	q1 = fmod_pow2(fmod_pow2(q1, D) + D, D);
	double fx = floor(q0 * n), fy = floor(q1 * n);
	// finite inputs give fx in [0, nx]; keep non-finite ones in bounds
	if (!(fx >= 0.0 && fx <= (double) nx)) fx = 0.0;
	if (!(fy >= 0.0 && fy <= (double) ny)) fy = 0.0;
	ix0 = (int) (((unsigned) fx + nx - 1) & (nx - 1));
	iy0 = (int) (((unsigned) fy + ny - 1) & (ny - 1));
	double rx[4], ry[4];
#pragma unroll
	for (int d = 0; d < 4; ++d) {
		double x = (fx + d - 1.0) * h, y = (fy + d - 1.0) * h;
		if (comp == 0) y += 0.5 * h; else x += 0.5 * h;
		rx[d] = x - cx;
		ry[d] = y - cy;
	}
	window_deltas(rx, ry, n, dx, dy);
}

// Per-launch constants of the 2-D window, formed once on the host by the
// same IEEE operations every lane would otherwise repeat (divisions and
// conversions of uniform values cost a lane several instructions each).
struct WinGeom {
	double h, half_h, nd, Wd, invW, Dd, invD, nxd, nyd, vol;
	unsigned n, nx, ny;
};

inline WinGeom win_geom(unsigned n, unsigned W, unsigned D, unsigned nx, unsigned ny)
{
	WinGeom g;
	g.h = 1.0 / n;
	g.half_h = 0.5 * g.h;
	g.nd = (double) n;
	g.Wd = (double) W;
	g.invW = 1.0 / g.Wd;
	g.Dd = (double) D;
	g.invD = 1.0 / g.Dd;
	g.nxd = (double) nx;
	g.nyd = (double) ny;
	g.vol = 1. / (n * n);
	g.n = n;
	g.nx = nx;
	g.ny = ny;
	return g;
}

// fmod_pow2 with the period and its reciprocal given
__device__ inline double fmod_pow2g(double x, double pd, double inv) { return x - pd * trunc(x * inv); }

// f[d] = delta(r[d]), d = 0..3: as window_deltas, for one axis
__device__ inline void axis_deltas(const double r[4], unsigned n, double f[4])
{
	double rr4[4], f4[4];
#pragma unroll
	for (int d = 0; d < 4; ++d) rr4[d] = r[d];
	unsigned bad = delta_fast_n<4>(rr4, n, f4);
#pragma unroll
	for (int d = 0; d < 4; ++d) f[d] = f4[d];
	while (bad) {
		const int i = __builtin_ctz(bad);
		bad &= bad - 1;
		double rr = r[0];
#pragma unroll
		for (int d = 1; d < 4; ++d) rr = i == d ? r[d] : rr;
		const double v = delta_exact_inl(rr, n);
#pragma unroll
		for (int d = 0; d < 4; ++d) f[d] = i == d ? v : f[d];
	}
}

// One axis of ib_window with the launch's constants (the same operations):
// p = 0 gives dx, p = 1 dy; both give the base cell.  A lane pair (one lane
// per axis) forms a marker's window with half the work per lane.
__device__ inline void ib_window_axis(int comp, double cx, double cy, const WinGeom& g, int p, int& ix0, int& iy0,
                                      double f[4])
{
	double q0 = cx, q1 = cy;
	if (comp == 0) q1 -= g.half_h; else q0 -= g.half_h;
	q0 = fmod_pow2g(fmod_pow2g(q0, g.Wd, g.invW) + g.Wd, g.Wd, g.invW);
	q1 = fmod_pow2g(fmod_pow2g(q1, g.Dd, g.invD) + g.Dd, g.Dd, g.invD);
	double fx = floor(q0 * g.nd), fy = floor(q1 * g.nd);
	if (!(fx >= 0.0 && fx <= g.nxd)) fx = 0.0;
	if (!(fy >= 0.0 && fy <= g.nyd)) fy = 0.0;
	ix0 = (int) (((unsigned) fx + g.nx - 1) & (g.nx - 1));
	iy0 = (int) (((unsigned) fy + g.ny - 1) & (g.ny - 1));
	const double base = p ? fy : fx, c = p ? cy : cx;
	// comp 0 shifts the y taps by h/2, comp 1 the x taps
	const bool shift = (p == 0) == (comp != 0);
	double r[4];
#pragma unroll
	for (int d = 0; d < 4; ++d) {
		double x = (base + d - 1.0) * g.h;
		if (shift) x += g.half_h;
		r[d] = x - c;
	}
	axis_deltas(r, g.n, f);
}

// ib_window with the launch's constants (the same operations)
__device__ inline void ib_window_g(int comp, double cx, double cy, const WinGeom& g, int& ix0, int& iy0,
                                   double dx[4], double dy[4])
{
	double q0 = cx, q1 = cy;
	if (comp == 0) q1 -= g.half_h; else q0 -= g.half_h;
	q0 = fmod_pow2g(fmod_pow2g(q0, g.Wd, g.invW) + g.Wd, g.Wd, g.invW);
	q1 = fmod_pow2g(fmod_pow2g(q1, g.Dd, g.invD) + g.Dd, g.Dd, g.invD);
	double fx = floor(q0 * g.nd), fy = floor(q1 * g.nd);
	if (!(fx >= 0.0 && fx <= g.nxd)) fx = 0.0;
	if (!(fy >= 0.0 && fy <= g.nyd)) fy = 0.0;
	ix0 = (int) (((unsigned) fx + g.nx - 1) & (g.nx - 1));
	iy0 = (int) (((unsigned) fy + g.ny - 1) & (g.ny - 1));
	double rx[4], ry[4];
#pragma unroll
	for (int d = 0; d < 4; ++d) {
		double x = (fx + d - 1.0) * g.h, y = (fy + d - 1.0) * g.h;
		if (comp == 0) y += g.half_h; else x += g.half_h;
		rx[d] = x - cx;
		ry[d] = y - cy;
	}
	window_deltas(rx, ry, g.n, dx, dy);
}

// sum_taps (w * vol) * field[idx], taps in the reference order (x fastest).
__device__ inline double ib_gather(const double* __restrict__ f, unsigned nx, unsigned ny, int ix0, int iy0,
                                   const double dx[4], const double dy[4], double vol)
{
	double value = 0.0;
#pragma unroll
	for (int b = 0; b < 4; ++b) {
		int iy = iy0 + b; if (iy >= (int) ny) iy -= ny;
#pragma unroll
		for (int a = 0; a < 4; ++a) {
			int ix = ix0 + a; if (ix >= (int) nx) ix -= nx;
			double w = dx[a] * dy[b];
			value += w * vol * f[ix + (size_t) nx * iy];
		}
	}
	return value;
}

// ------------------------------------------------------ lane exchange ----
// Whole-wave DPP shifts by one lane (GFX9 wave_shl:1 / wave_shr:1): lane k
// receives lane k+1's (SHL) or lane k-1's (SHR) value; the end lane keeps its
// own.  No LDS traffic.
constexpr int DPP_WAVE_SHL1 = 0x130;
constexpr int DPP_WAVE_SHR1 = 0x138;

template<int CTRL>
__device__ inline double lane_shift(double v)
{
	int lo = __double2loint(v), hi = __double2hiint(v);
	lo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xF, 0xF, false);
	hi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xF, 0xF, false);
	return __hiloint2double(hi, lo);
}

template<int CTRL>
__device__ inline float lane_shift(float v)
{
	return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Two adjacent cells of one row, loaded and stored as one vector.
template<typename T> struct Pair;
template<> struct Pair<double> { using type = double2; };
template<> struct Pair<float> { using type = float2; };

// g in [-n, 2n) folded into [0, n)
__device__ inline int wrap_once(int g, int n)
{
	g += (g < 0) ? n : 0;
	g -= (g >= n) ? n : 0;
	return g;
}

// Column and row of cell i of a grid nx wide.  Every grid here is a power of
// two wide (the context rejects other sizes), so this is a mask and a shift
// -- the 64-bit i % nx and i / nx it replaces are long division sequences.
__device__ inline void cell_xy(size_t i, int nx, int& gx, int& gy)
{
	gx = (int) ((unsigned) i & (unsigned) (nx - 1));
	gy = (int) (i >> __builtin_ctz((unsigned) nx));
}

// XCD-aware block order: consecutive block ids go to different XCDs (round
// robin over 8), so map them back to spatially consecutive tiles per XCD.
__device__ inline unsigned xcd_tile(unsigned bid, unsigned nblk)
{
	const unsigned X = 8;
	if (nblk % X) return bid;
	return (bid % X) * (nblk / X) + bid / X;
}

// The same for any block count: XCD x receives blocks x, x + 8, ... (there
// are (nblk - x + 7) / 8 of them) and gives them the consecutive indices
// after those of XCDs 0 .. x - 1.  A bijection of 0 .. nblk - 1.
__device__ inline unsigned xcd_remap(unsigned bid, unsigned nblk)
{
	const unsigned X = 8, x = bid % X, q = nblk / X, r = nblk % X;
	// XCDs below r receive q + 1 blocks, the others q
	const unsigned before = x * q + (x < r ? x : r);
	return before + bid / X;
}

}  // namespace afm
