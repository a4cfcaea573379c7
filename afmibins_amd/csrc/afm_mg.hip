// afm_mg.hip -- geometric multigrid for the doubly periodic cell-centred
// Poisson problem on gfx950 (mg/solver.h:35-271 restated).
//
// The algorithm is the reference's: nested-iteration start, mean removal,
// then V(nu1, nu2) cycles on the residual equation until RMS(r) <= tol, with
// 16-tap full-weighting restriction, bilinear prolongation and the coarsest
// level solved by mean removal + CG.  The smoother is red-black Gauss-Seidel
// (the reference's lexicographic sweep is inherently serial); each row update
// uses the reference's arithmetic (alg/smooth.h:39-50) and stencil order.
//
// Levels with n >= 64 run as tiled global kernels:
//   k_smooth        nu sweeps per launch, temporally blocked (halo 2 nu) in
//                   registers, optionally fused with "x += P e" or "x = P e";
//   k_resid_restr   residual b - L x on a fine tile in LDS, restricted;
//   k_update_resid  x += e, r = rhs - L x, deterministic RMS, stop test.
// Levels with n <= 32 run inside one workgroup (k_coarse), all in LDS.
// The stop test runs on the GPU (SolveCtl); every kernel of a cycle checks
// ctl->done first, so the host enqueues cycles in batches and synchronises
// once per batch.
#include <type_traits>
#include "afm_device.h"

namespace afm {

constexpr int RTX = 32, RTY = 8;         // restriction tile (coarse cells)
constexpr int CMAXL = 6;                 // finest level handled in LDS (64x64)
constexpr int FCG_BLOCKS = 4096;         // most blocks (and the partial-sum stride) of the flexible-CG update
constexpr int CTHREADS = 1024;

enum { M_NORMAL = 0, M_ZERO = 1, M_PADD = 2, M_PSET = 3 };

template<typename T>
__device__ inline T prolong_at(const T* __restrict__ xc, int gx, int gy, int hx, int hy)
{
	// mg/solver.h:141-148 with the periodic (x, z) plane: fine cell 2a+1 has
	// primary coarse a, fine cell 2a+2 has primary a+1.
	int px, ox, py, oy;
	if (gx & 1) { px = (gx - 1) >> 1; ox = px + 1; if (ox == hx) ox = 0; }
	else { px = gx >> 1; if (px == hx) px = 0; ox = px - 1; if (ox < 0) ox += hx; }
	if (gy & 1) { py = (gy - 1) >> 1; oy = py + 1; if (oy == hy) oy = 0; }
	else { py = gy >> 1; if (py == hy) py = 0; oy = py - 1; if (oy < 0) oy += hy; }
	T P = xc[px + (size_t) hx * py], Sx = xc[ox + (size_t) hx * py], Sy = xc[px + (size_t) hx * oy],
	  Dg = xc[ox + (size_t) hx * oy];
	return T(9. / 16.) * P + T(3. / 16.) * (Sx + Sy) + T(1. / 16.) * Dg;
}

// ---------------------------------------------------------------- smooth --
// Register-resident red-black smoother: S sweeps (2 S half-sweeps) per launch.
//
// A block covers a region of 128 x LY cells (LY = NW * RPW) whose corner is
// (X0 - H, Y0 - H) and writes its inner (128 - 2H) x (LY - 2H) cells (tiles
// at the far edges are clipped).  Lane k of wave w holds region columns 2k
// and 2k+1 of rows RPW*w .. RPW*w + RPW - 1, both x and b, in registers.  A
// red/black update of the point in column 2k+o reads its other-colour E/W
// neighbours from its own pair and from lane k-1 or k+1 (a DPP wave shift),
// its N/S neighbours from the rows above and below; the first and last rows
// of each wave go through LDS once per half-sweep.  Every update reads only
// neighbours one cell further out, so after half-sweep ph the cells within
// H-1-ph of the output are exact and the ring outside may hold anything
// (including the lane-shift fill at the region's edge columns and the unset
// LDS rows above and below).
//
// The update is the reference's row arithmetic (alg/smooth.h:39-50) in the
// stored column order (stencil5), with "/ diagonal" as "* (1/diagonal)": the
// diagonal -4 n^2 is a power of two, so the two round identically.
// MODE: M_NORMAL x = xin; M_ZERO x = 0; M_PADD x = xin + P xc; M_PSET x = P xc
// (P: bilinear prolongation from the coarse level, mg/solver.h:71-151).
//
// Epilogues (halo H = 2S + 2, so x is exact on the output ring 2):
//   RESTRICT  the residual b - L x on the output ring 1, restricted with the
//             16-tap full weighting (mg/solver.h:153-181) into the coarse tile;
//   UPDATE    (finest level, after the V-cycle) xs' = xs + x and the outer
//             residual r = rhs - L xs' in FP64 on the output tile, with the
//             block's sum of r^2 (mg/solver.h:226-233).
enum { EPI_STORE = 0, EPI_RESTRICT = 1, EPI_UPDATE = 2 };

// bnd[(slot * 2 + which) * 64 + lane][col]: slot w+1 holds wave w's first
// (which 0) and last (which 1) rows; slots 0 and NW+1 are never written.
template<typename T>
__device__ inline T& bnd_at(T* bnd, int slot, int which, int lane, int col)
{
	return bnd[((slot * 2 + which) * 64 + lane) * 2 + col];
}

template<int NW> constexpr int bnd_size() { return (NW + 2) * 2 * 64 * 2; }

// Neighbours E, W, N, S of the cell (row i, column o) of a lane's strip v:
// the pair partner, the adjacent lane's column (DPP), the rows above and
// below (the strip itself or the boundary rows up / dn of the adjacent waves).
template<typename T, int RPW>
__device__ inline void strip_nbrs(const T (&v)[RPW][2], int i, int o, T up, T dn, T& E, T& W, T& N, T& S)
{
	if (o == 0) { W = lane_shift<DPP_WAVE_SHR1>(v[i][1]); E = v[i][1]; }
	else { W = v[i][0]; E = lane_shift<DPP_WAVE_SHL1>(v[i][0]); }
	S = i > 0 ? v[i - 1][o] : up;
	N = i < RPW - 1 ? v[i + 1][o] : dn;
}

template<typename T, int RPW>
__device__ inline void bnd_publish(T* bnd, const T (&v)[RPW][2], int w, int lane)
{
	bnd_at(bnd, w + 1, 0, lane, 0) = v[0][0]; bnd_at(bnd, w + 1, 0, lane, 1) = v[0][1];
	bnd_at(bnd, w + 1, 1, lane, 0) = v[RPW - 1][0]; bnd_at(bnd, w + 1, 1, lane, 1) = v[RPW - 1][1];
}

// One lane's strip of a tile: coordinates and the x / b registers.
template<typename T, int S, int H, int NW, int RPW, bool EDGE>
struct Strip {
	using T2 = typename Pair<T>::type;
	static constexpr int LY = NW * RPW;
	static_assert(RPW % 2 == 0 && H % 2 == 0, "strips start on even rows and columns (fixed colour per register)");
	int lane, w, gxu, gx, gyb;
	T xr[RPW][2], br[RPW][2];

	__device__ Strip(int X0, int Y0, int nx)
	{
		lane = threadIdx.x & 63;
		w = threadIdx.x >> 6;
		gxu = X0 - H + 2 * lane;                // first column of the pair, unwrapped (even)
		gx = EDGE ? wrap_once(gxu, nx) : gxu;
		gyb = Y0 - H + RPW * w;                 // first row of the wave, unwrapped (even)
	}
	__device__ int row(int i, int ny) const { return EDGE ? wrap_once(gyb + i, ny) : gyb + i; }
	// is (row i, column pair) part of the tile's output?
	__device__ bool out_col(int nx) const { return lane >= H / 2 && lane < 64 - H / 2 && gxu < nx; }
	__device__ bool out_row(int i, int ny) const
	{
		const int ly = RPW * w + i;
		return ly >= H && ly < LY - H && gyb + i < ny;
	}
	// the 5-point stencil at (i, o) with the periodic wrap flags of EDGE tiles
	template<bool DIAG, typename U>
	__device__ U stencil(int i, int o, U off, U dg, U c, U E, U W, U N, U Sv, int nx, int ny) const
	{
		if (EDGE) {
			const int px = wrap_once(gxu + o, nx), py = wrap_once(gyb + i, ny);
			return stencil5p2<U, DIAG>(off, dg, c, E, W, N, Sv, px == nx - 1, px == 0, py == ny - 1, py == 0);
		}
		return stencil5p2<U, DIAG>(off, dg, c, E, W, N, Sv, false, false, false, false);
	}

	template<int MODE>
	__device__ void load(const T* __restrict__ xin, const T* __restrict__ b, const T* __restrict__ xc, int nx, int ny)
	{
		if (MODE == M_NORMAL || MODE == M_PADD) {
#pragma unroll
			for (int i = 0; i < RPW; ++i) {
				const T2 v = *(const T2*) (xin + (size_t) nx * row(i, ny) + gx);
				xr[i][0] = v.x; xr[i][1] = v.y;
			}
		}
#pragma unroll
		for (int i = 0; i < RPW; ++i) {
			const T2 v = *(const T2*) (b + (size_t) nx * row(i, ny) + gx);
			br[i][0] = v.x; br[i][1] = v.y;
		}
		if (MODE == M_ZERO) {
#pragma unroll
			for (int i = 0; i < RPW; ++i) xr[i][0] = xr[i][1] = T(0);
		}
		if (MODE == M_PADD || MODE == M_PSET) {
			// coarse rows gyb/2 - 1 .. gyb/2 + RPW/2, columns m-1 .. m+1 (periodic)
			const int hx = nx >> 1, hy = ny >> 1, m = gxu / 2;
			int cc[3];
#pragma unroll
			for (int j = 0; j < 3; ++j) cc[j] = wrap_once(m - 1 + j, hx);
			T cw[RPW / 2 + 2][3];
#pragma unroll
			for (int r = 0; r < RPW / 2 + 2; ++r) {
				const size_t crow = (size_t) hx * wrap_once(gyb / 2 - 1 + r, hy);
#pragma unroll
				for (int j = 0; j < 3; ++j) cw[r][j] = xc[crow + cc[j]];
			}
#pragma unroll
			for (int i = 0; i < RPW; ++i) {
				// even fine row: primary coarse row i/2 + 1, other i/2; odd: (i+1)/2, (i+3)/2
				const int pr = (i & 1) ? (i + 1) / 2 : i / 2 + 1, orr = (i & 1) ? (i + 3) / 2 : i / 2;
#pragma unroll
				for (int o = 0; o < 2; ++o) {
					// even fine column: primary m, other m-1; odd: primary m, other m+1
					const int oc = o ? 2 : 0;
					const T P = cw[pr][1], Sx = cw[pr][oc], Sy = cw[orr][1], Dg = cw[orr][oc];
					const T p = T(9. / 16.) * P + T(3. / 16.) * (Sx + Sy) + T(1. / 16.) * Dg;
					xr[i][o] = (MODE == M_PADD) ? xr[i][o] + p : p;
				}
			}
		}
	}

	// 2S red-black half-sweeps; bnd: LDS boundary-row exchange.  With SUMS,
	// s64[i] receives off times the FP64 sum of the four neighbours the last
	// half-sweep (colour 1) read for row i's colour-1 cell (its final
	// colour-0 neighbours), for fcg_dots_edges.
	template<bool SUMS = false>
	__device__ void sweeps(T* bnd, int nx, int ny, T off, T dg, T inv, double off64 = 0.0, double* s64 = nullptr)
	{
		bnd_publish<T, RPW>(bnd, xr, w, lane);
		__syncthreads();
#pragma unroll
		for (int ph = 0; ph < 2 * S; ++ph) {
			const int c = ph & 1;               // colour (gx + gy) & 1 updated now
			// the cell of colour c in row i of this wave is column o = (c + i) & 1
			const T up = bnd_at(bnd, w, 1, lane, c);          // row above the wave, column c
			const T dn = bnd_at(bnd, w + 2, 0, lane, 1 - c);  // row below the wave, column 1 - c
#pragma unroll
			for (int i = 0; i < RPW; ++i) {
				const int o = (c + i) & 1;
				T Ea, We, Nv, Sv;
				strip_nbrs<T, RPW>(xr, i, o, up, dn, Ea, We, Nv, Sv);
				const T v = stencil<false, T>(i, o, off, dg, T(0), Ea, We, Nv, Sv, nx, ny);
				xr[i][o] = (br[i][o] - v) * inv;
				if (SUMS && ph == 2 * S - 1)
					s64[i] = off64 * (((double) Ea + (double) We) + ((double) Nv + (double) Sv));
			}
			if (ph < 2 * S - 1) {
				bnd_at(bnd, w + 1, 0, lane, c) = xr[0][c];
				bnd_at(bnd, w + 1, 1, lane, 1 - c) = xr[RPW - 1][1 - c];
				__syncthreads();
			}
		}
	}

	__device__ void store(T* __restrict__ xout, int nx, int ny) const
	{
		if (!out_col(nx)) return;
#pragma unroll
		for (int i = 0; i < RPW; ++i)
			if (out_row(i, ny)) {
				T2 v;
				v.x = xr[i][0]; v.y = xr[i][1];
				*(T2*) (xout + (size_t) nx * (gyb + i) + gx) = v;
			}
	}

	// residual b - L x (mg/laplacian.h rows in stored order, diagonal
	// included) on every cell, exact on the output ring 1, restricted into
	// rc; b is overwritten.  bx, brs: two more LDS exchange buffers.
	__device__ void restrict_to(T* __restrict__ rc, T* bx, T* brs, int X0, int Y0, int nx, int ny, T off, T dg)
	{
		bnd_publish<T, RPW>(bx, xr, w, lane);
		__syncthreads();
		T up[2], dn[2];
#pragma unroll
		for (int o = 0; o < 2; ++o) { up[o] = bnd_at(bx, w, 1, lane, o); dn[o] = bnd_at(bx, w + 2, 0, lane, o); }
		T (&rr)[RPW][2] = br;
#pragma unroll
		for (int i = 0; i < RPW; ++i)
#pragma unroll
			for (int o = 0; o < 2; ++o) {
				T Ea, We, Nv, Sv;
				strip_nbrs<T, RPW>(xr, i, o, up[o], dn[o], Ea, We, Nv, Sv);
				rr[i][o] = br[i][o] - stencil<true, T>(i, o, off, dg, xr[i][o], Ea, We, Nv, Sv, nx, ny);
			}
		bnd_publish<T, RPW>(brs, rr, w, lane);
		__syncthreads();
		// coarse cell (j, q) of the region takes fine rows 2q-1 .. 2q+2 and
		// columns 2j-1 .. 2j+2: lane j-1's second column, this lane's pair,
		// lane j+1's first column.  Taps in the reference order (x fastest).
		T ru[2], rd[2];
#pragma unroll
		for (int o = 0; o < 2; ++o) { ru[o] = bnd_at(brs, w, 1, lane, o); rd[o] = bnd_at(brs, w + 2, 0, lane, o); }
		const int cnx = nx >> 1, cny = ny >> 1;
		const int j = lane, cgx = (X0 - H) / 2 + j;
		const bool colok = j >= H / 2 && j < 64 - H / 2 && cgx < cnx;
#pragma unroll
		for (int qq = 0; qq < RPW / 2; ++qq) {
			const int q = RPW / 2 * w + qq;     // coarse row of the region
			T o = T(0);
#pragma unroll
			for (int dz = 0; dz < 4; ++dz) {
				const int i = 2 * qq - 1 + dz;    // fine row of the strip, -1 .. RPW
				const int ic = i < 0 ? 0 : (i >= RPW ? RPW - 1 : i);
				const T c0 = i < 0 ? ru[0] : (i >= RPW ? rd[0] : rr[ic][0]);
				const T c1 = i < 0 ? ru[1] : (i >= RPW ? rd[1] : rr[ic][1]);
				const T cl = lane_shift<DPP_WAVE_SHR1>(c1), cr = lane_shift<DPP_WAVE_SHL1>(c0);
				const int wz = 4 - abs(2 * (dz - 2) + 1);
				o += T((1 * wz) / 64.) * cl;
				o += T((3 * wz) / 64.) * c0;
				o += T((3 * wz) / 64.) * c1;
				o += T((1 * wz) / 64.) * cr;
			}
			const int cgy = (Y0 - H) / 2 + q;
			if (colok && q >= H / 2 && q < LY / 2 - H / 2 && cgy < cny) rc[(size_t) cnx * cgy + cgx] = o;
		}
	}

	// xs' = xs + x on the output ring 1, r = rhs - L xs' on the output tile
	// (the reference's update and residual, mg/solver.h:226-233, in FP64).
	// Writes xs', r and returns this lane's sum of r^2 (rows, then columns).
	template<typename TR>
	__device__ double update(const double* __restrict__ xs, const double* __restrict__ rhs, double* __restrict__ xs_out,
	                         TR* __restrict__ res, double* bx, int nx, int ny, double off, double dg, bool write) const
	{
		double xn[RPW][2];
		const bool ring1_col = lane >= H / 2 - 1 && lane <= 64 - H / 2;
#pragma unroll
		for (int i = 0; i < RPW; ++i) {
			const int ly = RPW * w + i;
			if (ring1_col && ly >= H - 1 && ly <= LY - H) {
				const double2 v = *(const double2*) (xs + (size_t) nx * row(i, ny) + gx);
				xn[i][0] = v.x + (double) xr[i][0];
				xn[i][1] = v.y + (double) xr[i][1];
			} else {
				xn[i][0] = xn[i][1] = 0.0;
			}
		}
		bnd_publish<double, RPW>(bx, xn, w, lane);
		__syncthreads();
		double up[2], dn[2];
#pragma unroll
		for (int o = 0; o < 2; ++o) { up[o] = bnd_at(bx, w, 1, lane, o); dn[o] = bnd_at(bx, w + 2, 0, lane, o); }
		double acc = 0.0;
		const bool oc = out_col(nx);
#pragma unroll
		for (int i = 0; i < RPW; ++i) {
			double rv[2];
#pragma unroll
			for (int o = 0; o < 2; ++o) {
				double Ea, We, Nv, Sv;
				strip_nbrs<double, RPW>(xn, i, o, up[o], dn[o], Ea, We, Nv, Sv);
				rv[o] = stencil<true, double>(i, o, off, dg, xn[i][o], Ea, We, Nv, Sv, nx, ny);
			}
			if (oc && out_row(i, ny)) {
				const size_t at = (size_t) nx * (gyb + i) + gx;
				const double2 r0 = *(const double2*) (rhs + at);
				const double ra = r0.x - rv[0], rb = r0.y - rv[1];
				if (write) {
					double2 xo;
					xo.x = xn[i][0]; xo.y = xn[i][1];
					*(double2*) (xs_out + at) = xo;
					typename Pair<TR>::type ro;
					ro.x = (TR) ra; ro.y = (TR) rb;
					*(typename Pair<TR>::type*) (res + at) = ro;
				}
				acc += ra * ra;
				acc += rb * rb;
			}
		}
		return acc;
	}

	// Flexible CG: z = x (exact on the output ring 2) as the preconditioned
	// residual.  Stores z on the output tile and adds this lane's z.r_k (r_k
	// as the cycle took it: b, in registers), z.r_{k-1} (rkm1, rounded the
	// same way) and z.(L z), products and sums in FP64 (rows, then columns).
	// Stores z and sums z.r_k, z.r_{k-1} and z.Lz without exchanging z's
	// boundary rows: every grid edge joins a colour-0 and a colour-1 cell, so
	// z.Lz = sum dg z^2 + 2 sum over colour-1
	// cells of z times off (sum of its neighbours), whose sums the last
	// half-sweep kept (sweeps<true>); each edge is counted once, by the tile
	// that writes its colour-1 cell.  Products and sums in FP64.
	__device__ void fcg_dots_edges(T* __restrict__ z, const T* __restrict__ rkm1, int nx, int ny, double dg,
	                               const double* s64, double (&acc)[3]) const
	{
		const bool oc = out_col(nx);
#pragma unroll
		for (int i = 0; i < RPW; ++i) {
			if (oc && out_row(i, ny)) {
				const size_t at = (size_t) nx * (gyb + i) + gx;
				T2 zo;
				zo.x = xr[i][0]; zo.y = xr[i][1];
				*(T2*) (z + at) = zo;
				const T2 q = *(const T2*) (rkm1 + at);
				const double z0 = (double) xr[i][0], z1 = (double) xr[i][1];
				// row i's colour-1 cell is column (1 + i) & 1 (rows and columns of a
				// strip start even)
				const double zb = ((1 + i) & 1) ? z1 : z0;
				acc[0] += z0 * (double) br[i][0];
				acc[0] += z1 * (double) br[i][1];
				acc[1] += z0 * (double) q.x;
				acc[1] += z1 * (double) q.y;
				acc[2] += dg * z0 * z0;
				acc[2] += dg * z1 * z1;
				acc[2] += 2.0 * zb * s64[i];
			}
		}
	}
};

// Waves per block and rows per wave.  SH = 0 for the large levels (tall
// 1024-thread tiles: few halo rows per output row); SH = 1 for levels of at
// most SMALL_CELLS cells, whose launches are few blocks on an idle chip:
// there a block's serial chain of row updates sets the time, so each wave
// holds fewer rows; SH = 2 for plain smoothing launches (M_NORMAL, the
// smoother entry point) at the large levels, which stream fastest in
// 256-thread tiles.
template<int S, int SH = 0> struct SmoothShape;
template<> struct SmoothShape<1, 0> { static constexpr int NW = 6, RPW = 6; };
template<> struct SmoothShape<2, 0> { static constexpr int NW = 16, RPW = 4; };
template<> struct SmoothShape<1, 1> { static constexpr int NW = 8, RPW = 2; };
template<> struct SmoothShape<2, 1> { static constexpr int NW = 8, RPW = 2; };
template<> struct SmoothShape<1, 2> { static constexpr int NW = 6, RPW = 6; };
template<> struct SmoothShape<2, 2> { static constexpr int NW = 4, RPW = 10; };
// the flexible CG's finest post-smoothing (k_smooth_fcg): 512-thread blocks
// of 8 rows per wave; its 80-odd VGPRs allow one 1024-thread block per CU
// but two of these (2048^2: 720 blocks either way, 26.5 -> 22.0 us)
#ifndef AFM_SHAPE_FCG_NW
#define AFM_SHAPE_FCG_NW 8
#endif
#ifndef AFM_SHAPE_FCG_RPW
#define AFM_SHAPE_FCG_RPW 8
#endif
template<> struct SmoothShape<1, 3> { static constexpr int NW = 6, RPW = 6; };
template<> struct SmoothShape<2, 3> { static constexpr int NW = AFM_SHAPE_FCG_NW, RPW = AFM_SHAPE_FCG_RPW; };
constexpr size_t SMALL_CELLS = (size_t) 512 * 512;
// the prolongation + post-smoothing launch at levels above SMALL_CELLS up to
// MID_CELLS (1024^2 at the benchmark's 2048^2): 256-thread blocks of 6 rows
// per wave, 576 blocks at 1024^2 instead of 171 (8.1 -> 6.6 us per launch;
// the pre-smoothing launch there keeps the 16-wave tiles: 7.5 against
// 7.6-8.6 us for the shorter ones)
template<> struct SmoothShape<1, 5> { static constexpr int NW = 4, RPW = 6; };
template<> struct SmoothShape<2, 5> { static constexpr int NW = 4, RPW = 6; };
constexpr size_t MID_CELLS = (size_t) 1024 * 1024;

// halo of a launch: 2 cells per sweep, 2 more for a residual epilogue
template<int S, int EPI> constexpr int smooth_halo() { return 2 * S + (EPI != EPI_STORE ? 2 : 0); }
template<int S, int EPI> constexpr int smooth_txo() { return 128 - 2 * smooth_halo<S, EPI>(); }
template<int S, int EPI, int SH = 0>
constexpr int smooth_tyo() { return SmoothShape<S, SH>::NW * SmoothShape<S, SH>::RPW - 2 * smooth_halo<S, EPI>(); }

template<int S, int EPI, int SH = 0>
static inline unsigned smooth_blocks(int nx, int ny)
{
	constexpr int TXO = smooth_txo<S, EPI>(), TYO = smooth_tyo<S, EPI, SH>();
	return (unsigned) (((nx + TXO - 1) / TXO) * ((ny + TYO - 1) / TYO));
}

// tile origin of this block and whether its region wraps
template<int S, int EPI, int SH = 0>
__device__ inline bool smooth_tile_origin(int nx, int ny, int& X0, int& Y0)
{
	constexpr int H = smooth_halo<S, EPI>(), TXO = smooth_txo<S, EPI>(), TYO = smooth_tyo<S, EPI, SH>();
	constexpr int LY = SmoothShape<S, SH>::NW * SmoothShape<S, SH>::RPW;
	const int ntx = (nx + TXO - 1) / TXO;
	const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
	X0 = (int) (tile % ntx) * TXO;
	Y0 = (int) (tile / ntx) * TYO;
	return X0 - H < 0 || X0 - H + 128 > nx || Y0 - H < 0 || Y0 - H + LY > ny;
}

template<typename T, int S, int MODE, int EPI, bool EDGE, int SH>
__device__ inline void smooth_body(const T* __restrict__ xin, const T* __restrict__ b, const T* __restrict__ xc,
                                   T* __restrict__ xout, T* __restrict__ rc, int nx, int ny, T off, T dg, T inv,
                                   int X0, int Y0, T* bnd)
{
	constexpr int NW = SmoothShape<S, SH>::NW, RPW = SmoothShape<S, SH>::RPW, H = smooth_halo<S, EPI>();
	Strip<T, S, H, NW, RPW, EDGE> st(X0, Y0, nx);
	st.template load<MODE>(xin, b, xc, nx, ny);
	st.sweeps(bnd, nx, ny, off, dg, inv);
	st.store(xout, nx, ny);
	if (EPI == EPI_RESTRICT) st.restrict_to(rc, bnd + bnd_size<NW>(), bnd + 2 * bnd_size<NW>(), X0, Y0, nx, ny, off, dg);
}

// The multigrid stop test of the previous V-cycle (k_update_fin's work),
// folded into the first launch of the next cycle: block 0 alone runs it and
// nothing else in that launch reads the control block.  The cycle index
// comes from the host, so no block reads a count being written.
struct MGFold {
	SolveCtl* ctl;            // nullptr: no test in this launch
	const double* partials;   // the update kernel's r^2 per block
	unsigned nblk;
	int it;                   // cycles completed once the tested one is counted
	double N;                 // cells of the finest level
	int fcg;                  // flexible CG: also p.r and p.Lp (columns 1, 2, stride FCG_BLOCKS)
};

// red_total over a 256-thread block, computed by the first four waves of a
// larger one (same order, so the same bits as the one-block finalisers);
// NV columns of partials, `stride` apart; totals in thread 0
template<int NV>
__device__ inline void red_totals_256(const double* __restrict__ partials, size_t stride, unsigned nblk,
                                      double (&t)[NV])
{
	__shared__ double wsum[NV][4];
	const int tid = (int) threadIdx.x;
#pragma unroll
	for (int j = 0; j < NV; ++j) {
		double s = tid < 256 ? strided_sum(partials + j * stride, tid, 256, nblk) : 0.0;
		for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
		if (tid < 256 && (tid & 63) == 0) wsum[j][tid >> 6] = s;
	}
	__syncthreads();
#pragma unroll
	for (int j = 0; j < NV; ++j) {
		t[j] = 0.0;
		if (tid == 0)
			for (int w = 0; w < 4; ++w) t[j] += wsum[j][w];
	}
}

__device__ inline double red_total_256(const double* __restrict__ partials, unsigned nblk)
{
	double t[1];
	red_totals_256<1>(partials, 0, nblk, t);
	return t[0];
}

// The stop test after a flexible-CG update (r.r in column 0) and the two
// carried scalars p_{k}.r_{k+1}, p_k.L p_k for the next iteration; thread 0
// writes.  Same sums, same order, in the fold and in k_fcg_fin.
__device__ inline void fcg_stop_write(SolveCtl* c, const double (&t)[3], int it, double N)
{
	const double rms = sqrt(t[0]) / sqrt(N);
	c->count = it;
	c->res = rms;
	c->frho = t[1];
	c->fdel = t[2];
	c->done = !(rms > c->tol && it < c->max_iter);
}

__device__ inline void mg_fold_test(const MGFold& f)
{
	__shared__ int already;
	if (threadIdx.x == 0) already = f.ctl->done;
	__syncthreads();
	if (already) return;
	if (f.fcg) {
		double t[3];
		red_totals_256<3>(f.partials, FCG_BLOCKS, f.nblk, t);
		if (threadIdx.x == 0) fcg_stop_write(f.ctl, t, f.it, f.N);
		return;
	}
	const double tot = red_total_256(f.partials, f.nblk);
	if (threadIdx.x == 0) {
		SolveCtl* c = f.ctl;
		const double rms = sqrt(tot) / sqrt(f.N);
		c->count = f.it;
		c->res = rms;
		c->done = !(rms > c->tol && f.it < c->max_iter);
	}
}

template<typename T, int S, int MODE, int EPI, int SH>
__global__ __launch_bounds__((64 * SmoothShape<S, SH>::NW)) void k_smooth(const T* __restrict__ xin, const T* __restrict__ b,
                                                                    const T* __restrict__ xc, T* __restrict__ xout,
                                                                    T* __restrict__ rc, int nx, int ny, T off, T dg,
                                                                    T inv, const int* __restrict__ done, MGFold fold)
{
	constexpr int NW = SmoothShape<S, SH>::NW;
	__shared__ T bnd[(EPI == EPI_RESTRICT ? 3 : 1) * bnd_size<NW>()];
	if (done && *done) return;
	int X0, Y0;
	if (smooth_tile_origin<S, EPI, SH>(nx, ny, X0, Y0))
		smooth_body<T, S, MODE, EPI, true, SH>(xin, b, xc, xout, rc, nx, ny, off, dg, inv, X0, Y0, bnd);
	else
		smooth_body<T, S, MODE, EPI, false, SH>(xin, b, xc, xout, rc, nx, ny, off, dg, inv, X0, Y0, bnd);
	// after the block's own tile: block 0 is among the first to start, so the
	// test overlaps the rest of the grid
	if (fold.ctl && blockIdx.x == 0) mg_fold_test(fold);
}

// Finest-level post-smoothing fused with the outer update: e = smooth(e_pre +
// P ec) against b (the V-cycle right-hand side, the current residual), then
// xs' = xs + e, r = rhs - L xs' and the block's sum of r^2 in partials.
template<typename TE, typename TR, int S>
__global__ __launch_bounds__(64 * SmoothShape<S>::NW) void k_smooth_update(
	const TE* __restrict__ epre, const TE* __restrict__ b, const TE* __restrict__ ec, const double* __restrict__ xs,
	const double* __restrict__ rhs, double* __restrict__ xs_out, TR* __restrict__ res, int nx, int ny, TE off, TE dg,
	TE inv, double off64, double dg64, const SolveCtl* __restrict__ ctl, double* __restrict__ partials)
{
	constexpr int NW = SmoothShape<S>::NW, RPW = SmoothShape<S>::RPW, H = smooth_halo<S, EPI_UPDATE>();
	__shared__ TE bnd[bnd_size<NW>()];
	__shared__ double bnd64[bnd_size<NW>()];
	// the stop flag only gates the stores of the new iterate and residual, so
	// its load is never waited for before the work
	const int stop = ctl->done;
	int X0, Y0;
	double acc[1];
	if (smooth_tile_origin<S, EPI_UPDATE>(nx, ny, X0, Y0)) {
		Strip<TE, S, H, NW, RPW, true> st(X0, Y0, nx);
		st.template load<M_PADD>(epre, b, ec, nx, ny);
		st.sweeps(bnd, nx, ny, off, dg, inv);
		acc[0] = st.template update<TR>(xs, rhs, xs_out, res, bnd64, nx, ny, off64, dg64, stop == 0);
	} else {
		Strip<TE, S, H, NW, RPW, false> st(X0, Y0, nx);
		st.template load<M_PADD>(epre, b, ec, nx, ny);
		st.sweeps(bnd, nx, ny, off, dg, inv);
		acc[0] = st.template update<TR>(xs, rhs, xs_out, res, bnd64, nx, ny, off64, dg64, stop == 0);
	}
	reduce_partial<1>(acc, partials);
}

// ---------------------------------------------------------- flexible CG --
// The mixed form's outer iteration: CG in FP64 on L x = b (mean removed),
// preconditioned by one V(nu1, nu2) cycle in FP32 (z = V(r)).  The V-cycle
// is not a symmetric operator (red-then-black in both smoothing legs) and
// its FP32 rounding varies with r, so the search direction uses the
// Polak-Ribiere / flexible coefficient (Golub and Ye, "Inexact
// preconditioned conjugate gradient method with inner-outer iteration",
// SIAM J. Sci. Comput. 21, 1999):
//   z_k = V(r_k);  beta_k = z_k.(r_k - r_{k-1}) / z_{k-1}.r_{k-1};
//   p_k = z_k + beta_k p_{k-1};  alpha_k = p_k.r_k / p_k.L p_k;
//   x_{k+1} = x_k + alpha_k p_k;  r_{k+1} = b - L x_{k+1}  (the true
//   residual, so the stop test is the reference's RMS(b - L x) <= tol).
// Two launches per iteration besides the V-cycle: its last post-smoothing
// (k_smooth_fcg, or k_fcg_dots under a coarse-kernel cycle) stores z and
// sums z.r_k, z.r_{k-1}, z.L z; k_fcg_update forms beta and alpha from
// those sums and the carried scalars (p.L p expanded through L p_k = L z_k
// + beta q_{k-1}, q_{k-1} = (r_{k-1} - r_k) / alpha_{k-1}, L symmetric), then
// p, x and r in one pass.  The residual is formed in FP64 (its sum of
// squares, the stop test, p.r) and stored rounded to the cycle's type: the
// cycle takes it so, and beta's and alpha's z.r terms use that same rounded
// vector (products and sums in FP64).  The search direction p is formed in
// FP64 and stored rounded to the cycle's type too (the next iteration reads
// it so; every tile forms its halo's p from the same stored values, so the
// tiles agree): the recurrence for p.L p then holds to about 1e-7 relative.
// These roundings change only the step coefficients, by about 1e-7
// relative, and the search directions by FP32 rounding; the solution and
// the tested residual stay FP64.  All sums are fixed-order.

// block sums with an explicit column stride (the dots launch can have more
// blocks than MAX_RED_BLOCKS)
template<int NV>
__device__ inline void reduce_partial_s(double (&v)[NV], double* partials, size_t stride)
{
	block_sum<NV>(v);
	if (threadIdx.x == 0) {
#pragma unroll
		for (int j = 0; j < NV; ++j) partials[j * stride + blockIdx.x] = v[j];
	}
}

// Finest-level post-smoothing of the preconditioner cycle: z = smooth(e_pre
// + P ec) against b (the cycle's right-hand side, r_k rounded), stored, with
// the block's z.r_k, z.r_{k-1} and z.L z (at k = 0 rkm1 is b, and the
// column unused).
template<typename TE, int S>
__global__ __launch_bounds__((64 * SmoothShape<S, 3>::NW)) void k_smooth_fcg(
	const TE* __restrict__ epre, const TE* __restrict__ b, const TE* __restrict__ ec, TE* __restrict__ z,
	const TE* __restrict__ rkm1, int nx, int ny, TE off, TE dg, TE inv, double off64, double dg64,
	double* __restrict__ partials, unsigned stride)
{
	constexpr int NW = SmoothShape<S, 3>::NW, RPW = SmoothShape<S, 3>::RPW, H = smooth_halo<S, EPI_UPDATE>();
	__shared__ TE bnd[bnd_size<NW>()];
	int X0, Y0;
	double acc[3] = {0.0, 0.0, 0.0};
	double s64[RPW];   // the colour-1 cells' neighbour sums (fcg_dots_edges)
	if (smooth_tile_origin<S, EPI_UPDATE, 3>(nx, ny, X0, Y0)) {
		Strip<TE, S, H, NW, RPW, true> st(X0, Y0, nx);
		st.template load<M_PADD>(epre, b, ec, nx, ny);
		st.template sweeps<true>(bnd, nx, ny, off, dg, inv, off64, s64);
		st.fcg_dots_edges(z, rkm1, nx, ny, dg64, s64, acc);
	} else {
		Strip<TE, S, H, NW, RPW, false> st(X0, Y0, nx);
		st.template load<M_PADD>(epre, b, ec, nx, ny);
		st.template sweeps<true>(bnd, nx, ny, off, dg, inv, off64, s64);
		st.fcg_dots_edges(z, rkm1, nx, ny, dg64, s64, acc);
	}
	reduce_partial_s<3>(acc, partials, stride);
}

// the same sums for a z that is already stored (grids of the coarse kernel)
template<typename TE>
__global__ __launch_bounds__(256) void k_fcg_dots(const TE* __restrict__ z, const TE* __restrict__ rk,
                                                  const TE* __restrict__ rkm1, int nx, int ny, double off,
                                                  double dg, double* __restrict__ partials)
{
	const size_t N = (size_t) nx * ny;
	double acc[3] = {0.0, 0.0, 0.0};
	for (size_t i = blockIdx.x * 256 + threadIdx.x; i < N; i += (size_t) gridDim.x * 256) {
		int gx, gy;
		cell_xy(i, nx, gx, gy);
		const size_t ie = gx == nx - 1 ? i - (nx - 1) : i + 1, iw = gx == 0 ? i + (nx - 1) : i - 1;
		const size_t in_ = gy == ny - 1 ? (size_t) gx : i + nx, is = gy == 0 ? i + N - nx : i - nx;
		const double zc = (double) z[i];
		const double lz = stencil5p2<double, true>(off, dg, zc, (double) z[ie], (double) z[iw], (double) z[in_],
		                                         (double) z[is], gx == nx - 1, gx == 0, gy == ny - 1, gy == 0);
		acc[0] += zc * (double) rk[i];
		acc[1] += zc * (double) rkm1[i];
		acc[2] += zc * lz;
	}
	reduce_partial<3>(acc, partials);
}

template<typename TE>
struct FcgArgs {
	const TE* z;              // z_k
	const TE* p_old;          // p_{k-1}; nullptr at k == 0
	TE* p_new;                // p_k, stored in the cycle's type
	const double* x_old;      // x_k; nullptr when x_0 = 0 (k == 0 without a start)
	double* x_new;            // x_{k+1}
	const double* b;          // right-hand side, mean removed
	TE* r;                    // r_{k+1} rounded to TE (the next cycle's right-hand side)
	SolveCtl* ctl;
	const double* fpart;      // the dots launch's sums, 3 columns `fstride` apart
	unsigned fstride, fnblk;
	double* gpart;            // this launch's r.r, p.r, p.L p (stride FCG_BLOCKS)
	int k;
};

// beta_k and alpha_k from the dots launch's partials (every block sums them
// in the same order, so all get the same bits) and the carried scalars;
// block 0 records (z.r_k, alpha_k) in the parity slot k & 1, which no block
// of this launch reads.  Returns the stop flag as the launch found it.
template<typename TE, bool FIRST>
__device__ inline bool fcg_coefficients(const FcgArgs<TE>& a, double& beta, double& alpha)
{
	__shared__ double coef[2];
	__shared__ int stop;
	SolveCtl* c = a.ctl;
	const int tid = (int) threadIdx.x, nt = (int) blockDim.x;
	double gam = 0.0, alp = 1.0, del = 0.0, rho = 0.0;
	if (tid == 0) {
		stop = c->done;
		if (!FIRST) {
			gam = c->fg[(a.k - 1) & 1][0];
			alp = c->fg[(a.k - 1) & 1][1];
			del = c->fdel;
			rho = c->frho;
		}
	}
	double s[3] = {0.0, 0.0, 0.0};
	for (unsigned q0 = 0; q0 < a.fnblk; q0 += 2 * nt) {
		// two partials per thread and column in flight
		double v[3][2];
#pragma unroll
		for (int u = 0; u < 2; ++u) {
			const unsigned q = q0 + u * nt + tid;
#pragma unroll
			for (int j = 0; j < 3; ++j) v[j][u] = q < a.fnblk ? a.fpart[(size_t) j * a.fstride + q] : 0.0;
		}
#pragma unroll
		for (int j = 0; j < 3; ++j) { s[j] += v[j][0]; s[j] += v[j][1]; }
	}
	block_sum<3>(s);
	if (tid == 0) {
		double b = 0.0, pLp = s[2], pr = s[0];
		if (!FIRST) {
			b = (s[0] - s[1]) / gam;
			const double zq = (s[1] - s[0]) / alp;
			pLp = s[2] + 2.0 * b * zq + b * b * del;
			pr = s[0] + b * rho;
		}
		const double al = pr / pLp;
		coef[0] = b;
		coef[1] = al;
		if (blockIdx.x == 0 && !stop) { c->fg[a.k & 1][0] = s[0]; c->fg[a.k & 1][1] = al; }
	}
	__syncthreads();
	beta = coef[0];
	alpha = coef[1];
	return stop != 0;
}

// The update over tiles of 128 x 32 cells in registers, like the smoother:
// lane k of wave w holds columns 2k, 2k+1 of rows 4w .. 4w+3 of a region
// whose outer 2 columns and 1 row on each side are halo (a lane holds a
// column pair); p_k and x_{k+1} are formed on the whole region, E/W
// neighbours by DPP, N/S through LDS boundary rows, and r_{k+1} = b - L
// x_{k+1} on the inner 124 x 30 cells (1.10x the algorithmic reads of z, p
// and x; a 2-row halo, 124 x 28, read 1.18x: 37.5 -> 37.2 us).  Every cell of the
// grid is written by one tile; a halo cell's p and x are formed from the
// same inputs by the same operations, so its value equals the stored one.
// Two 512-thread blocks per CU walk tiles blockIdx.x, + gridDim.x, ... (one
// coefficient prologue per block; the partial sums per block follow that
// fixed order): while one block computes, the other's loads are in flight.
// At 2048^2 the 1173 short tiles spread over the blocks more evenly than
// taller ones, which outweighs their more halo reads.
constexpr int FU_NW = 8, FU_RPW = 4, FU_H = 2, FU_HY = 1, FU_LY = FU_NW * FU_RPW;
constexpr int FU_TXO = 128 - 2 * FU_H, FU_TYO = FU_LY - 2 * FU_HY;

static inline unsigned fcg_update_blocks(int nx, int ny)
{
	return (unsigned) (((nx + FU_TXO - 1) / FU_TXO) * ((ny + FU_TYO - 1) / FU_TYO));
}

template<typename TE, bool FIRST, bool XZ>
__global__ __launch_bounds__(64 * FU_NW) void k_fcg_update(FcgArgs<TE> a, int nx, int ny, double off, double dg)
{
	using TE2 = typename Pair<TE>::type;
	__shared__ double bp[bnd_size<FU_NW>()], bx[bnd_size<FU_NW>()];
	double beta, alpha;
	if (fcg_coefficients<TE, FIRST>(a, beta, alpha)) return;
	// b is the right-hand side with its mean (k_mean); r = (b - mean) - L x
	const double bmean = a.ctl->mean;
	const int lane = (int) threadIdx.x & 63, w = (int) threadIdx.x >> 6;
	const int ntx = (nx + FU_TXO - 1) / FU_TXO;
	const unsigned ntiles = (unsigned) (ntx * ((ny + FU_TYO - 1) / FU_TYO));
	double acc[3] = {0.0, 0.0, 0.0};
	for (unsigned t = blockIdx.x; t < ntiles; t += gridDim.x) {
		const unsigned tile = xcd_tile(t, ntiles);
		const int X0 = (int) (tile % ntx) * FU_TXO, Y0 = (int) (tile / ntx) * FU_TYO;
		const int gxu = X0 - FU_H + 2 * lane, gx = wrap_once(gxu, nx);
		const int gyb = Y0 - FU_HY + FU_RPW * w;
		double pr[FU_RPW][2], xr[FU_RPW][2];
		{
			TE2 zv[FU_RPW], pv[FU_RPW];
			double2 xv[FU_RPW];
#pragma unroll
			for (int i = 0; i < FU_RPW; ++i) {
				const size_t at = (size_t) nx * wrap_once(gyb + i, ny) + gx;
				zv[i] = *(const TE2*) (a.z + at);
				if (!FIRST) pv[i] = *(const TE2*) (a.p_old + at);
				if (!XZ) xv[i] = *(const double2*) (a.x_old + at);
			}
#pragma unroll
			for (int i = 0; i < FU_RPW; ++i) {
				const double z0 = (double) zv[i].x, z1 = (double) zv[i].y;
				pr[i][0] = FIRST ? z0 : z0 + beta * (double) pv[i].x;
				pr[i][1] = FIRST ? z1 : z1 + beta * (double) pv[i].y;
				xr[i][0] = XZ ? alpha * pr[i][0] : xv[i].x + alpha * pr[i][0];
				xr[i][1] = XZ ? alpha * pr[i][1] : xv[i].y + alpha * pr[i][1];
			}
		}
		bnd_publish<double, FU_RPW>(bp, pr, w, lane);
		bnd_publish<double, FU_RPW>(bx, xr, w, lane);
		__syncthreads();
		double pu[2], pd[2], xu[2], xd[2];
#pragma unroll
		for (int o = 0; o < 2; ++o) {
			pu[o] = bnd_at(bp, w, 1, lane, o); pd[o] = bnd_at(bp, w + 2, 0, lane, o);
			xu[o] = bnd_at(bx, w, 1, lane, o); xd[o] = bnd_at(bx, w + 2, 0, lane, o);
		}
		const bool oc = lane >= FU_H / 2 && lane < 64 - FU_H / 2 && gxu < nx;
#pragma unroll
		for (int i = 0; i < FU_RPW; ++i) {
			// the lane shifts run with every lane active (a DPP read of a lane
			// masked off returns the reader's own value), the output test after
			double Lx[2], Lp[2];
#pragma unroll
			for (int o = 0; o < 2; ++o) {
				double E, W, Nn, S;
				const int px = wrap_once(gx + o, nx), py = wrap_once(gyb + i, ny);
				const bool we = px == nx - 1, ww = px == 0, wn = py == ny - 1, ws = py == 0;
				strip_nbrs<double, FU_RPW>(xr, i, o, xu[o], xd[o], E, W, Nn, S);
				Lx[o] = stencil5p2<double, true>(off, dg, xr[i][o], E, W, Nn, S, we, ww, wn, ws);
				strip_nbrs<double, FU_RPW>(pr, i, o, pu[o], pd[o], E, W, Nn, S);
				Lp[o] = stencil5p2<double, true>(off, dg, pr[i][o], E, W, Nn, S, we, ww, wn, ws);
			}
			const int ly = FU_RPW * w + i;
			if (!(oc && ly >= FU_HY && ly < FU_LY - FU_HY && gyb + i < ny)) continue;
			const size_t at = (size_t) nx * (gyb + i) + gx;
			const double2 bv = *(const double2*) (a.b + at);
			double rv[2];
#pragma unroll
			for (int o = 0; o < 2; ++o) {
				rv[o] = ((o ? bv.y : bv.x) - bmean * 1.0) - Lx[o];
				acc[0] += rv[o] * rv[o];
				acc[1] += pr[i][o] * rv[o];
				acc[2] += pr[i][o] * Lp[o];
			}
			TE2 po;
			double2 xo;
			po.x = (TE) pr[i][0]; po.y = (TE) pr[i][1];
			xo.x = xr[i][0]; xo.y = xr[i][1];
			*(TE2*) (a.p_new + at) = po;
			*(double2*) (a.x_new + at) = xo;
			TE2 ro;
			ro.x = (TE) rv[0]; ro.y = (TE) rv[1];
			*(TE2*) (a.r + at) = ro;
		}
		__syncthreads();     // the next tile rewrites the boundary rows
	}
	reduce_partial_s<3>(acc, a.gpart, FCG_BLOCKS);
}

// The same update, one cell per thread (grids below the tile size).
template<typename TE, bool FIRST, bool XZ>
__global__ __launch_bounds__(256) void k_fcg_update_1d(FcgArgs<TE> a, int nx, int ny, double off, double dg)
{
	double beta, alpha;
	if (fcg_coefficients<TE, FIRST>(a, beta, alpha)) return;
	// b is the right-hand side with its mean (k_mean); r = (b - mean) - L x
	const double bmean = a.ctl->mean;
	const size_t N = (size_t) nx * ny;
	double acc[3] = {0.0, 0.0, 0.0};
	for (size_t i = blockIdx.x * 256 + threadIdx.x; i < N; i += (size_t) gridDim.x * 256) {
		int gx, gy;
		cell_xy(i, nx, gx, gy);
		const size_t nb[5] = {i, gx == nx - 1 ? i - (nx - 1) : i + 1, gx == 0 ? i + (nx - 1) : i - 1,
		                      gy == ny - 1 ? (size_t) gx : i + nx, gy == 0 ? i + N - nx : i - nx};
		double zv[5], pv[5], xv[5];
#pragma unroll
		for (int t = 0; t < 5; ++t) zv[t] = (double) a.z[nb[t]];
		if (!FIRST) {
#pragma unroll
			for (int t = 0; t < 5; ++t) pv[t] = (double) a.p_old[nb[t]];
		}
		if (!XZ) {
#pragma unroll
			for (int t = 0; t < 5; ++t) xv[t] = a.x_old[nb[t]];
		}
		const double bi = a.b[i] - bmean * 1.0;
#pragma unroll
		for (int t = 0; t < 5; ++t) {
			const double p = FIRST ? zv[t] : zv[t] + beta * pv[t];
			pv[t] = p;
			xv[t] = XZ ? alpha * p : xv[t] + alpha * p;
		}
		const bool we = gx == nx - 1, ww = gx == 0, wn = gy == ny - 1, ws = gy == 0;
		const double Lx = stencil5p2<double, true>(off, dg, xv[0], xv[1], xv[2], xv[3], xv[4], we, ww, wn, ws);
		const double Lp = stencil5p2<double, true>(off, dg, pv[0], pv[1], pv[2], pv[3], pv[4], we, ww, wn, ws);
		const double r = bi - Lx;
		a.p_new[i] = (TE) pv[0];
		a.x_new[i] = xv[0];
		a.r[i] = (TE) r;
		acc[0] += r * r;
		acc[1] += pv[0] * r;
		acc[2] += pv[0] * Lp;
	}
	reduce_partial_s<3>(acc, a.gpart, FCG_BLOCKS);
}

// the stop test after update `it - 1` and the carried scalars, one block
// (the last iteration of a batch, before the host reads the control block)
__global__ __launch_bounds__(256) void k_fcg_fin(int it, unsigned nblk, double N, SolveCtl* __restrict__ ctl,
                                                 const double* __restrict__ gpart)
{
	if (ctl->done) return;
	double t[3];
	red_totals_256<3>(gpart, FCG_BLOCKS, nblk, t);
	if (threadIdx.x == 0) fcg_stop_write(ctl, t, it, N);
}

__global__ void k_zero(double* __restrict__ x, size_t N)
{
	const size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i < N) x[i] = 0.0;
}

// ------------------------------------------------------ residual+restrict --
// RESID: coarse[c] = sum_taps (w/64) * (b - L x)[fine]; else sum of x itself.
// With cp != nullptr the fine input is also copied to cp (nested iteration).
// Without RESID the fine input may be FP64 for an FP32 restriction (TS):
// each value is rounded once on load, and cp receives the rounded copy.
template<typename T, bool RESID, typename TS = T>
__global__ __launch_bounds__(256) void k_resid_restr(const TS* __restrict__ x, const T* __restrict__ b,
                                                     T* __restrict__ rc, T* __restrict__ cp, int nx, int ny,
                                                     T off, T dg, const int* __restrict__ done)
{
	constexpr int FX = 2 * RTX + 2, FY = 2 * RTY + 2;   // fine cells feeding the tile
	constexpr int LX = FX + 2, LY = FY + 2;             // plus the stencil halo
	__shared__ T xs[RESID ? LX * LY : 1];
	__shared__ T rs[FX * FY];
	if (done && *done) return;
	const int cnx = nx >> 1;
	const unsigned ntx = cnx / RTX;
	const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
	const int C0x = (tile % ntx) * RTX, C0y = (tile / ntx) * RTY;
	const int F0x = 2 * C0x - 1, F0y = 2 * C0y - 1;     // first fine cell of the taps
	const int tid = threadIdx.x;
	// all global loads of a thread are issued before it touches LDS
	constexpr int NXV = (LX * LY + 255) / 256, NRV = (FX * FY + 255) / 256;
	T bv[NRV];
#pragma unroll
	for (int k = 0; k < NRV; ++k) {
		const int p = tid + 256 * k;
		if (p < FX * FY) {
			int gx = F0x + p % FX, gy = F0y + p / FX;
			gx += (gx < 0) ? nx : 0; gx -= (gx >= nx) ? nx : 0;
			gy += (gy < 0) ? ny : 0; gy -= (gy >= ny) ? ny : 0;
			bv[k] = RESID ? b[gx + (size_t) nx * gy] : (T) x[gx + (size_t) nx * gy];
		}
	}
	if (RESID) {
		T xv[NXV];
#pragma unroll
		for (int k = 0; k < NXV; ++k) {
			const int p = tid + 256 * k;
			if (p < LX * LY) {
				int gx = F0x - 1 + p % LX, gy = F0y - 1 + p / LX;
				gx += (gx < 0) ? nx : 0; gx -= (gx >= nx) ? nx : 0;
				gy += (gy < 0) ? ny : 0; gy -= (gy >= ny) ? ny : 0;
				xv[k] = x[gx + (size_t) nx * gy];
			}
		}
#pragma unroll
		for (int k = 0; k < NXV; ++k)
			if (tid + 256 * k < LX * LY) xs[tid + 256 * k] = xv[k];
		__syncthreads();
	}
#pragma unroll
	for (int k = 0; k < NRV; ++k) {
		const int p = tid + 256 * k;
		if (p < FX * FY) {
			const int lx = p % FX, ly = p / FX;
			if (RESID) {
				int gx = F0x + lx, gy = F0y + ly;
				gx += (gx < 0) ? nx : 0; gx -= (gx >= nx) ? nx : 0;
				gy += (gy < 0) ? ny : 0; gy -= (gy >= ny) ? ny : 0;
				const int li = (lx + 1) + LX * (ly + 1);
				T Lx = stencil5p2<T, true>(off, dg, xs[li], xs[li + 1], xs[li - 1], xs[li + LX], xs[li - LX],
				                         gx == nx - 1, gx == 0, gy == ny - 1, gy == 0);
				rs[p] = bv[k] - Lx;
			} else {
				rs[p] = bv[k];
			}
		}
	}
	__syncthreads();
	const int cxl = tid % RTX, cyl = tid / RTX;
	T o = T(0);
#pragma unroll
	for (int i = 0; i < 16; ++i) {
		const int dx = (i % 4) - 2, dy = i / 4 - 2;
		const int wgt = (4 - abs(2 * dx + 1)) * (4 - abs(2 * dy + 1));
		o += T(wgt / 64.) * rs[(2 * cxl + dx + 2) + FX * (2 * cyl + dy + 2)];
	}
	rc[(C0x + cxl) + (size_t) cnx * (C0y + cyl)] = o;
	if (cp) {
		// each fine cell of the tile's own 2x2 blocks, copied once
		for (int p = tid; p < 4 * RTX * RTY; p += 256) {
			int fx = 2 * C0x + p % (2 * RTX), fy = 2 * C0y + p / (2 * RTX);
			cp[fx + (size_t) nx * fy] = (T) x[fx + (size_t) nx * fy];
		}
	}
}

// ------------------------------------------------------- update + residual --
// xs_new = xs + e (e may be null), r = rhs' - L xs_new, where rhs' = rhs -
// mean when `sub_mean` (written back to rhs when it is 1, not when 2).
// Deterministic RMS of r and the stop test (mg/solver.h:226-233) in the
// last block.  XS / E: xs / e given
// (template flags, so a cell's loads issue together; update_resid_launch
// picks the variant from the pointers).
template<typename TE, typename TR, bool XS, bool E>
__global__ __launch_bounds__(256) void k_update_resid(const double* __restrict__ xs, const TE* __restrict__ e,
                                                      double* __restrict__ rhs, double* __restrict__ xs_out,
                                                      TR* __restrict__ res, int nx, int ny, double off, double dg,
                                                      int sub_mean, int first, SolveCtl* __restrict__ ctl,
                                                      double* __restrict__ partials)
{
	if (!first && ctl->done) return;
	const size_t N = (size_t) nx * ny;
	const double mean = ctl->mean;
	// acc[1]: sum of (rhs - mean)^2 for a relative tolerance (with sub_mean:
	// the same terms in the same order as a separate pass over rhs would add)
	double acc[2] = {0.0, 0.0};
	// 32-bit cell indices (fields have fewer than 2^32 cells)
	for (unsigned i = blockIdx.x * 256 + threadIdx.x; i < (unsigned) N; i += gridDim.x * 256) {
		const int gx = (int) (i & (unsigned) (nx - 1)), gy = (int) (i >> __builtin_ctz((unsigned) nx));
		const unsigned ie = gx == nx - 1 ? i - (nx - 1) : i + 1;
		const unsigned iw = gx == 0 ? i + (nx - 1) : i - 1;
		const unsigned in_ = gy == ny - 1 ? (unsigned) gx : i + nx;
		const unsigned is = gy == 0 ? (unsigned) (i + N - nx) : i - nx;
		// xs == nullptr: the iterate is zero and e (below) is the whole start
		double c = 0.0, Ev = 0.0, Wv = 0.0, Nv = 0.0, Sv = 0.0;
		if (XS) { c = xs[i]; Ev = xs[ie]; Wv = xs[iw]; Nv = xs[in_]; Sv = xs[is]; }
		if (E) {
			c = c + (double) e[i]; Ev = Ev + (double) e[ie]; Wv = Wv + (double) e[iw];
			Nv = Nv + (double) e[in_]; Sv = Sv + (double) e[is];
			xs_out[i] = c;
		}
		double r0 = rhs[i];
		if (sub_mean) {
			r0 = r0 - mean * 1.0;
			if (sub_mean == 1) rhs[i] = r0;
			acc[1] += r0 * r0;
		}
		double Lx = stencil5p2<double, true>(off, dg, c, Ev, Wv, Nv, Sv, gx == nx - 1, gx == 0, gy == ny - 1, gy == 0);
		double r = r0 - Lx;
		res[i] = (TR) r;
		acc[0] += r * r;
	}
	if (sub_mean) reduce_partial<2>(acc, partials);
	else {
		double a1[1] = {acc[0]};
		reduce_partial<1>(a1, partials);
	}
}

// k_update_resid on cell pairs along x (nx even): thread k takes cells 2k
// and 2k + 1 of a row, with 16-byte loads of the pair in rows z - 1, z,
// z + 1 and single loads of its two outer neighbours, each value read
// once, and 16-byte stores.  Every cell takes the arithmetic of
// k_update_resid (the same bits of xs_out, rhs and res); the partial sums
// add the pair's terms in cell order.
template<typename TE, typename TR, bool XS, bool E>
__global__ __launch_bounds__(256) void k_update_resid_pairs(const double* __restrict__ xs, const TE* __restrict__ e,
                                                            double* __restrict__ rhs, double* __restrict__ xs_out,
                                                            TR* __restrict__ res, int nx, int ny, double off,
                                                            double dg, int sub_mean, int first,
                                                            SolveCtl* __restrict__ ctl, double* __restrict__ partials)
{
	using TE2 = typename Pair<TE>::type;
	using TR2 = typename Pair<TR>::type;
	if (!first && ctl->done) return;
	const unsigned N = (unsigned) nx * (unsigned) ny, unx = (unsigned) nx, P = N / 2;
	const double mean = ctl->mean;
	double acc[2] = {0.0, 0.0};
	// consecutive blocks per XCD (26.9 -> 25.3 us at 2048^2)
	for (unsigned k = xcd_remap(blockIdx.x, gridDim.x) * 256 + threadIdx.x; k < P; k += gridDim.x * 256) {
		const unsigned i = 2 * k;
		const int x = (int) (i & (unx - 1)), z = (int) (i >> __builtin_ctz(unx));
		const unsigned rz = i - (unsigned) x;
		const unsigned rm = z == 0 ? rz + (N - unx) : rz - unx;
		const unsigned rp = z == ny - 1 ? rz - (N - unx) : rz + unx;
		const unsigned xm = x == 0 ? unx - 1 : (unsigned) x - 1;
		const unsigned x2 = x + 2 == nx ? 0u : (unsigned) x + 2;
		// per cell o: centre, E, W, N (z + 1), S (z - 1)
		double c[2] = {0.0, 0.0}, Ev[2] = {0.0, 0.0}, Wv[2] = {0.0, 0.0}, Nv[2] = {0.0, 0.0}, Sv[2] = {0.0, 0.0};
		if (XS) {
			const double2 X = *(const double2*) (xs + i), Xm = *(const double2*) (xs + rm + x),
			              Xp = *(const double2*) (xs + rp + x);
			const double xw = xs[rz + xm], xe = xs[rz + x2];
			c[0] = X.x; Ev[0] = X.y; Wv[0] = xw; Nv[0] = Xp.x; Sv[0] = Xm.x;
			c[1] = X.y; Ev[1] = xe; Wv[1] = X.x; Nv[1] = Xp.y; Sv[1] = Xm.y;
		}
		if (E) {
			const TE2 Y = *(const TE2*) (e + i), Ym = *(const TE2*) (e + rm + x), Yp = *(const TE2*) (e + rp + x);
			const double yw = (double) e[rz + xm], ye = (double) e[rz + x2];
			c[0] = c[0] + (double) Y.x; Ev[0] = Ev[0] + (double) Y.y; Wv[0] = Wv[0] + yw;
			Nv[0] = Nv[0] + (double) Yp.x; Sv[0] = Sv[0] + (double) Ym.x;
			c[1] = c[1] + (double) Y.y; Ev[1] = Ev[1] + ye; Wv[1] = Wv[1] + (double) Y.x;
			Nv[1] = Nv[1] + (double) Yp.y; Sv[1] = Sv[1] + (double) Ym.y;
			*(double2*) (xs_out + i) = make_double2(c[0], c[1]);
		}
		const double2 B = *(const double2*) (rhs + i);
		double r0[2] = {B.x, B.y};
		if (sub_mean) {
			r0[0] = r0[0] - mean * 1.0;
			r0[1] = r0[1] - mean * 1.0;
			if (sub_mean == 1) *(double2*) (rhs + i) = make_double2(r0[0], r0[1]);
			acc[1] += r0[0] * r0[0];
			acc[1] += r0[1] * r0[1];
		}
		const bool wn = z == ny - 1, ws = z == 0;
		const double Lx0 = stencil5p2<double, true>(off, dg, c[0], Ev[0], Wv[0], Nv[0], Sv[0], false, x == 0, wn, ws);
		const double Lx1 = stencil5p2<double, true>(off, dg, c[1], Ev[1], Wv[1], Nv[1], Sv[1], x + 1 == nx - 1, false,
		                                            wn, ws);
		const double r[2] = {r0[0] - Lx0, r0[1] - Lx1};
		TR2 ro;
		ro.x = (TR) r[0]; ro.y = (TR) r[1];
		*(TR2*) (res + i) = ro;
		acc[0] += r[0] * r[0];
		acc[0] += r[1] * r[1];
	}
	if (sub_mean) reduce_partial<2>(acc, partials);
	else {
		double a1[1] = {acc[0]};
		reduce_partial<1>(a1, partials);
	}
}

template<typename TE, typename TR>
static void update_resid_launch(Ctx* c, unsigned grid, const double* xs, const TE* e, double* rhs, double* xs_out,
                                TR* res, double off, double dg, int sub_mean, int first, SolveCtl* ctl,
                                double* partials)
{
	const int nx = (int) c->nx, ny = (int) c->ny;
	// cell pairs along x when the rows have even length (every power of two
	// from 2)
	const bool pairs = nx % 2 == 0;
#define AFM_UR(X_, E_)                                                                                           \
	do {                                                                                                         \
		if (pairs)                                                                                               \
			hipLaunchKernelGGL((k_update_resid_pairs<TE, TR, X_, E_>), dim3(grid), dim3(256), 0, c->stream, xs, e, \
			                   rhs, xs_out, res, nx, ny, off, dg, sub_mean, first, ctl, partials);                \
		else                                                                                                     \
			hipLaunchKernelGGL((k_update_resid<TE, TR, X_, E_>), dim3(grid), dim3(256), 0, c->stream, xs, e, rhs,  \
			                   xs_out, res, nx, ny, off, dg, sub_mean, first, ctl, partials);                     \
	} while (0)
	if (xs && e) AFM_UR(true, true);
	else if (xs) AFM_UR(true, false);
	else if (e) AFM_UR(false, true);
	else AFM_UR(false, false);
#undef AFM_UR
}

// RMS of the residual and the stop test (mg/solver.h:228), one block.
// first == 2: the first call with a relative tolerance, which is rel times
// the RMS of rhs - mean summed by k_update_resid (partials[MAX_RED_BLOCKS..])
__global__ __launch_bounds__(256) void k_update_fin(int first, unsigned nblk, size_t N, SolveCtl* __restrict__ ctl,
                                                    const double* __restrict__ partials, double rel = 0.0)
{
	if (!first && ctl->done) return;
	double tot[2];
	if (first == 2) red_total<2>(partials, nblk, tot);
	else {
		double t1[1];
		red_total<1>(partials, nblk, t1);
		tot[0] = t1[0];
	}
	if (threadIdx.x == 0) {
		if (first == 2) ctl->tol = rel * (sqrt(tot[1]) / sqrt((double) N));
		double rms = sqrt(tot[0]) / sqrt((double) N);
		int count = first ? 0 : ctl->count + 1;
		ctl->count = count;
		ctl->res = rms;
		ctl->done = !(rms > ctl->tol && count < ctl->max_iter);
	}
}

// sum(rhs * 1) -> mean (pass 0); RMS(rhs - mean) -> relative tol (pass 1)
__global__ __launch_bounds__(256) void k_mean(const double* __restrict__ rhs, size_t N, int pass,
                                              SolveCtl* __restrict__ ctl, double* __restrict__ partials,
                                              double tol = 0.0, int max_iter = 0)
{
	// pass 0 also starts the solve's control block (nothing in this launch
	// reads it)
	if (pass == 0 && blockIdx.x == 0 && threadIdx.x == 0) {
		ctl->tol = tol;
		ctl->max_iter = max_iter;
		ctl->count = 0;
		ctl->done = 0;
		ctl->err = 0;
		ctl->res = 0.0;
	}
	double acc[1] = {0.0};
	const double mean = pass ? ctl->mean : 0.0;
	for (size_t i = blockIdx.x * 256 + threadIdx.x; i < N; i += (size_t) gridDim.x * 256) {
		if (pass == 0) acc[0] += rhs[i] * 1.0;
		else { double t = rhs[i] - mean * 1.0; acc[0] += t * t; }
	}
	reduce_partial<1>(acc, partials);
}

__global__ __launch_bounds__(256) void k_mean_fin(int pass, double rel, unsigned nblk, size_t N,
                                                  SolveCtl* __restrict__ ctl, const double* __restrict__ partials)
{
	double tot[1];
	red_total<1>(partials, nblk, tot);
	if (threadIdx.x == 0) {
		if (pass == 0) ctl->mean = tot[0] / (double) N;
		else ctl->tol = rel * (sqrt(tot[0]) / sqrt((double) N));
	}
}

template<typename T>
__global__ void k_copy_cvt(const T* __restrict__ src, double* __restrict__ dst, size_t N)
{
	size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i < N) dst[i] = (double) src[i];
}

// ------------------------------------------------------------- coarse -----
// Levels 0..Lt (n <= 64) of the unit square in one workgroup, all in LDS.
// Levels 6 and 5 run on all CTHREADS threads with block barriers; levels
// <= 4 run in wave 0 alone, where LDS operations complete in issue order and
// no block barrier is needed.  Levels are template parameters, so every loop
// bound is a constant and nothing spills.
//
// Per-point arithmetic is the global kernels': the 9-slot stencil order for
// n >= 4, the reference's folded 2x2 rows at n = 2 (mg/laplacian.h:62-95
// with nx = nz = 2), and "/ diagonal" as "* (1/diagonal)" (the diagonal is
// -4 n^2, a power of two).  "x += P e" is fused into the first red half-sweep:
// red updates read black neighbours as x + P e, black updates never read
// their own old value.
struct CoarseLDS {
	double X[5461], B[5461], TMP[4096];   // levels 0..6 stacked; TMP: one residual
	const double* op3;                     // level 3's V-cycle as a matrix, or nullptr
};

constexpr int lvl_off(int l) { return ((1 << (2 * l)) - 1) / 3; }   // sum_{j<l} 4^j

template<int L> struct CLv { static constexpr int NT = L >= 5 ? CTHREADS : 64; };

// run a wave-0 level from code that every thread executes
#define C_WAVE0(call)                  \
	do {                               \
		if (threadIdx.x < 64) { call; } \
		__syncthreads();               \
	} while (0)

template<int NT>
__device__ inline void csync()
{
	if constexpr (NT > 64) {
		__syncthreads();
	} else {
		// one wave: order LDS traffic without a block barrier
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
	}
}

// A cell's four stencil neighbours in stencil5's summation order (no
// diagonal term): the wrap flags of a cell only fix the order in which its
// neighbours are added -- the vertical terms of slots 0-1, the horizontal
// pair, the vertical terms of slots 6-7 -- so the order is found with a few
// integer selects instead of stencil5's eight conditional double adds.
struct CellOps {
	int i, a0, a1, a2, a3;
};

template<int L>
__device__ inline CellOps cell_ops(int gx, int gy)
{
	constexpr int nx = 1 << L, msk = nx - 1;
	const bool we = gx == msk, ww = gx == 0, wn = gy == msk, ws = gy == 0;
	const int iE = ((gx + 1) & msk) + (gy << L), iW = ((gx - 1) & msk) + (gy << L);
	const int iN = gx + (((gy + 1) & msk) << L), iS = gx + (((gy - 1) & msk) << L);
	// horizontal pair: E before W at either x edge, W before E inside
	const int h0 = (we || ww) ? iE : iW, h1 = (we || ww) ? iW : iE;
	CellOps o;
	o.i = gx + (gy << L);
	// y = ny-1: N, S first; y = 0: N, S last; else S first, N last
	if (wn) { o.a0 = iN; o.a1 = iS; o.a2 = h0; o.a3 = h1; }
	else if (ws) { o.a0 = h0; o.a1 = h1; o.a2 = iN; o.a3 = iS; }
	else { o.a0 = iS; o.a1 = h0; o.a2 = h1; o.a3 = iN; }
	return o;
}

// red-black sweeps at level L; zero: x starts at 0; xc: x += P xc first
template<int L>
__device__ void c_smooth(double* x, const double* b, bool zero, const double* xc, int sweeps)
{
	constexpr int NT = CLv<L>::NT;
	if constexpr (L == 1) {
		// 2x2 with the folded rows: (x, y) couples to xo = (1-x, y) and
		// yo = (x, 1-y), weight 2 off = 8, diagonal -16; the stored order puts
		// xo first in row y = 0 and yo first in row y = 1.
		const int i = threadIdx.x;
		for (int s = 0; s < sweeps; ++s)
			for (int color = 0; color < 2; ++color) {
				const bool first = s == 0 && color == 0;
				if (i < 4 && (((i & 1) + (i >> 1)) & 1) == color) {
					double value;
					if (first && zero) value = 0.0;
					else {
						double xo = x[i ^ 1], yo = x[i ^ 2];
						if (first && xc) {
							xo = xo + prolong_at(xc, (i ^ 1) & 1, (i ^ 1) >> 1, 1, 1);
							yo = yo + prolong_at(xc, (i ^ 2) & 1, (i ^ 2) >> 1, 1, 1);
						}
						value = 0.0;
						if ((i >> 1) == 0) { value += 8.0 * xo; value += 8.0 * yo; }
						else { value += 8.0 * yo; value += 8.0 * xo; }
					}
					x[i] = (b[i] - value) * (1.0 / -16.0);
				}
				csync<NT>();
			}
	} else {
		constexpr int nx = 1 << L, msk = nx - 1, hx = nx >> 1, N2 = (nx * nx) >> 1;
		constexpr double off = double(nx * nx), dg = -4.0 * off, inv = 1.0 / dg;
		for (int s = 0; s < sweeps; ++s)
			for (int color = 0; color < 2; ++color) {
				const bool first = s == 0 && color == 0;
				if (L >= 5 && !(first && (zero || xc))) {
					// the issue-bound levels (all threads, several cells each)
#pragma unroll
					for (int k0 = 0; k0 < N2; k0 += NT) {
						const int k = k0 + (int) threadIdx.x;
						if (N2 % NT != 0 && k >= N2) break;
						const int gy = k >> (L - 1), gx = ((k & (hx - 1)) << 1) + ((color + gy) & 1);
						const CellOps o = cell_ops<L>(gx, gy);
						double value = 0.0;
						value = fma_p2(off, x[o.a0], value);   // off = n^2: exact products
						value = fma_p2(off, x[o.a1], value);
						value = fma_p2(off, x[o.a2], value);
						value = fma_p2(off, x[o.a3], value);
						x[o.i] = (b[o.i] - value) * inv;
					}
					csync<NT>();
					continue;
				}
#pragma unroll
				for (int k0 = 0; k0 < N2; k0 += NT) {
					const int k = k0 + (int) threadIdx.x;
					if (N2 % NT != 0 && k >= N2) break;
					const int gy = k >> (L - 1), gx = ((k & (hx - 1)) << 1) + ((color + gy) & 1);
					const int i = gx + (gy << L);
					double value;
					if (first && zero) value = 0.0;
					else {
						const int ex = (gx + 1) & msk, wx = (gx - 1) & msk, ny_ = (gy + 1) & msk, sy = (gy - 1) & msk;
						double E = x[ex + (gy << L)], Wv = x[wx + (gy << L)], Nv = x[gx + (ny_ << L)], Sv = x[gx + (sy << L)];
						if (first && xc) {
							E = E + prolong_at(xc, ex, gy, hx, hx);
							Wv = Wv + prolong_at(xc, wx, gy, hx, hx);
							Nv = Nv + prolong_at(xc, gx, ny_, hx, hx);
							Sv = Sv + prolong_at(xc, gx, sy, hx, hx);
						}
						value = stencil5p2<double, false>(off, dg, 0.0, E, Wv, Nv, Sv, gx == msk, gx == 0, gy == msk, gy == 0);
					}
					x[i] = (b[i] - value) * inv;
				}
				csync<NT>();
			}
	}
}

// dst(level L-1) = R src, src = b - L x when x != nullptr (mg/solver.h:153-181)
template<int L>
__device__ void c_restrict(CoarseLDS& S, const double* x, const double* src, double* dst)
{
	constexpr int NT = CLv<L>::NT;
	constexpr int nx = 1 << L, N = nx * nx, cn = nx >> 1, msk = nx - 1;
	constexpr double off = double(nx * nx);
	const double* in = src;
	if (x) {
#pragma unroll
		for (int i0 = 0; i0 < N; i0 += NT) {
			const int i = i0 + (int) threadIdx.x;
			if (N % NT != 0 && i >= N) break;
			const int gx = i & msk, gy = i >> L;
			double y;
			if constexpr (L >= 2) {
				const int ex = (gx + 1) & msk, wx = (gx - 1) & msk, ny_ = (gy + 1) & msk, sy = (gy - 1) & msk;
				y = stencil5p2<double, true>(off, -4.0 * off, x[i], x[ex + (gy << L)], x[wx + (gy << L)],
				                           x[gx + (ny_ << L)], x[gx + (sy << L)], gx == msk, gx == 0, gy == msk, gy == 0);
			} else {
				// folded 2x2 rows in stored order: (0,0) d,xo,yo  (1,0) xo,d,yo
				// (0,1) yo,d,xo  (1,1) yo,xo,d
				const double c = x[i], xo = x[i ^ 1], yo = x[i ^ 2];
				y = 0.0;
				if (i == 0) { y += -16.0 * c; y += 8.0 * xo; y += 8.0 * yo; }
				else if (i == 1) { y += 8.0 * xo; y += -16.0 * c; y += 8.0 * yo; }
				else if (i == 2) { y += 8.0 * yo; y += -16.0 * c; y += 8.0 * xo; }
				else { y += 8.0 * yo; y += 8.0 * xo; y += -16.0 * c; }
			}
			S.TMP[i] = src[i] - y;
		}
		csync<NT>();
		in = S.TMP;
	}
#pragma unroll
	for (int c0 = 0; c0 < cn * cn; c0 += NT) {
		const int c = c0 + (int) threadIdx.x;
		if ((cn * cn) % NT != 0 && c >= cn * cn) break;
		const int xcc = c & (cn - 1), zcc = c >> (L - 1);
		const int xcf = 2 * xcc + 1, zcf = 2 * zcc + 1;
		double o = 0.0;
#pragma unroll
		for (int t = 0; t < 16; ++t) {
			const int dx = (t % 4) - 2, dz = t / 4 - 2;
			const int xx = (xcf + dx + nx) & msk, zz = (zcf + dz + nx) & msk;
			const int w = (4 - abs(2 * dx + 1)) * (4 - abs(2 * dz + 1));
			o += w / 64. * in[xx + (zz << L)];
		}
		dst[c] = o;
	}
	csync<NT>();
}

// dst(level L) (+)= P xc(level L-1)
template<int L>
__device__ void c_prolong(const double* xc, double* dst, bool add)
{
	constexpr int NT = CLv<L>::NT;
	constexpr int nx = 1 << L, N = nx * nx, h = nx >> 1;
#pragma unroll
	for (int i0 = 0; i0 < N; i0 += NT) {
		const int i = i0 + (int) threadIdx.x;
		if (N % NT != 0 && i >= N) break;
		double v = prolong_at(xc, i & (nx - 1), i >> L, h, h);
		dst[i] = add ? dst[i] + v : v;
	}
	csync<NT>();
}

// mg/solver.h:264-271 on the 1x1 coarsest grid: mean removal leaves 0 and
// CG stops at once (the 1x1 laplacian has no stored entries).
__device__ inline void c_solve0(CoarseLDS& S)
{
	if (threadIdx.x == 0) {
		double c = (1.0 * S.B[0]) / 1u;
		S.B[0] -= 1.0 * c;
		S.X[0] = S.B[0];
	}
	csync<64>();
}

// X_3 = M B_3 with M the V-cycle of levels 3..0 (linear in B_3: smoothing
// from zero, restriction, the coarsest solve's mean removal and prolongation
// are), built from unit vectors by k_coarse_op3 with the same arithmetic;
// only the order of the sums differs.  Wave 0: lane i forms row i, reading
// column j of all rows as one contiguous 512-byte load.
__device__ inline void c_apply_op3(CoarseLDS& S)
{
	const int i = (int) threadIdx.x;
	const double* b = S.B + lvl_off(3);
	double acc = 0.0;
#pragma unroll 16
	for (int j = 0; j < 64; ++j) acc = fma(S.op3[j * 64 + i], b[j], acc);
	S.X[lvl_off(3) + i] = acc;
	csync<64>();
}

// recurse(L, B_L) -> X_L (mg/solver.h:199-216).  L <= 4: wave 0 only;
// L >= 5: every thread.
template<int L>
__device__ void c_recurse(CoarseLDS& S, int nu1, int nu2)
{
	if constexpr (L == 3) {
		if (S.op3) {
			c_apply_op3(S);
			return;
		}
	}
	if constexpr (L == 0) {
		c_solve0(S);
	} else {
		c_smooth<L>(S.X + lvl_off(L), S.B + lvl_off(L), true, nullptr, nu1);
		c_restrict<L>(S, S.X + lvl_off(L), S.B + lvl_off(L), S.B + lvl_off(L - 1));
		if constexpr (L - 1 >= 5) c_recurse<L - 1>(S, nu1, nu2);
		else if constexpr (L == 5) C_WAVE0(c_recurse<L - 1>(S, nu1, nu2));
		else c_recurse<L - 1>(S, nu1, nu2);
		if constexpr (L >= 2) {
			c_smooth<L>(S.X + lvl_off(L), S.B + lvl_off(L), false, S.X + lvl_off(L - 1), nu2);
		} else {
			c_prolong<1>(S.X, S.X + lvl_off(1), true);
			c_smooth<1>(S.X + lvl_off(1), S.B + lvl_off(1), false, nullptr, nu2);
		}
	}
}

// level dispatch for wave 0 (levels 0..4)
__device__ inline void c_recurse_w0(CoarseLDS& S, int l, int nu1, int nu2)
{
	switch (l) {
	case 0: c_recurse<0>(S, nu1, nu2); break;
	case 1: c_recurse<1>(S, nu1, nu2); break;
	case 2: c_recurse<2>(S, nu1, nu2); break;
	case 3: c_recurse<3>(S, nu1, nu2); break;
	default: c_recurse<4>(S, nu1, nu2); break;
	}
}

// recurse(l) called by every thread
__device__ inline void c_recurse_all(CoarseLDS& S, int l, int nu1, int nu2)
{
	if (l >= 6) c_recurse<6>(S, nu1, nu2);
	else if (l == 5) c_recurse<5>(S, nu1, nu2);
	else C_WAVE0(c_recurse_w0(S, l, nu1, nu2));
}

// Nested-iteration steps at level L (mg/solver.h:183-197).  Down: B_{L-1} =
// R B_L.  Up: recurse(L-1); X_L = P X_{L-1}; smooth X_L against B_L (still
// the restricted right-hand side); B_L = X_L, the nested result that the
// next level's recursion takes as its right-hand side.
template<int L>
__device__ void c_nested_down(CoarseLDS& S)
{
	c_restrict<L>(S, nullptr, S.B + lvl_off(L), S.B + lvl_off(L - 1));
}

template<int L>
__device__ void c_nested_up(CoarseLDS& S, int nu1, int nu2)
{
	constexpr int NT = CLv<L>::NT;
	if constexpr (L - 1 >= 5) c_recurse<L - 1>(S, nu1, nu2);
	else if constexpr (L == 5) C_WAVE0(c_recurse<L - 1>(S, nu1, nu2));
	else c_recurse<L - 1>(S, nu1, nu2);
	c_prolong<L>(S.X + lvl_off(L - 1), S.X + lvl_off(L), false);
	c_smooth<L>(S.X + lvl_off(L), S.B + lvl_off(L), false, nullptr, nu2);
	for (int i = threadIdx.x; i < (1 << (2 * L)); i += NT) S.B[lvl_off(L) + i] = S.X[lvl_off(L) + i];
	csync<NT>();
}

// the whole nested iteration over levels 1..Lt (called by every thread)
__device__ void c_nested(CoarseLDS& S, int Lt, int nu1, int nu2)
{
	if (Lt >= 6) c_nested_down<6>(S);
	if (Lt >= 5) c_nested_down<5>(S);
	if (threadIdx.x < 64) {
		for (int k = Lt < 4 ? Lt : 4; k >= 1; --k) {
			switch (k) {
			case 1: c_nested_down<1>(S); break;
			case 2: c_nested_down<2>(S); break;
			case 3: c_nested_down<3>(S); break;
			default: c_nested_down<4>(S); break;
			}
		}
		for (int k = 1; k <= (Lt < 4 ? Lt : 4); ++k) {
			switch (k) {
			case 1: c_nested_up<1>(S, nu1, nu2); break;
			case 2: c_nested_up<2>(S, nu1, nu2); break;
			case 3: c_nested_up<3>(S, nu1, nu2); break;
			default: c_nested_up<4>(S, nu1, nu2); break;
			}
		}
	}
	__syncthreads();
	if (Lt >= 5) c_nested_up<5>(S, nu1, nu2);
	if (Lt >= 6) c_nested_up<6>(S, nu1, nu2);
}

// mode 0: recurse(Lt, in) -> out;  1: nested(Lt, in); recurse -> out;
// mode 2: nested(Lt, in) -> out (the approximate solution itself).
template<typename TI>
__global__ __launch_bounds__(CTHREADS) void k_coarse(int mode, int Lt, const TI* __restrict__ in, TI* __restrict__ out,
                                                     int nu1, int nu2, const int* __restrict__ done,
                                                     const double* __restrict__ op3)
{
	__shared__ CoarseLDS S;
	if (done && *done) return;
	if (threadIdx.x == 0) S.op3 = op3;
	const int N = 1 << (2 * Lt), o = lvl_off(Lt);
	for (int i = threadIdx.x; i < N; i += CTHREADS) S.B[o + i] = (double) in[i];
	__syncthreads();
	if (mode >= 1) c_nested(S, Lt, nu1, nu2);
	if (mode <= 1) {
		c_recurse_all(S, Lt, nu1, nu2);
		for (int i = threadIdx.x; i < N; i += CTHREADS) out[i] = (TI) S.X[o + i];
	} else {
		for (int i = threadIdx.x; i < N; i += CTHREADS) out[i] = (TI) S.B[o + i];
	}
}

// Column j of level 3's V-cycle matrix: recurse(3) on the unit vector e_j,
// by the recursion c_apply_op3 stands in for.  One 64-thread block per column.
__global__ __launch_bounds__(64) void k_coarse_op3(int nu1, int nu2, double* __restrict__ op3)
{
	__shared__ CoarseLDS S;
	const int i = (int) threadIdx.x, j = (int) blockIdx.x;
	if (i == 0) S.op3 = nullptr;
	S.B[lvl_off(3) + i] = i == j ? 1.0 : 0.0;
	csync<64>();
	c_recurse<3>(S, nu1, nu2);
	op3[j * 64 + i] = S.X[lvl_off(3) + i];
}

// Level 6 of the FP32 cycles (V(2, 2)) as one 4096 x 4096 matrix: the
// recursion of levels 6..0 (k_coarse, mode 0) is a fixed linear map of the
// level's right-hand side, the same for every grid of 64^2 cells or more
// (the level's spacing is fixed), so it is formed once per context and
// applied as a matrix-vector product by the whole chip instead of by one
// workgroup.  Column j: recurse(6) on e_j, with level 3's matrix as the
// cycles use it (one block per column, FP32, by rows: m[i * 4096 + j] =
// M[i][j]); then each row is stored in FP16 scaled by a power of two that
// puts its largest entry in [2^14, 2^15) (33.5 MB: half the bytes per
// product, and half the Infinity Cache the finest level's vectors share
// with it).  Rounding the entries to 11 bits perturbs the coarse correction
// by about 5e-4 relative, far below what the V-cycle leaves (its
// convergence factor is about 0.03).
constexpr int C6N = 1 << (2 * CMAXL);    // level 6's cells

__global__ __launch_bounds__(CTHREADS) void k_coarse_op6(int nu1, int nu2, const double* __restrict__ op3,
                                                         float* __restrict__ m)
{
	__shared__ CoarseLDS S;
	const int j = (int) blockIdx.x, o = lvl_off(CMAXL);
	if (threadIdx.x == 0) S.op3 = op3;
	for (int i = threadIdx.x; i < C6N; i += CTHREADS) S.B[o + i] = i == j ? 1.0 : 0.0;
	__syncthreads();
	c_recurse<CMAXL>(S, nu1, nu2);
	for (int i = threadIdx.x; i < C6N; i += CTHREADS) m[(size_t) i * C6N + j] = (float) S.X[o + i];
}

// Row i of m to FP16 times 2^e_i (largest magnitude to [2^14, 2^15); a zero
// row keeps e_i = 0); inv[i] = 2^-e_i.  One 256-thread block per row.
__global__ __launch_bounds__(256) void k_coarse_op6_pack(const float* __restrict__ m, _Float16* __restrict__ op6,
                                                         float* __restrict__ inv)
{
	__shared__ float smax[256];
	const size_t row = (size_t) blockIdx.x * C6N;
	float mx = 0.f;
	for (int j = threadIdx.x; j < C6N; j += 256) mx = fmaxf(mx, fabsf(m[row + j]));
	smax[threadIdx.x] = mx;
	__syncthreads();
	for (int h = 128; h >= 1; h >>= 1) {
		if ((int) threadIdx.x < h) smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + h]);
		__syncthreads();
	}
	int e = 0;
	if (smax[0] > 0.f) {
		int ex;
		(void) frexpf(smax[0], &ex);     // smax = f 2^ex, f in [0.5, 1)
		e = 15 - ex;                      // smax 2^e in [2^14, 2^15)
	}
	const float sc = ldexpf(1.f, e);
	for (int j = threadIdx.x; j < C6N; j += 256) op6[row + j] = (_Float16) (m[row + j] * sc);
	if (threadIdx.x == 0) inv[blockIdx.x] = ldexpf(1.f, -e);
}

// out = M6 in: one wave per row (16 rows per 1024-thread block), the row
// read with 16-byte loads (eight FP16 entries per lane and load, 1 KB per
// wave), in staged in LDS.  Each lane sums its 64 products (FP32) in order,
// the wave adds the lanes' sums in a fixed butterfly and scales by the row's
// 2^-e: the same bits on every call.
typedef _Float16 afm_half8 __attribute__((ext_vector_type(8)));

__global__ __launch_bounds__(1024) void k_coarse_mv6(const _Float16* __restrict__ op6, const float* __restrict__ inv,
                                                     const float* __restrict__ in, float* __restrict__ out,
                                                     const int* __restrict__ done)
{
	__shared__ float4 sb[C6N / 4];
	if (done && *done) return;
	sb[threadIdx.x] = ((const float4*) in)[threadIdx.x];
	__syncthreads();
	const int lane = threadIdx.x & 63, row = blockIdx.x * 16 + (threadIdx.x >> 6);
	const afm_half8* __restrict__ m = (const afm_half8*) (op6 + (size_t) row * C6N);
	afm_half8 v[8];
#pragma unroll
	for (int k = 0; k < 8; ++k) v[k] = m[k * 64 + lane];
	float acc = 0.f;
#pragma unroll
	for (int k = 0; k < 8; ++k) {
		const float4 b0 = sb[(k * 64 + lane) * 2], b1 = sb[(k * 64 + lane) * 2 + 1];
		acc = __builtin_fmaf((float) v[k][0], b0.x, acc);
		acc = __builtin_fmaf((float) v[k][1], b0.y, acc);
		acc = __builtin_fmaf((float) v[k][2], b0.z, acc);
		acc = __builtin_fmaf((float) v[k][3], b0.w, acc);
		acc = __builtin_fmaf((float) v[k][4], b1.x, acc);
		acc = __builtin_fmaf((float) v[k][5], b1.y, acc);
		acc = __builtin_fmaf((float) v[k][6], b1.z, acc);
		acc = __builtin_fmaf((float) v[k][7], b1.w, acc);
	}
#pragma unroll
	for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, 64);
	if (lane == 0) out[row] = acc * inv[row];
}
static_assert(C6N == 4096, "k_coarse_mv6 maps 1024 threads x 4 values onto level 6");

// ------------------------------------------------------------- host side --
static inline unsigned red_blocks(size_t N)
{
	size_t b = (N + 255) / 256;
	return (unsigned) (b < (size_t) MAX_RED_BLOCKS ? b : (size_t) MAX_RED_BLOCKS);
}

// sweeps red-black sweeps at level l, x0 given by `mode`; with rc != nullptr
// also rc = R (b - L x) on level l-1 (mode must be M_ZERO then).
template<typename T>
static int smooth_launch(Ctx* c, unsigned l, int mode, int sweeps, const T* xin, const T* b, const T* xc, T* xout,
                         const int* done, T* rc = nullptr, const MGFold* fold = nullptr)
{
	const MGFold fd = fold ? *fold : MGFold{nullptr, nullptr, 0, 0, 0.0};
	const LevelBuf& L = c->lv[l];
	const double n2 = (double) L.n * (double) L.n;
	const T off = (T) n2, dg = (T) (-6.0 * n2 + n2 + n2), inv = T(1) / dg;   // dg = -4 n^2, a power of two
	const int nx = (int) L.nx, ny = (int) L.ny;
#define AFM_SM2(S_, M_, E_, SH_)                                                                               \
	hipLaunchKernelGGL((k_smooth<T, S_, M_, E_, SH_>), dim3(smooth_blocks<S_, E_, SH_>(nx, ny)),              \
	                   dim3(64 * SmoothShape<S_, SH_>::NW), 0, c->stream, xin, b, xc, xout, rc, nx, ny, off, dg, inv, done,  \
	                   fd)
#define AFM_SM(S_, M_, E_)                          \
	do {                                            \
		if (small) AFM_SM2(S_, M_, E_, 1);          \
		else AFM_SM2(S_, M_, E_, 0);                \
	} while (0)
#define AFM_SMP(S_)                                         \
	do {                                                    \
		if (small) AFM_SM2(S_, M_PADD, EPI_STORE, 1);       \
		else if (mid) AFM_SM2(S_, M_PADD, EPI_STORE, 5);    \
		else AFM_SM2(S_, M_PADD, EPI_STORE, 0);             \
	} while (0)
#define AFM_SMN(S_)                                         \
	do {                                                    \
		if (small) AFM_SM2(S_, M_NORMAL, EPI_STORE, 1);     \
		else AFM_SM2(S_, M_NORMAL, EPI_STORE, 2);           \
	} while (0)
	const bool small = (size_t) nx * ny <= SMALL_CELLS, mid = !small && (size_t) nx * ny <= MID_CELLS;
	if (rc) {
		if (mode != M_ZERO) { set_error("fused restriction needs a zero start"); return -1; }
		if (sweeps == 1) AFM_SM(1, M_ZERO, EPI_RESTRICT);
		else if (sweeps == 2) AFM_SM(2, M_ZERO, EPI_RESTRICT);
		else { set_error("smoother supports 1 or 2 sweeps per launch"); return -1; }
	} else if (sweeps == 1) {
		switch (mode) { case M_NORMAL: AFM_SMN(1); break; case M_ZERO: AFM_SM(1, M_ZERO, EPI_STORE); break;
		                case M_PADD: AFM_SMP(1); break; default: AFM_SM(1, M_PSET, EPI_STORE); }
	} else if (sweeps == 2) {
		switch (mode) { case M_NORMAL: AFM_SMN(2); break; case M_ZERO: AFM_SM(2, M_ZERO, EPI_STORE); break;
		                case M_PADD: AFM_SMP(2); break; default: AFM_SM(2, M_PSET, EPI_STORE); }
	} else {
		set_error("smoother supports 1 or 2 sweeps per launch");
		return -1;
	}
#undef AFM_SM
#undef AFM_SMP
#undef AFM_SMN
#undef AFM_SM2
	AFM_LAUNCH_CHECK();
	return 0;
}

// blocks of the fused post-smoothing + update launch (one partial sum each)
static inline unsigned update_blocks(int nx, int ny, int sweeps)
{
	return sweeps == 1 ? smooth_blocks<1, EPI_UPDATE>(nx, ny) : smooth_blocks<2, EPI_UPDATE>(nx, ny);
}
// blocks of the flexible CG's post-smoothing launch (three partial sums each)
static inline unsigned fcg_blocks(int nx, int ny, int sweeps)
{
	return sweeps == 1 ? smooth_blocks<1, EPI_UPDATE, 3>(nx, ny) : smooth_blocks<2, EPI_UPDATE, 3>(nx, ny);
}

// nu sweeps as launches of at most 2 sweeps; the first launch applies `mode`.
template<typename T>
static int smooth_n(Ctx* c, unsigned l, int mode, int nu, const T* xin, const T* b, const T* xc, T* xout, T* tmp,
                    const int* done)
{
	if (nu == 0) {
		// x = xin (+ P xc) with no sweep: a 0-sweep launch is not provided, so
		// express it through the restriction-free path of one tile copy.
		set_error("nu == 0 is not supported by the GPU smoother");
		return -1;
	}
	const T* src = xin;
	int left = nu;
	int m = mode;
	T* bufs[2] = {xout, tmp};
	// arrange for the final launch to land in xout
	int launches = (nu + 1) / 2;
	int cur = (launches % 2 == 1) ? 0 : 1;
	while (left > 0) {
		int s = left >= 2 ? 2 : 1;
		T* dst = bufs[cur];
		AFM_TRY(smooth_launch<T>(c, l, m, s, src, b, xc, dst, done));
		src = dst;
		cur ^= 1;
		left -= s;
		m = M_NORMAL;
	}
	return 0;
}

template<typename T>
static int resid_restrict_launch(Ctx* c, unsigned l, bool resid, const T* x, const T* b, T* rc, T* cp, const int* done)
{
	const LevelBuf& L = c->lv[l];
	const T off = (T) (double) (L.n * L.n), dg = (T) (-4.0 * (double) (L.n * L.n));
	dim3 g(((L.nx / 2) / RTX) * ((L.ny / 2) / RTY));
	if (resid)
		hipLaunchKernelGGL((k_resid_restr<T, true>), g, dim3(256), 0, c->stream, x, b, rc, cp, (int) L.nx, (int) L.ny, off, dg, done);
	else
		hipLaunchKernelGGL((k_resid_restr<T, false>), g, dim3(256), 0, c->stream, x, b, rc, cp, (int) L.nx, (int) L.ny, off, dg, done);
	AFM_LAUNCH_CHECK();
	return 0;
}

template<typename T> static T* lvX(LevelBuf& L, int ab);
template<> double* lvX<double>(LevelBuf& L, int ab) { return ab ? L.Xb : L.Xa; }
template<> float* lvX<float>(LevelBuf& L, int ab) { return ab ? L.Xb32 : L.Xa32; }
template<typename T> static T* lvB(LevelBuf& L);
template<> double* lvB<double>(LevelBuf& L) { return L.B; }
template<> float* lvB<float>(LevelBuf& L) { return L.B32; }

template<typename T>
static int coarse_launch(Ctx* c, int mode, int Lt, const T* in, T* out, int nu1, int nu2, const int* done)
{
	// the FP32-cycle forms (preconditioner and FP32 refinement) take levels
	// 3..0 as one matrix product; the FP64 cycles keep the recursion, the
	// reference's operation order
	if (std::is_same<T, float>::value && mode == 0 && Lt == CMAXL && c->op6 && nu1 == 2 && nu2 == 2) {
		hipLaunchKernelGGL(k_coarse_mv6, dim3(C6N / 16), dim3(1024), 0, c->stream, c->op6, c->op6_inv,
		                   (const float*) in, (float*) out, done);
		AFM_LAUNCH_CHECK();
		return 0;
	}
	const double* op3 = (std::is_same<T, float>::value && Lt >= 4) ? c->op3[(nu1 - 1) * 2 + (nu2 - 1)] : nullptr;
	hipLaunchKernelGGL((k_coarse<T>), dim3(1), dim3(CTHREADS), 0, c->stream, mode, Lt, in, out, nu1, nu2, done, op3);
	AFM_LAUNCH_CHECK();
	return 0;
}

// The outer update fused into the finest post-smoothing (mg_solve).
template<typename T>
struct UpdateArgs {
	const double* xs;       // current iterate
	const double* rhs;      // right-hand side (mean removed)
	double* xs_out;         // next iterate
	T* res_out;             // next residual (the next V-cycle's right-hand side)
	SolveCtl* ctl;
	double* partials;       // one r^2 sum per block
	int slot;               // timing event pair of the update (kind 0), or -1
	const MGFold* fold;     // the previous cycle's stop test, or nullptr
	// flexible CG: the last post-smoothing stores z and sums z.r_k, z.r_{k-1},
	// z.L z instead of the update (xs, rhs, xs_out, res_out unused)
	int fcg;
	const T* rkm1;
	T* z;
	double* fpart;
	unsigned fstride;
	int tslot = -1;         // timing event pair of the finest smoothing launches (kinds 1, 2), or -1
};

// record one timing event of kind `kind`, slot `slot` (start 0, end 1)
static inline int kt_record(Ctx* c, int kind, int slot, int end)
{
	if (slot < 0) return 0;
	AFM_HIP(hipEventRecord(c->kt_ev[kind][2 * slot + end], c->stream));
	return 0;
}

// add the elapsed time of kind `kind`, slot `slot` to the totals
static inline int kt_accumulate(Ctx* c, int kind, int slot)
{
	float ms = 0.f;
	AFM_HIP(hipEventElapsedTime(&ms, c->kt_ev[kind][2 * slot], c->kt_ev[kind][2 * slot + 1]));
	c->kt_ms[kind] += ms;
	++c->kt_launches[kind];
	return 0;
}

// recurse(l, B_l) -> X_l (final in Xb); rhs of level l given explicitly.
// With up (and l above the LDS levels), the last post-smoothing instead
// produces the outer update and residual of mg_solve.  Within a solve only
// that launch (and its finaliser) consult the stop flag: the other launches
// write level scratch only, so a cycle enqueued after convergence runs them
// harmlessly instead of waiting on the flag at the start of every launch.
template<typename T>
static int vcycle(Ctx* c, unsigned l, const T* rhs_l, int nu1, int nu2, const int* done,
                  const UpdateArgs<T>* up = nullptr)
{
	const unsigned lc = c->lc;
	if (l <= lc) return coarse_launch<T>(c, 0, (int) l, rhs_l, lvX<T>(c->lv[l], 1), nu1, nu2, done);
	if (up) done = nullptr;
	for (unsigned k = l; k > lc; --k) {
		const T* bk = (k == l) ? rhs_l : lvB<T>(c->lv[k]);
		// pre-smoothing from zero, residual and restriction in one launch
		const int ts = (k == l && up) ? up->tslot : -1;
		AFM_TRY(kt_record(c, 2, ts, 0));
		AFM_TRY(smooth_launch<T>(c, k, M_ZERO, nu1, nullptr, bk, nullptr, lvX<T>(c->lv[k], 0), done,
		                         lvB<T>(c->lv[k - 1]), (k == l && up) ? up->fold : nullptr));
		AFM_TRY(kt_record(c, 2, ts, 1));
	}
	AFM_TRY(coarse_launch<T>(c, 0, (int) lc, lvB<T>(c->lv[lc]), lvX<T>(c->lv[lc], 1), nu1, nu2, done));
	for (unsigned k = lc + 1; k <= l; ++k) {
		const T* bk = (k == l) ? rhs_l : lvB<T>(c->lv[k]);
		if (k == l && up) {
			const LevelBuf& L = c->lv[k];
			const double n2 = (double) L.n * (double) L.n;
			const T off = (T) n2, dg = (T) (-6.0 * n2 + n2 + n2), inv = T(1) / dg;
			const int nx = (int) L.nx, ny = (int) L.ny;
#define AFM_SU(S_)                                                                                              \
			hipLaunchKernelGGL((k_smooth_update<T, T, S_>), dim3(smooth_blocks<S_, EPI_UPDATE>(nx, ny)),       \
			                   dim3(64 * SmoothShape<S_>::NW), 0, c->stream, lvX<T>(c->lv[k], 0), bk,          \
			                   lvX<T>(c->lv[k - 1], 1), up->xs, up->rhs, up->xs_out, up->res_out, nx, ny, off, dg, \
			                   inv, n2, -4.0 * n2, up->ctl, up->partials)
#define AFM_SF(S_)                                                                                              \
			hipLaunchKernelGGL((k_smooth_fcg<T, S_>), dim3(smooth_blocks<S_, EPI_UPDATE, 3>(nx, ny)),             \
			                   dim3(64 * SmoothShape<S_, 3>::NW), 0, c->stream, lvX<T>(c->lv[k], 0), bk,              \
			                   lvX<T>(c->lv[k - 1], 1), up->z, up->rkm1, nx, ny, off, dg, inv, n2, -4.0 * n2,       \
			                   up->fpart, up->fstride)
			const int slot = up->slot;
			AFM_TRY(kt_record(c, 0, slot, 0));
			if (up->fcg) {
				AFM_TRY(kt_record(c, 1, up->tslot, 0));
				if (nu2 == 1) AFM_SF(1);
				else AFM_SF(2);
				AFM_TRY(kt_record(c, 1, up->tslot, 1));
			} else if (nu2 == 1) {
				AFM_SU(1);
			} else {
				AFM_SU(2);
			}
			AFM_TRY(kt_record(c, 0, slot, 1));
#undef AFM_SU
#undef AFM_SF
			AFM_LAUNCH_CHECK();
		} else {
			AFM_TRY(smooth_n<T>(c, k, M_PADD, nu2, lvX<T>(c->lv[k], 0), bk, lvX<T>(c->lv[k - 1], 1),
			                    lvX<T>(c->lv[k], 1), lvX<T>(c->lv[k], 0), done));
		}
	}
	return 0;
}

// rc = R x at level l-1 in FP32 from an FP64 fine field; cp = x rounded.
static int restrict_cvt_launch(Ctx* c, unsigned l, const double* x, float* rc, float* cp)
{
	const LevelBuf& L = c->lv[l];
	dim3 g(((L.nx / 2) / RTX) * ((L.ny / 2) / RTY));
	hipLaunchKernelGGL((k_resid_restr<float, false, double>), g, dim3(256), 0, c->stream, x, (const float*) nullptr, rc,
	                   cp, (int) L.nx, (int) L.ny, 0.f, 0.f, (const int*) nullptr);
	AFM_LAUNCH_CHECK();
	return 0;
}

// nested(T, rhs) -> XSa (mg/solver.h:183-197 unrolled over the levels).
static int nested(Ctx* c, double* rhs, int nu1, int nu2)
{
	const unsigned T = c->levels, lc = c->lc;
	if (T <= lc) return coarse_launch<double>(c, 2, (int) T, rhs, c->XSa, nu1, nu2, nullptr);
	// down: CP_l = B_l, B_{l-1} = R B_l  (CP_T is rhs itself)
	for (unsigned k = T; k > lc; --k) {
		const double* src = (k == T) ? rhs : c->lv[k].B;
		double* cp = (k == T) ? nullptr : c->lv[k].CP;
		AFM_TRY(resid_restrict_launch<double>(c, k, false, src, nullptr, c->lv[k - 1].B, cp, nullptr));
	}
	AFM_TRY(coarse_launch<double>(c, 1, (int) lc, c->lv[lc].B, c->lv[lc].Xb, nu1, nu2, nullptr));
	for (unsigned k = lc + 1; k <= T; ++k) {
		double* dst = (k == T) ? c->XSa : c->lv[k].B;
		const double* cpk = (k == T) ? rhs : c->lv[k].CP;
		AFM_TRY(smooth_n<double>(c, k, M_PSET, nu2, nullptr, cpk, c->lv[k - 1].Xb, dst, c->lv[k].Xa, nullptr));
		if (k < T) AFM_TRY(vcycle<double>(c, k, c->lv[k].B, nu1, nu2, nullptr));
	}
	return 0;
}

// The mixed form's start (levels above the LDS ones): the same nested
// iteration in FP32.  A start only has to be close; the FP64 refinement
// that follows sets the accuracy.  rhs is rounded once, in the first
// restriction, which keeps the rounded copy as the finest right-hand side;
// the level copies (CP) reuse the FP64 buffers.  Result: FP32, finest
// level, in RES32b.
static int nested32(Ctx* c, const double* rhs, int nu1, int nu2)
{
	const unsigned T = c->levels, lc = c->lc;
	float* bT = c->lv[T].B32;
	auto cp32 = [c](unsigned k) { return reinterpret_cast<float*>(c->lv[k].CP); };
	AFM_TRY(restrict_cvt_launch(c, T, rhs, c->lv[T - 1].B32, bT));
	for (unsigned k = T - 1; k > lc; --k)
		AFM_TRY(resid_restrict_launch<float>(c, k, false, c->lv[k].B32, nullptr, c->lv[k - 1].B32, cp32(k), nullptr));
	AFM_TRY(coarse_launch<float>(c, 1, (int) lc, c->lv[lc].B32, c->lv[lc].Xb32, nu1, nu2, nullptr));
	for (unsigned k = lc + 1; k <= T; ++k) {
		float* dst = (k == T) ? c->RES32b : c->lv[k].B32;
		const float* cpk = (k == T) ? bT : cp32(k);
		AFM_TRY(smooth_n<float>(c, k, M_PSET, nu2, nullptr, cpk, c->lv[k - 1].Xb32, dst, c->lv[k].Xa32, nullptr));
		if (k < T) AFM_TRY(vcycle<float>(c, k, c->lv[k].B32, nu1, nu2, nullptr));
	}
	return 0;
}

int mg_setup(Ctx* c)
{
	int cus = 0;
	AFM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
	c->n_cu = std::max(1, std::min(cus, FCG_BLOCKS));
	unsigned T = 0;
	while ((1u << T) < c->n) ++T;
	c->levels = T;
	c->lc = T < (unsigned) CMAXL ? T : (unsigned) CMAXL;
	c->lv.assign(T + 1, LevelBuf());
	for (unsigned l = 0; l <= T; ++l) {
		LevelBuf& L = c->lv[l];
		L.n = 1u << l; L.nx = L.n; L.ny = L.n;
		size_t N = (size_t) L.nx * L.ny;
		if (l < c->lc) continue;                         // LDS only
		AFM_TRY(c->alloc((void**) &L.Xb, N * 8));
		AFM_TRY(c->alloc((void**) &L.B, N * 8));
		AFM_TRY(c->alloc((void**) &L.Xb32, N * 4));
		AFM_TRY(c->alloc((void**) &L.B32, N * 4));
		if (l > c->lc) {
			AFM_TRY(c->alloc((void**) &L.Xa, N * 8));
			AFM_TRY(c->alloc((void**) &L.CP, N * 8));
			AFM_TRY(c->alloc((void**) &L.Xa32, N * 4));
		}
	}
	// the FP32-cycle forms' level-3 operators, one per (nu1, nu2), built here
	// so that no solve allocates (solves may be captured in a graph)
	if (c->levels >= 4)
		for (int k = 0; k < 4; ++k) {
			AFM_TRY(c->alloc((void**) &c->op3[k], 64 * 64 * sizeof(double)));
			hipLaunchKernelGGL(k_coarse_op3, dim3(64), dim3(64), 0, c->stream, 1 + k / 2, 1 + k % 2, c->op3[k]);
			AFM_LAUNCH_CHECK();
		}
	// and level 6's matrix for V(2, 2) (after level 3's, which it uses)
	if (c->levels >= (unsigned) CMAXL) {
		AFM_TRY(c->alloc((void**) &c->op6, (size_t) C6N * C6N * sizeof(_Float16)));
		AFM_TRY(c->alloc((void**) &c->op6_inv, (size_t) C6N * sizeof(float)));
		float* m = nullptr;
		AFM_HIP(hipMalloc((void**) &m, (size_t) C6N * C6N * sizeof(float)));
		hipLaunchKernelGGL(k_coarse_op6, dim3(C6N), dim3(CTHREADS), 0, c->stream, 2, 2, c->op3[3], m);
		hipLaunchKernelGGL(k_coarse_op6_pack, dim3(C6N), dim3(256), 0, c->stream, m, c->op6, c->op6_inv);
		const hipError_t le = hipGetLastError();
		const hipError_t se = hipStreamSynchronize(c->stream);
		(void) hipFree(m);
		AFM_HIP(le);
		AFM_HIP(se);
	}
	AFM_TRY(c->alloc((void**) &c->XSa, c->N * 8));
	AFM_TRY(c->alloc((void**) &c->XSb, c->N * 8));
	AFM_TRY(c->alloc((void**) &c->RES, c->N * 8));
	AFM_TRY(c->alloc((void**) &c->RESb, c->N * 8));
	AFM_TRY(c->alloc((void**) &c->RES32, c->N * 4));
	AFM_TRY(c->alloc((void**) &c->RES32b, c->N * 4));
	{
		// one partial sum per block of the largest reduction of a solve
		size_t nb = MAX_RED_BLOCKS;
		nb = std::max<size_t>(nb, update_blocks((int) c->nx, (int) c->ny, 1));
		nb = std::max<size_t>(nb, update_blocks((int) c->nx, (int) c->ny, 2));
		nb = std::max<size_t>(nb, fcg_blocks((int) c->nx, (int) c->ny, 1));
		nb = std::max<size_t>(nb, fcg_blocks((int) c->nx, (int) c->ny, 2));
		AFM_TRY(c->alloc((void**) &c->upartials, nb * 8));
		// flexible CG: the dots launch's three columns, its search directions
		// and the update launch's sums (alternating by iteration)
		c->fstride = (unsigned) nb;
		AFM_TRY(c->alloc((void**) &c->fpartials, 3 * nb * 8));
		for (int j = 0; j < 2; ++j) {
			AFM_TRY(c->alloc((void**) &c->gpartials[j], 3 * (size_t) FCG_BLOCKS * 8));
			AFM_TRY(c->alloc((void**) &c->P64[j], c->N * 8));
		}
	}
	return 0;
}

static int read_ctl(Ctx* c, int i)
{
	AFM_HIP(hipMemcpyAsync(&c->hctl[i], &c->ctl[i], sizeof(SolveCtl), hipMemcpyDeviceToHost, c->stream));
	AFM_HIP(hipStreamSynchronize(c->stream));
	return 0;
}

// The solution and the solve's end, shared by both outer iterations: x is
// where the last update left it (the caller takes it there with sol).
static int mg_finish(Ctx* c, double* rhs, const double* x, int* cycles, double* final_rms, const double** sol)
{
	const int cnt = c->hctl[0].count;
	if (sol) *sol = x;
	else AFM_HIP(hipMemcpyAsync(rhs, x, c->N * 8, hipMemcpyDeviceToDevice, c->stream));
	c->last_mg_cycles = cnt;
	if (cycles) *cycles = cnt;
	if (final_rms) *final_rms = c->hctl[0].res;
	// the stop flag also fires at the cycle limit: convergence is the test
	if (!(c->hctl[0].res <= c->hctl[0].tol)) { set_error("multigrid did not converge within max_cycles"); return -4; }
	return 0;
}

// Flexible CG preconditioned by one V-cycle in T per iteration (see
// k_fcg_update), from x_0 = 0; the stop test and tolerance as mg_solve's.
// k_mean has run (ctl started, mean in ctl).
template<typename T>
static int fcg_solve(Ctx* c, double* rhs, double tol, int tol_mode, int nu1, int nu2, int max_cycles, int* cycles,
                     double* final_rms, const double** sol, const double* x0)
{
	constexpr bool F32 = std::is_same<T, float>::value;
	SolveCtl* dctl = &c->ctl[0];
	const size_t N = c->N;
	const unsigned rb = red_blocks(N), Lt = c->levels;
	const int nx = (int) c->nx, ny = (int) c->ny;
	// the update: register tiles where the grid has at least one (and the
	// tiles' partial sums fit), else one cell per thread
	const bool tiled = nx >= 128 && ny >= FU_LY;
	const size_t gb0 = (N + 255) / 256;
	const unsigned gb = tiled ? std::min(fcg_update_blocks(nx, ny), 2u * (unsigned) c->n_cu)
	                          : (unsigned) (gb0 < (size_t) FCG_BLOCKS ? gb0 : (size_t) FCG_BLOCKS);
	const double off = (double) (c->n * c->n), dg = -4.0 * off;
	double* xs[2] = {c->XSa, c->XSb};
	// r_k, rounded to the cycle's type, alternating: the cycle of iteration
	// k reads r[k & 1], the update writes r[(k + 1) & 1]
	T* r[2];
	if (F32) { r[0] = (T*) c->RES32; r[1] = (T*) c->RES32b; }
	else { r[0] = (T*) c->RES; r[1] = (T*) c->RESb; }
	// r_0 = rhs - mean - L x_0; rhs keeps its mean (the updates subtract it,
	// with the same bits, instead of reading rhs - mean back)
	// (mg/solver.h:224-225) and r_0, rounded, as the first cycle's right-hand
	// side; the relative tolerance
	update_resid_launch<T, T>(c, rb, x0, (const T*) nullptr, rhs, (double*) nullptr,
	                          r[0], off, dg, 2, 1, dctl, c->partials);
	hipLaunchKernelGGL(k_update_fin, dim3(1), dim3(256), 0, c->stream, tol_mode == 1 ? 2 : 1, rb, N, dctl, c->partials,
	                   tol);
	AFM_LAUNCH_CHECK();
	const int* done = &dctl->done;
	const bool fused = Lt > c->lc;     // else every cycle is one coarse-kernel launch
	const unsigned fnblk = fused ? fcg_blocks(nx, ny, nu2) : rb;
	const unsigned fstride = fused ? c->fstride : (unsigned) MAX_RED_BLOCKS;
	T* z = lvX<T>(c->lv[Lt], 1);
	int enq = 0;
	int batch = c->last_mg_cycles > 0 ? c->last_mg_cycles : 2;
	for (;;) {
		if (batch > max_cycles - enq) batch = max_cycles - enq;
		const int enq0 = enq;
		for (int s = 0; s < batch; ++s) {
			const int k = enq;
			const T* rk = r[k & 1];
			const T* rkm1 = k ? r[(k + 1) & 1] : rk;      // r_{k-1} (at k = 0 unused)
			if (fused) {
				// the stop test of iteration k-1 rides on this cycle's first
				// launch (all but a batch's first iteration)
				const MGFold fold{dctl, c->gpartials[(k + 1) & 1], gb, k, (double) N, 1};
				UpdateArgs<T> up{};
				up.ctl = dctl;
				up.slot = -1;
				up.fold = s > 0 ? &fold : nullptr;
				up.fcg = 1;
				up.tslot = (c->kt_on && s < Ctx::KT_SLOTS) ? s : -1;
				up.rkm1 = rkm1;
				up.z = z;
				up.fpart = c->fpartials;
				up.fstride = fstride;
				AFM_TRY(vcycle<T>(c, Lt, rk, nu1, nu2, done, &up));
			} else {
				AFM_TRY(vcycle<T>(c, Lt, rk, nu1, nu2, done));
				hipLaunchKernelGGL((k_fcg_dots<T>), dim3(rb), dim3(256), 0, c->stream, z, rk, rkm1, nx, ny, off, dg,
				                   c->fpartials);
			}
			const FcgArgs<T> a{z, k ? (const T*) c->P64[(k + 1) & 1] : nullptr, (T*) c->P64[k & 1], k ? xs[k & 1] : x0,
			                   xs[(k + 1) & 1], rhs, r[(k + 1) & 1], dctl, c->fpartials, fstride, fnblk,
			                   c->gpartials[k & 1], k};
			const int slot = (c->kt_on && k >= 1 && s < Ctx::KT_SLOTS) ? s : -1;
			AFM_TRY(kt_record(c, 0, slot, 0));
			const dim3 blk(tiled ? 64 * FU_NW : 256);
#define AFM_FU(F_, Z_)                                                                                    \
			do {                                                                                          \
				if (tiled) hipLaunchKernelGGL((k_fcg_update<T, F_, Z_>), dim3(gb), blk, 0, c->stream, a, nx, ny, off, dg);    \
				else hipLaunchKernelGGL((k_fcg_update_1d<T, F_, Z_>), dim3(gb), blk, 0, c->stream, a, nx, ny, off, dg);   \
			} while (0)
			if (k > 0) AFM_FU(false, false);
			else if (x0) AFM_FU(true, false);
			else AFM_FU(true, true);
#undef AFM_FU
			AFM_TRY(kt_record(c, 0, slot, 1));
			if (!fused)
				hipLaunchKernelGGL(k_fcg_fin, dim3(1), dim3(256), 0, c->stream, k + 1, gb, (double) N, dctl,
				                   c->gpartials[k & 1]);
			AFM_LAUNCH_CHECK();
			++enq;
		}
		if (fused) {
			hipLaunchKernelGGL(k_fcg_fin, dim3(1), dim3(256), 0, c->stream, enq, gb, (double) N, dctl,
			                   c->gpartials[(enq + 1) & 1]);
			AFM_LAUNCH_CHECK();
		}
		AFM_TRY(read_ctl(c, 0));
		if (c->kt_on) {
			// launches of iterations that ran (later ones exit at once, or
			// run on scratch: the cycle of iteration k ran if k < count)
			for (int s = 0; s < batch && s < Ctx::KT_SLOTS; ++s) {
				const int k = enq0 + s;
				if (k >= c->hctl[0].count) continue;
				if (k >= 1) AFM_TRY(kt_accumulate(c, 0, s));
				if (fused) {
					AFM_TRY(kt_accumulate(c, 1, s));
					AFM_TRY(kt_accumulate(c, 2, s));
				}
			}
		}
		if (c->hctl[0].done || enq >= max_cycles) break;
		batch = 1;
	}
	const int cnt = c->hctl[0].count;
	if (cnt == 0) {
		// r_0 already met the tolerance: x = x_0
		if (x0) AFM_HIP(hipMemcpyAsync(xs[0], x0, N * 8, hipMemcpyDeviceToDevice, c->stream));
		else hipLaunchKernelGGL(k_zero, dim3((unsigned) ((N + 255) / 256)), dim3(256), 0, c->stream, xs[0], N);
		AFM_LAUNCH_CHECK();
	}
	return mg_finish(c, rhs, xs[cnt & 1], cycles, final_rms, sol);
}

int mg_solve(Ctx* c, double* rhs, double tol, int tol_mode, int nu1, int nu2, int precision, int max_cycles,
             int* cycles, double* final_rms, const double** sol, const double* x0, bool sums_ready)
{
	if (nu1 < 1 || nu1 > 2 || nu2 < 1 || nu2 > 2) { set_error("nu1, nu2 must be 1 or 2 (one fused launch per smoothing step)"); return -1; }
	if (precision < 0 || precision > 3) {
		set_error("precision must be 0 (fp64), 1 (fp32 V-cycle preconditioning fp64 CG), 2 (fp32 V-cycles in fp64 "
		          "refinement) or 3 (fp64 V-cycle preconditioning fp64 CG)");
		return -1;
	}
	SolveCtl* dctl = &c->ctl[0];
	const size_t N = c->N;
	const unsigned rb = red_blocks(N);
	const unsigned T = c->levels;
	const double off = (double) (c->n * c->n), dg = -4.0 * off;
	// the mean before anything else, as the oracle computes it
	// (mg/solver.h:224 does the mean after the nested start; the value is the
	// same because rhs is not touched there); for a relative tolerance the RMS
	// of rhs - mean is summed by the start's residual launch below
	if (!sums_ready)
		hipLaunchKernelGGL(k_mean, dim3(rb), dim3(256), 0, c->stream, rhs, N, 0, dctl, c->partials,
		                   tol_mode == 1 ? 0.0 : tol, max_cycles);
	hipLaunchKernelGGL(k_mean_fin, dim3(1), dim3(256), 0, c->stream, 0, tol, rb, N, dctl, c->partials);
	AFM_LAUNCH_CHECK();
	if (precision == 1) return fcg_solve<float>(c, rhs, tol, tol_mode, nu1, nu2, max_cycles, cycles, final_rms, sol, x0);
	if (precision == 3) return fcg_solve<double>(c, rhs, tol, tol_mode, nu1, nu2, max_cycles, cycles, final_rms, sol, x0);
	// the reference's iteration (precision 0), or the same in FP32 V-cycles
	// (2): nested start and its residual, mean removed from rhs in place
	if (precision == 2 && T > c->lc) {
		AFM_TRY(nested32(c, rhs, nu1, nu2));
		update_resid_launch<float, float>(c, rb, (const double*) nullptr, (const float*) c->RES32b, rhs, c->XSa,
		                                  c->RES32, off, dg, 1, 1, dctl, c->partials);
	} else {
		AFM_TRY(nested(c, rhs, nu1, nu2));
		if (precision == 0)
			update_resid_launch<double, double>(c, rb, c->XSa, (const double*) nullptr, rhs, c->XSa,
			                                    c->RES, off, dg, 1, 1, dctl, c->partials);
		else
			update_resid_launch<float, float>(c, rb, c->XSa, (const float*) nullptr, rhs, c->XSa,
			                                  c->RES32, off, dg, 1, 1, dctl, c->partials);
	}
	hipLaunchKernelGGL(k_update_fin, dim3(1), dim3(256), 0, c->stream, tol_mode == 1 ? 2 : 1, rb, N, dctl, c->partials,
	                   tol);
	AFM_LAUNCH_CHECK();
	const int* done = &dctl->done;
	// above the LDS levels the update rides on the last post-smoothing
	// launch, with the residual double-buffered (read as b, written as r)
	const bool fused = T > c->lc;
	const unsigned ub = fused ? update_blocks((int) c->nx, (int) c->ny, nu2) : rb;
	int enq = 0;   // cycles enqueued
	int batch = c->last_mg_cycles > 0 ? c->last_mg_cycles : 2;
	double* xs[2] = {c->XSa, c->XSb};
	double* res64[2] = {c->RES, c->RESb};
	float* res32[2] = {c->RES32, c->RES32b};
	for (;;) {
		if (batch > max_cycles - enq) batch = max_cycles - enq;
		const int enq0 = enq;
		for (int s = 0; s < batch; ++s) {
			const int slot = (c->kt_on && fused && s < Ctx::KT_SLOTS) ? s : -1;
			double* cur = xs[enq & 1];
			double* nxt = xs[(enq + 1) & 1];
			const int rc = fused ? (enq & 1) : 0, rn = fused ? ((enq + 1) & 1) : 0;
			// the stop test of the cycle before rides on this cycle's first
			// launch (all but the batch's first cycle; the batch's last cycle
			// is tested by its own launch below, before the host reads it)
			const MGFold fold{dctl, c->upartials, ub, enq, (double) N};
			const MGFold* fp = s > 0 ? &fold : nullptr;
			if (precision == 0) {
				if (fused) {
					UpdateArgs<double> up{cur, rhs, nxt, res64[rn], dctl, c->upartials, slot, fp};
					up.tslot = slot;
					AFM_TRY(vcycle<double>(c, T, res64[rc], nu1, nu2, done, &up));
				} else {
					AFM_TRY(vcycle<double>(c, T, res64[rc], nu1, nu2, done));
					update_resid_launch<double, double>(c, rb, cur, (const double*) c->lv[T].Xb, rhs, nxt,
					                                    res64[rn], off, dg, 0, 0, dctl, c->upartials);
				}
			} else {
				if (fused) {
					UpdateArgs<float> up{cur, rhs, nxt, res32[rn], dctl, c->upartials, slot, fp};
					up.tslot = slot;
					AFM_TRY(vcycle<float>(c, T, res32[rc], nu1, nu2, done, &up));
				} else {
					AFM_TRY(vcycle<float>(c, T, res32[rc], nu1, nu2, done));
					update_resid_launch<float, float>(c, rb, cur, (const float*) c->lv[T].Xb32, rhs, nxt,
					                                  res32[rn], off, dg, 0, 0, dctl, c->upartials);
				}
			}
			if (!fused) hipLaunchKernelGGL(k_update_fin, dim3(1), dim3(256), 0, c->stream, 0, ub, N, dctl, c->upartials);
			AFM_LAUNCH_CHECK();
			++enq;
		}
		if (fused) {
			hipLaunchKernelGGL(k_update_fin, dim3(1), dim3(256), 0, c->stream, 0, ub, N, dctl, c->upartials);
			AFM_LAUNCH_CHECK();
		}
		AFM_TRY(read_ctl(c, 0));
		if (c->kt_on && fused) {
			// launches of cycles that ran (later ones exit at once)
			for (int s = 0; s < batch && s < Ctx::KT_SLOTS && enq0 + s < c->hctl[0].count; ++s) {
				AFM_TRY(kt_accumulate(c, 0, s));
				AFM_TRY(kt_accumulate(c, 2, s));
			}
		}
		if (c->hctl[0].done || enq >= max_cycles) break;
		batch = 1;                  // beyond the prediction, one cycle at a time
	}
	// the solution replaces rhs (mg/solver.h:218-237), unless the caller
	// takes it where it is (valid until the context's next solve)
	return mg_finish(c, rhs, xs[c->hctl[0].count & 1], cycles, final_rms, sol);
}

// One V-cycle recurse(T, rhs) on the residual equation, result in place.
int mg_vcycle(Ctx* c, double* rhs_inout, int nu1, int nu2)
{
	AFM_TRY(vcycle<double>(c, c->levels, rhs_inout, nu1, nu2, nullptr));
	AFM_HIP(hipMemcpyAsync(rhs_inout, c->lv[c->levels].Xb, c->N * 8, hipMemcpyDeviceToDevice, c->stream));
	return 0;
}

// ---- single operators at a global level (n >= 64), for parity tests ----
__global__ void k_lap_apply(const double* __restrict__ x, double* __restrict__ y, int nx, int ny, double off, double dg)
{
	size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	size_t N = (size_t) nx * ny;
	if (i >= N) return;
	int gx, gy;
	cell_xy(i, nx, gx, gy);
	size_t ie = gx == nx - 1 ? i - (nx - 1) : i + 1, iw = gx == 0 ? i + (nx - 1) : i - 1;
	size_t in_ = gy == ny - 1 ? gx : i + nx, is = gy == 0 ? i + N - nx : i - nx;
	y[i] = stencil5p2<double, true>(off, dg, x[i], x[ie], x[iw], x[in_], x[is], gx == nx - 1, gx == 0, gy == ny - 1, gy == 0);
}

static int check_level(Ctx* c, unsigned level)
{
	// the tiled kernels need n >= 64 (their regions wrap at most once)
	if (level > c->levels || level < 6) { set_error("level must be a level of the context with n >= 64"); return -1; }
	return 0;
}

int mg_apply_laplacian(Ctx* c, unsigned level, const double* x, double* y)
{
	if (level > c->levels) { set_error("level out of range"); return -1; }
	const LevelBuf& L = c->lv[level];
	size_t N = (size_t) L.nx * L.ny;
	if (L.nx < 3) { set_error("laplacian apply needs n >= 4"); return -1; }
	hipLaunchKernelGGL(k_lap_apply, dim3((N + 255) / 256), dim3(256), 0, c->stream, x, y, (int) L.nx, (int) L.ny,
	                   (double) (L.n * L.n), -4.0 * (double) (L.n * L.n));
	AFM_LAUNCH_CHECK();
	return 0;
}

int mg_smooth(Ctx* c, unsigned level, const double* x_in, const double* b, double* x_out, int sweeps, int zero_init)
{
	AFM_TRY(check_level(c, level));
	if (sweeps < 1 || sweeps > 2) { set_error("sweeps must be 1 or 2"); return -1; }
	return smooth_launch<double>(c, level, zero_init ? M_ZERO : M_NORMAL, sweeps, x_in, b, nullptr, x_out, nullptr);
}

int mg_restrict(Ctx* c, unsigned level, const double* fine, double* coarse)
{
	AFM_TRY(check_level(c, level));
	return resid_restrict_launch<double>(c, level, false, fine, nullptr, coarse, nullptr, nullptr);
}

// fine = P coarse, overwritten, followed by no sweep: realised as the
// M_PSET smoother path would smooth, so use a plain kernel instead.
__global__ void k_prolong(const double* __restrict__ xc, double* __restrict__ xf, int nx, int ny)
{
	size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x;
	if (i >= (size_t) nx * ny) return;
	xf[i] = prolong_at(xc, (int) (i % nx), (int) (i / nx), nx >> 1, ny >> 1);
}

int mg_prolong(Ctx* c, unsigned level, const double* coarse, double* fine)
{
	if (level > c->levels || level < 1) { set_error("level out of range"); return -1; }
	const LevelBuf& L = c->lv[level];
	size_t N = (size_t) L.nx * L.ny;
	hipLaunchKernelGGL(k_prolong, dim3((N + 255) / 256), dim3(256), 0, c->stream, coarse, fine, (int) L.nx, (int) L.ny);
	AFM_LAUNCH_CHECK();
	return 0;
}

}  // namespace afm
