// afm_ib.hip -- Lagrangian <-> Eulerian transfer on gfx950.
//
// Interpolation (ib/grid.h:68-84, marker targets): one lane per marker, its
// eight delta factors, 16 taps summed in the reference's order.  Same
// arithmetic as the reference, so the result is bit-exact against the CPU
// restatement.
//
// Spreading (ib/grid.h:68-84, grid targets): the reference gathers, for
// every grid point, the markers found by a kd-tree box query.  Here every
// (marker, component) is binned by the 4x4 block of its window's base cell:
// counts, a chunked exclusive scan, an unordered fill, then each record is
// written to its bin in ascending marker index (its rank in the bin).  Four
// lanes per bin (one per cell row) then merge the four bins whose base cells
// can reach its 16 cells and sum each cell's terms (w * ds) * F in ascending
// marker order.
// No float atomics and no capacity limit: the result is identical run to run
// and equals a sequential scatter over markers in index order.
#include "afm_device.h"

namespace afm {

constexpr int SB = 4;             // spread bin edge (window base cells)
constexpr int SCAN_BLOCKS = 1024; // at most this many scan chunks
constexpr int GCLS = 4;           // density classes of gather tiles (records reaching the tile:
                                  // >= 1024, >= 128, >= 1, none), launched densest first
constexpr int GB = 8;        // gather block edge, bins
constexpr int GCAP = 2048;   // records staged per block
constexpr int GMIN = 64;     // fewer: short lists, not worth the copy

__global__ void k_zero_int(int* __restrict__ a, long count)
{
	const long i = blockIdx.x * (long) blockDim.x + threadIdx.x;
	if (i < count) a[i] = 0;
}

__global__ void k_delta(const double* __restrict__ r, long count, unsigned n, double* __restrict__ out)
{
	long i = blockIdx.x * (long) blockDim.x + threadIdx.x;
	if (i < count) out[i] = delta(r[i], n);
}

// One lane per marker: its window's eight delta factors (two batches of
// four fast paths), then the 16 taps summed in the reference's order (tap
// a + 4 b, x fastest).  (Two lanes per marker, one axis each, were slower:
// 26.8 against 21.4 us at 262,144 markers.)
__global__ void k_interp(int comp, const double* __restrict__ f, const double* __restrict__ X, long m, WinGeom g,
                         double* __restrict__ U)
{
	const long k = blockIdx.x * (long) blockDim.x + threadIdx.x;
	if (k >= m) return;
	int ix0, iy0;
	double dx[4], dy[4];
	ib_window_g(comp, X[2 * k], X[2 * k + 1], g, ix0, iy0, dx, dy);
	U[k] = ib_gather(f, g.nx, g.ny, ix0, iy0, dx, dy, g.vol);
}

// X_h = X + half * (J_u(X) u, J_v(X) v).  The windows are the ones the
// spreading of the same step stored for X (rec[comp * m + k]): the same
// delta factors, not computed twice.
__global__ void k_ib_move_half(const double* __restrict__ u, const double* __restrict__ v, const double* __restrict__ X,
                               const MarkerRec* __restrict__ rec, long m, unsigned n, unsigned nx, unsigned ny,
                               double half, double* __restrict__ Xh)
{
	long k = blockIdx.x * (long) blockDim.x + threadIdx.x;
	if (k >= m) return;
	const double vol = 1. / (n * n);
	const MarkerRec& a = rec[k];
	const MarkerRec& b = rec[m + k];
	const double Ux = ib_gather(u, nx, ny, a.ix0, a.iy0, a.dx, a.dy, vol);
	const double Uy = ib_gather(v, nx, ny, b.ix0, b.iy0, b.dx, b.dy, vol);
	Xh[2 * k] = X[2 * k] + half * Ux;
	Xh[2 * k + 1] = X[2 * k + 1] + half * Uy;
}

// X = X + dt * (0.5 * (J(X_h) u_old + J(X_h) u)), windows of X_h from the
// step's second spreading.
__global__ void k_ib_move_full(const double* __restrict__ uo, const double* __restrict__ vo,
                               const double* __restrict__ u, const double* __restrict__ v,
                               const MarkerRec* __restrict__ rec, long m, unsigned n, unsigned nx, unsigned ny,
                               double dt, double* __restrict__ X)
{
	long k = blockIdx.x * (long) blockDim.x + threadIdx.x;
	if (k >= m) return;
	const double vol = 1. / (n * n);
	const MarkerRec& a = rec[k];
	const MarkerRec& b = rec[m + k];
	const double U0x = ib_gather(uo, nx, ny, a.ix0, a.iy0, a.dx, a.dy, vol);
	const double U1x = ib_gather(u, nx, ny, a.ix0, a.iy0, a.dx, a.dy, vol);
	const double U0y = ib_gather(vo, nx, ny, b.ix0, b.iy0, b.dx, b.dy, vol);
	const double U1y = ib_gather(v, nx, ny, b.ix0, b.iy0, b.dx, b.dy, vol);
	X[2 * k] = X[2 * k] + dt * (0.5 * (U0x + U1x));
	X[2 * k + 1] = X[2 * k + 1] + dt * (0.5 * (U0y + U1y));
}

// Per (marker, component), two lanes: lane 0 the window's base cell, bin
// and x factors, lane 1 the y factors and the source (the two halves of the
// record).  With R/k given, the source is the tether force F = -k (X - R);
// otherwise F[comp][k].
__global__ void k_spread_prep(int comp_mask, const double* __restrict__ Fx, const double* __restrict__ Fy,
                              const double* __restrict__ ds, const double* __restrict__ X,
                              const double* __restrict__ R, const double* __restrict__ kk, long m, WinGeom g,
                              int sbx, int sby, int nb1, MarkerRec* __restrict__ rec, int* __restrict__ counts,
                              int* __restrict__ tflag, int tpr, int ntile)
{
	const long t = blockIdx.x * (long) blockDim.x + threadIdx.x;
	const long k = t >> 1;
	const int p = (int) (t & 1);
	const int comp = blockIdx.y;
	if (k >= m || !((comp_mask >> comp) & 1)) return;
	const double cx = X[2 * k], cy = X[2 * k + 1];
	int ix0, iy0;
	double fct[4];
	ib_window_axis(comp, cx, cy, g, p, ix0, iy0, fct);
	MarkerRec& r = rec[comp * m + k];
	if (p == 0) {
		const int tile = ix0 / SB + sbx * (iy0 / SB);
		r.ix0 = ix0;
		r.iy0 = iy0;
		r.tile = tile;
		r.pad = 0;
#pragma unroll
		for (int d = 0; d < 4; ++d) r.dx[d] = fct[d];
		atomicAdd(&counts[comp * nb1 + tile], 1);
		// the gather tiles whose cells the window reaches: those of bins
		// (bx, by), (bx + 1, by), (bx, by + 1), (bx + 1, by + 1), periodic
		const int bx = ix0 / SB, by = iy0 / SB;
		const int bx1 = bx + 1 == sbx ? 0 : bx + 1, by1 = by + 1 == sby ? 0 : by + 1;
		int* tf = tflag + comp * ntile;
		tf[bx / GB + tpr * (by / GB)] = 1;
		tf[bx1 / GB + tpr * (by / GB)] = 1;
		tf[bx / GB + tpr * (by1 / GB)] = 1;
		tf[bx1 / GB + tpr * (by1 / GB)] = 1;
	} else {
#pragma unroll
		for (int d = 0; d < 4; ++d) r.dy[d] = fct[d];
		r.ds = ds[k];
		r.F = R ? (comp == 0 ? -kk[k] * (cx - R[2 * k]) : -kk[k] * (cy - R[2 * k + 1]))
		        : (comp == 0 ? Fx[k] : Fy[k]);
	}
}

// Exclusive scan of the bin counts in two launches: chunk sums (block c
// reduces bins [c C, (c+1) C), C = 256 * items), then block c adds the sums
// of the chunks before it and scans its own chunk.  Integer sums, so the
// result does not depend on the order of the additions.
__device__ inline int block_int_sum(int v, int* red)
{
	const int tid = threadIdx.x;
	red[tid] = v;
	__syncthreads();
	for (int d = 128; d > 0; d >>= 1) {
		if (tid < d) red[tid] += red[tid + d];
		__syncthreads();
	}
	const int s = red[0];
	__syncthreads();
	return s;
}

// A scan chunk is 256 * items consecutive bins, visited as `items` rows of
// 256 (thread t takes bin row * 256 + t, so a wave's loads and stores are
// contiguous).
__global__ __launch_bounds__(256) void k_spread_chunk(const int* __restrict__ counts, int* __restrict__ chunk,
                                                      int nbins, int items)
{
	__shared__ int red[256];
	const long lo = (long) blockIdx.x * 256 * items + threadIdx.x;
	int s = 0;
	for (int j = 0; j < items; ++j)
		if (lo + 256L * j < nbins) s += counts[lo + 256L * j];
	s = block_int_sum(s, red);
	if (threadIdx.x == 0) chunk[blockIdx.x] = s;
}

// inclusive scan of one int per thread over a 256-thread block; `total` the
// block's sum.  wsum: 4 ints of LDS, free again on return.
__device__ inline int block_int_scan(int v, int* wsum, int& total)
{
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
	for (int d = 1; d < 64; d <<= 1) {
		const int t = __shfl_up(v, d, 64);
		if (lane >= d) v += t;
	}
	if (lane == 63) wsum[w] = v;
	__syncthreads();
	int add = 0;
	for (int q = 0; q < w; ++q) add += wsum[q];
	total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
	__syncthreads();
	return v + add;
}

// offs = exclusive scan of counts (offs[nbins] = the total); cursor and
// counts cleared for the next spreading
__global__ __launch_bounds__(256) void k_spread_scan(int* __restrict__ counts, const int* __restrict__ chunk,
                                                     int* __restrict__ offs, int* __restrict__ cursor, int nbins,
                                                     int items, int* __restrict__ ccount)
{
	__shared__ int red[256];
	__shared__ int wsum[4];
	const int c = blockIdx.x, tid = threadIdx.x;
	// the gather order's class counts of this spreading start at zero
	if (ccount && c == 0 && tid < 2 * GCLS) ccount[tid] = 0;
	int base = 0;
	for (int q = tid; q < c; q += 256) base += chunk[q];
	base = block_int_sum(base, red);
	const long lo = (long) c * 256 * items + tid;
	for (int j = 0; j < items; ++j) {
		const long i = lo + 256L * j;
		const int v = i < nbins ? counts[i] : 0;
		int total;
		const int incl = block_int_scan(v, wsum, total);
		if (i < nbins) {
			offs[i] = base + incl - v;
			cursor[i] = 0;
			counts[i] = 0;   // the next spreading's counts start at zero
			if (i == nbins - 1) offs[nbins] = base + incl;
		}
		base += total;
	}
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

// Records in bin order, each bin in ascending marker index: one lane per
// (marker, component) counts the smaller marker indices in its bin's
// (unordered) list -- its rank -- and writes its record to that slot.
// Bins hold a few markers, so the count is a short scan of cached indices.
__global__ void k_spread_rank(int comp_mask, const int* __restrict__ offs, const int* __restrict__ list,
                              const MarkerRec* __restrict__ rec, long m, int nb1, BinRec* __restrict__ srec)
{
	const long k = blockIdx.x * (long) blockDim.x + threadIdx.x;
	const int comp = blockIdx.y;
	if (k >= m || !((comp_mask >> comp) & 1)) return;
	const MarkerRec& r = rec[comp * m + k];
	const int b = comp * nb1 + r.tile;
	const int lo = offs[b], hi = offs[b + 1];
	int rank = 0;
	for (int i = lo; i < hi; ++i) rank += list[i] < (int) k;
	BinRec o;
	o.key = (int) k;
	o.base = r.ix0 | (r.iy0 << 16);
	o.pad0 = o.pad1 = 0;
	o.ds = r.ds;
	o.F = r.F;
#pragma unroll
	for (int t = 0; t < 4; ++t) { o.dx[t] = r.dx[t]; o.dy[t] = r.dy[t]; }
	srec[lo + rank] = o;
}

// Spreading gather.  The 4x4 cells of bin (bx, by) take their terms from
// markers whose window base lies at most 3 cells before them, i.e. in bins
// (bx-1 .. bx) x (by-1 .. by) (periodic).  Four lanes per bin, one per cell
// row: each merges those four lists, each sorted by marker, into ascending
// marker order (the four lanes run the same merge in the same instructions)
// and adds each marker's terms ((dx[a] dy[b]) ds) F to the cells of its row
// that the window covers, so every cell sums its terms in ascending marker
// index.
//
// A block is 8 x 8 bins.  It first copies the keys of its 9 x 9 bins into
// LDS, so the merge, a chain of dependent key reads, runs on LDS; a record's
// factors are loaded from global memory one merge step ahead of their use.
// Blocks with fewer than GMIN or more than GCAP records, and grids of fewer
// than 8 x 8 bins, merge from global memory instead (same order, same sums).

// a record's values in registers (named fields, no arrays: the factor
// selection below must stay a choice among registers)
struct RecVals {
	int base;
	double ds, F, x0, x1, x2, x3, y0, y1, y2, y3;
};

__device__ inline RecVals load_rec(const BinRec* __restrict__ srec, int g)
{
	const double2* rv = reinterpret_cast<const double2*>(srec + g);
	const int4 kb = *reinterpret_cast<const int4*>(srec + g);
	const double2 sF = rv[1], x01 = rv[2], x23 = rv[3], y01 = rv[4], y23 = rv[5];
	RecVals v;
	v.base = kb.y;
	v.ds = sF.x; v.F = sF.y;
	v.x0 = x01.x; v.x1 = x01.y; v.x2 = x23.x; v.x3 = x23.y;
	v.y0 = y01.x; v.y1 = y01.y; v.y2 = y23.x; v.y3 = y23.y;
	return v;
}

// the record's terms added to the cells (X .. X+3, Yj) that its window covers
__device__ inline void add_terms(const RecVals r, int X, int Yj, int nx, int ny, double (&acc)[SB])
{
	// window column a = X + i - base_x, row b = Yj - base_y
	int ox = X - (r.base & 0xffff), b = Yj - (r.base >> 16);
	if (ox < -3) ox += nx;
	if (b < -3) b += ny;
	const bool vy = (unsigned) b < 4u;
	const double wy = b <= 0 ? r.y0 : (b == 1 ? r.y1 : (b == 2 ? r.y2 : r.y3));
	// every cell's term, kept only where the window covers the cell (a
	// select, not a branch: the lanes' windows differ)
#pragma unroll
	for (int i = 0; i < SB; ++i) {
		const int a = ox + i;
		const double wx = a <= 0 ? r.x0 : (a == 1 ? r.x1 : (a == 2 ? r.x2 : r.x3));
		const double w = wx * wy;
		const double sum = acc[i] + w * r.ds * r.F;
		acc[i] = (vy && (unsigned) a < 4u) ? sum : acc[i];
	}
}

// Merge of the four lists [head, end) in ascending key.  STAGED: positions
// index lkey / lidx in LDS (key, record index); else they index srec.
template<bool STAGED>
__device__ inline void merge_terms(const BinRec* __restrict__ srec, const int* lkey, const int* lidx, int (&head)[4],
                                   const int (&end)[4], int X, int Yj, int nx, int ny, double (&acc)[SB])
{
	int key[4];
#pragma unroll
	for (int l = 0; l < 4; ++l) key[l] = head[l] < end[l] ? (STAGED ? lkey[head[l]] : srec[head[l]].key) : 0x7fffffff;
	// take the smallest head (markers are distinct: no ties); its record
	// index, or -1 when every list is spent
	auto pick = [&]() -> int {
		int best = 0, bk = key[0];
#pragma unroll
		for (int l = 1; l < 4; ++l)
			if (key[l] < bk) { bk = key[l]; best = l; }
		if (bk == 0x7fffffff) return -1;
		int h = 0, e = 0;
#pragma unroll
		for (int l = 0; l < 4; ++l)
			if (l == best) { h = head[l]; e = end[l]; }
		const int nk = h + 1 < e ? (STAGED ? lkey[h + 1] : srec[h + 1].key) : 0x7fffffff;
#pragma unroll
		for (int l = 0; l < 4; ++l)
			if (l == best) { head[l] = h + 1; key[l] = nk; }
		return STAGED ? lidx[h] : h;
	};
	// the next record's loads go out before this one's terms are added; two
	// register sets in turn, so no copy waits on loads in flight
	int ga = pick();
	if (ga < 0) return;
	RecVals ra = load_rec(srec, ga), rb;
	for (;;) {
		const int gb = pick();
		rb = load_rec(srec, gb < 0 ? ga : gb);
		add_terms(ra, X, Yj, nx, ny, acc);
		if (gb < 0) break;
		ga = pick();
		ra = load_rec(srec, ga < 0 ? gb : ga);
		add_terms(rb, X, Yj, nx, ny, acc);
		if (ga < 0) break;
	}
}

// The gather's launch order: each tile (one thread) joins the list of its
// density class (the records that can reach its cells: those of the 9 x 9
// bins the gather stages); gather block i takes the i-th tile of the lists
// in class order, so the long merges of dense tiles start first instead of
// behind the sparse ones.  The order within a class depends on the atomics,
// but no tile's result depends on when or by which block it is computed.
// Slots are claimed per wave (ballot), then per block (LDS), so a class
// takes one global atomic per block.
__global__ __launch_bounds__(256) void k_gather_order(int comp_mask, const int* __restrict__ offs, int sbx, int sby,
                                                      int nb1, int ntiles, int* __restrict__ ccount,
                                                      int* __restrict__ order)
{
	__shared__ int bcount[GCLS], bbase[GCLS];
	const int comp = (int) blockIdx.y;
	if (!((comp_mask >> comp) & 1)) return;
	const int tile = (int) (blockIdx.x * 256 + threadIdx.x);
	if (threadIdx.x < GCLS) bcount[threadIdx.x] = 0;
	int cls = -1;
	if (tile < ntiles) {
		const int* o = offs + comp * nb1;
		const int bpr = (sbx + GB - 1) / GB;
		const int bx0 = GB * (tile % bpr), by0 = GB * (tile / bpr);
		int total = 0;
#pragma unroll
		for (int ly = 0; ly <= GB; ++ly) {
			// the wrapped bin left of the tile, then its GB bins (contiguous)
			const int gy = by0 - 1 + ly < 0 ? sby - 1 : by0 - 1 + ly;
			const int bl = (bx0 == 0 ? sbx - 1 : bx0 - 1) + sbx * gy, bb = bx0 + sbx * gy;
			total += (o[bl + 1] - o[bl]) + (o[bb + GB] - o[bb]);
		}
		cls = total >= 1024 ? 0 : total >= 128 ? 1 : total >= 1 ? 2 : 3;
	}
	__syncthreads();
	const int lane = (int) (threadIdx.x & 63);
	int slot = 0;
	for (int c2 = 0; c2 < GCLS; ++c2) {
		const unsigned long long mask = __ballot(cls == c2);
		if (!mask) continue;
		const int leader = __ffsll((long long) mask) - 1;
		int base = 0;
		if (lane == leader) base = atomicAdd(&bcount[c2], __popcll(mask));
		base = __shfl(base, leader);
		if (cls == c2) slot = base + __popcll(mask & ((1ull << lane) - 1ull));
	}
	__syncthreads();
	if (threadIdx.x < GCLS)
		bbase[threadIdx.x] = bcount[threadIdx.x] ? atomicAdd(&ccount[comp * GCLS + threadIdx.x], bcount[threadIdx.x]) : 0;
	__syncthreads();
	if (cls >= 0) order[(size_t) (comp * GCLS + cls) * ntiles + bbase[cls] + slot] = tile;
}

__global__ __launch_bounds__(256) void k_spread_gather(int comp_mask, const BinRec* __restrict__ srec,
                                                       const int* __restrict__ offs, unsigned nx, int sbx, int sby,
                                                       int nb1, double* __restrict__ fx, double* __restrict__ fy,
                                                       int accumulate, const int* __restrict__ ccount,
                                                       const int* __restrict__ order, int* __restrict__ tflag)
{
	constexpr int NR = 2 * (GB + 1);   // srec ranges staged per block
	__shared__ int lkey[GCAP], lidx[GCAP];
	__shared__ int rs[NR], rp[NR + 1];
	const int comp = blockIdx.y;
	if (!((comp_mask >> comp) & 1)) return;
	const int tid = (int) threadIdx.x;
	// consecutive blocks go round-robin to the 8 XCDs, which spreads the
	// long merges of a dense cluster over all of them
	const int bpr = (sbx + GB - 1) / GB;
	int tile = (int) blockIdx.x;
	if (order) {
		// the tile of this block in the densest-first order (uniform loads)
		const int nt = (int) gridDim.x;
		int cls = 0;
		while (cls < GCLS - 1 && tile >= ccount[comp * GCLS + cls]) tile -= ccount[comp * GCLS + cls++];
		tile = order[(size_t) (comp * GCLS + cls) * nt + tile];
	}
	const int bx0 = GB * (tile % bpr), by0 = GB * (tile / bpr);
	{
		// a tile no record reaches (flag unset by the window kernel) stores
		// zeros, or leaves the field as it is, without reading the bins
		int* tf = tflag + (size_t) comp * gridDim.x + tile;
		if (!*tf) {
			if (!accumulate) {
				const int q0 = tid >> 2, bxz = bx0 + (q0 & (GB - 1)), byz = by0 + q0 / GB;
				if (bxz < sbx && byz < sby) {
					double2* r2 = reinterpret_cast<double2*>((comp == 0 ? fx : fy) + (size_t) nx * (SB * byz + (tid & 3))
					                                         + SB * bxz);
					r2[0] = make_double2(0.0, 0.0);
					r2[1] = make_double2(0.0, 0.0);
				}
			}
			return;
		}
		// every thread has read the flag before it is cleared for the next call
		__syncthreads();
		if (tid == 0) *tf = 0;
	}
	const int* o = offs + comp * nb1;
	const int ny = SB * sby;
	// lane: cell row j of bin (tx, ty) of the block.  Its four bins' ranges
	// are read first, so their latency overlaps the block's staging
	const int j = tid & 3, q = tid >> 2;
	const int tx = q & (GB - 1), ty = q / GB;
	const int bx = bx0 + tx, by = by0 + ty;
	const bool live = bx < sbx && by < sby;
	int bs[4], bn[4];
#pragma unroll
	for (int l = 0; l < 4; ++l) {
		// bins (bx-1, by-1), (bx, by-1), (bx-1, by), (bx, by)
		const int gx = bx - 1 + (l & 1) < 0 ? sbx - 1 : bx - 1 + (l & 1);
		const int gy = by - 1 + (l >> 1) < 0 ? sby - 1 : by - 1 + (l >> 1);
		const int b = gx + sbx * gy;
		bs[l] = live ? o[b] : 0;
		bn[l] = live ? o[b + 1] - bs[l] : 0;
	}
	bool staged = sbx >= GB && sby >= GB;
	if (staged) {
		// local bin (lx, ly) in 0..GB is global (bx0 - 1 + lx, by0 - 1 + ly).
		// Per local row two srec ranges: bin lx = 0, then bins lx = 1..GB
		// (consecutive in a row, so contiguous in srec)
		if (tid < 64) {
			int cnt = 0;
			if (tid < NR) {
				const int ly = tid >> 1;
				const int gy = by0 - 1 + ly < 0 ? sby - 1 : by0 - 1 + ly;
				const int b = (tid & 1) ? bx0 + sbx * gy : (bx0 == 0 ? sbx - 1 : bx0 - 1) + sbx * gy;
				const int e = (tid & 1) ? o[b + GB] : o[b + 1];
				rs[tid] = o[b];
				cnt = e - o[b];
			}
			// inclusive scan over the wave
#pragma unroll
			for (int d = 1; d < 64; d <<= 1) {
				const int t = __shfl_up(cnt, d, 64);
				if (tid >= d) cnt += t;
			}
			if (tid < NR) rp[tid + 1] = cnt;
			if (tid == 0) rp[0] = 0;
		}
		__syncthreads();
		const int total = rp[NR];
		if (total == 0) {
			// no record reaches the block's cells (the block decides as one,
			// after its last barrier): zeros, or the field left as it is
			if (live && !accumulate) {
				double2* r2 = reinterpret_cast<double2*>((comp == 0 ? fx : fy) + (size_t) nx * (SB * by + j) + SB * bx);
				r2[0] = make_double2(0.0, 0.0);
				r2[1] = make_double2(0.0, 0.0);
			}
			return;
		}
		staged = total >= GMIN && total <= GCAP;
		if (staged) {
			for (int p = tid; p < total; p += 256) {
				int lo = 0, hi = NR - 1;   // last range with rp <= p
				while (lo < hi) {
					const int mid = (lo + hi + 1) >> 1;
					if (rp[mid] <= p) lo = mid; else hi = mid - 1;
				}
				const int g = rs[lo] + (p - rp[lo]);
				lkey[p] = srec[g].key;
				lidx[p] = g;
			}
		}
		__syncthreads();
	}
	if (!live) return;
	int head[4], end[4];
#pragma unroll
	for (int l = 0; l < 4; ++l) {
		if (staged) {
			const int lx = tx + (l & 1), ly = ty + (l >> 1);
			const int ri = 2 * ly + (lx > 0 ? 1 : 0);
			head[l] = rp[ri] + (bs[l] - rs[ri]);
		} else {
			head[l] = bs[l];
		}
		end[l] = head[l] + bn[l];
	}
	const int X = SB * bx, Yj = SB * by + j;
	double acc[SB];
#pragma unroll
	for (int i = 0; i < SB; ++i) acc[i] = 0.0;
	if (staged) merge_terms<true>(srec, lkey, lidx, head, end, X, Yj, (int) nx, ny, acc);
	else merge_terms<false>(srec, lkey, lidx, head, end, X, Yj, (int) nx, ny, acc);
	// the lane's 4 cells as two 16-byte pairs (fields are 16-byte aligned and
	// X is a multiple of 4)
	double2* r2 = reinterpret_cast<double2*>((comp == 0 ? fx : fy) + (size_t) nx * Yj + X);
	if (accumulate) {
		const double2 a = r2[0], b = r2[1];
		r2[0] = make_double2(a.x + acc[0], a.y + acc[1]);
		r2[1] = make_double2(b.x + acc[2], b.y + acc[3]);
	} else {
		r2[0] = make_double2(acc[0], acc[1]);
		r2[1] = make_double2(acc[2], acc[3]);
	}
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
	hipLaunchKernelGGL(k_interp, dim3((m + 255) / 256), dim3(256), 0, c->stream, comp, field, X, m,
	                   win_geom(c->n, c->W, c->D, c->nx, c->ny), U);
	AFM_LAUNCH_CHECK();
	return 0;
}

int launch_ib_move_half(Ctx* c, const double* u, const double* v, const double* X, long m, double half, double* Xh)
{
	if (m <= 0) return 0;
	hipLaunchKernelGGL(k_ib_move_half, dim3((m + 255) / 256), dim3(256), 0, c->stream, u, v, X, c->rec, m, c->n, c->nx,
	                   c->ny, half, Xh);
	AFM_LAUNCH_CHECK();
	return 0;
}

int launch_ib_move_full(Ctx* c, const double* uo, const double* vo, const double* u, const double* v,
                        const double* Xh, long m, double dt, double* X)
{
	if (m <= 0) return 0;
	(void) Xh;   // its windows are in c->rec
	hipLaunchKernelGGL(k_ib_move_full, dim3((m + 255) / 256), dim3(256), 0, c->stream, uo, vo, u, v, c->rec, m, c->n,
	                   c->nx, c->ny, dt, X);
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
	if (c->srec) { (void) hipFree(c->srec); c->srec = nullptr; }
	AFM_HIP(hipMalloc(&c->rec, sizeof(MarkerRec) * 2 * cap));
	AFM_HIP(hipMalloc(&c->bin_list, sizeof(int) * 2 * cap));
	AFM_HIP(hipMalloc(&c->srec, sizeof(BinRec) * 2 * cap));
	c->mcap = cap;
	return 0;
}

// the gather order's class counts and tile lists, both components
int spread_setup(Ctx* c)
{
	const size_t nt = (size_t) ((c->sbx + GB - 1) / GB) * ((c->sby + GB - 1) / GB);
	AFM_TRY(c->alloc((void**) &c->gorder, sizeof(int) * 2 * GCLS * (1 + nt)));
	AFM_TRY(c->alloc((void**) &c->tflag, sizeof(int) * 2 * nt));
	// the flags start unset once; each gather clears the ones it read
	AFM_HIP(hipMemsetAsync(c->tflag, 0, sizeof(int) * 2 * nt, c->stream));
	return 0;
}

int launch_spread(Ctx* c, int comp_mask, const double* Fx, const double* Fy, const double* ds, const double* X,
                  const double* R, const double* k, long m, double* fx, double* fy, int accumulate)
{
	const int nb1 = c->sbx * c->sby, nb = 2 * nb1;
	// chunk of 256 * items bins per scan block, at most SCAN_BLOCKS blocks
	const int items = (int) (((long) nb + 256L * SCAN_BLOCKS - 1) / (256L * SCAN_BLOCKS));
	const int nchunk = (int) (((long) nb + 256L * items - 1) / (256L * items));
	int* counts = c->bins;
	int* offs = counts + nb;           // nb + 1 entries
	int* cursor = offs + nb + 1;
	int* chunk = cursor + nb;          // nchunk entries
	const unsigned gblocks = (unsigned) (((c->sbx + GB - 1) / GB) * ((c->sby + GB - 1) / GB));
	// densest-first order where the tiles are staged (grids of >= 8 x 8 bins)
	// and markers are many per tile: its launch costs about 7 us, which the
	// reordering repays where clusters make long merges (262,144 markers on
	// 16,384 tiles: gather 91 -> 70 us), not with a few markers per tile
	// (16,384 on 4,096: 32 -> 27 us)
	const bool ordered = m >= 8L * (long) gblocks && c->gorder && c->sbx >= GB && c->sby >= GB;
	if (m > 0) {
		AFM_TRY(ensure_markers(c, m));
		dim3 g((m + 255) / 256, 2), g2((2 * m + 255) / 256, 2);
		hipLaunchKernelGGL(k_spread_prep, g2, dim3(256), 0, c->stream, comp_mask, Fx, Fy, ds, X, R, k, m,
		                   win_geom(c->n, c->W, c->D, c->nx, c->ny), c->sbx, c->sby, nb1, c->rec, counts, c->tflag,
		                   (c->sbx + GB - 1) / GB, (int) gblocks);
		AFM_LAUNCH_CHECK();
		hipLaunchKernelGGL(k_spread_chunk, dim3(nchunk), dim3(256), 0, c->stream, counts, chunk, nb, items);
		hipLaunchKernelGGL(k_spread_scan, dim3(nchunk), dim3(256), 0, c->stream, counts, chunk, offs, cursor, nb, items,
		                   ordered ? c->gorder : nullptr);
		AFM_LAUNCH_CHECK();
		hipLaunchKernelGGL(k_spread_fill, g, dim3(256), 0, c->stream, comp_mask, c->rec, m, nb1, offs, cursor,
		                   c->bin_list);
		AFM_LAUNCH_CHECK();
		hipLaunchKernelGGL(k_spread_rank, g, dim3(256), 0, c->stream, comp_mask, offs, c->bin_list, c->rec, m, nb1,
		                   c->srec);
		AFM_LAUNCH_CHECK();
	} else {
		// a kernel, not a memset, so that a captured graph replays it
		hipLaunchKernelGGL(k_zero_int, dim3((unsigned) ((nb + 1 + 255) / 256)), dim3(256), 0, c->stream, offs,
		                   (long) nb + 1);
		AFM_LAUNCH_CHECK();
	}
	if (ordered) {
		hipLaunchKernelGGL(k_gather_order, dim3((gblocks + 255) / 256, 2), dim3(256), 0, c->stream, comp_mask, offs,
		                   c->sbx, c->sby, nb1, (int) gblocks, c->gorder, c->gorder + 2 * GCLS);
		AFM_LAUNCH_CHECK();
	}
	hipLaunchKernelGGL(k_spread_gather, dim3(gblocks, 2), dim3(256), 0, c->stream, comp_mask, c->srec, offs,
	                   c->nx, c->sbx, c->sby, nb1, fx, fy, accumulate, ordered ? c->gorder : nullptr,
	                   ordered ? c->gorder + 2 * GCLS : nullptr, c->tflag);
	AFM_LAUNCH_CHECK();
	return 0;
}

}  // namespace afm
