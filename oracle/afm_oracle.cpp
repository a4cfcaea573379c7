// afm_oracle.cpp -- CPU restatement of the reference's IB hot path.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
// path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// may load it.  Nothing in afmibins_amd/ links, loads or calls it.
//
// Every function restates one reference routine (file:line under the
// reference's src/ tree) in the reference's own floating-point operation
// order, compiled with -ffp-contract=off and without -march=native so that no
// FMA contraction changes a rounding.
//
// Grid convention.  The reference is 3-D, x and z periodic, y walled, with
// fields stored x fastest: idx = x + nx*(y + ny*z) (ib/eulerian_grid.h:54-62).
// The doubly periodic 2-D mode used by the product is the reference's
// periodic (x, z) plane with y dropped; here it is stored as the 3-D layout
// with ny == 1, so "z" below is the 2-D "y".  Functions named *2 are the 2-D
// restatements, *3 the native 3-D ones (used to pin this oracle against the
// reference build in oracle/_ref).
//
// delta modes.  mode 0 evaluates the reference's delta exactly as compiled:
// x87 long-double argument, then glibc cos.  mode 1 replaces glibc cos (which
// is not correctly rounded: about 1.3e-3 of arguments in [0, 3.2) differ by
// one ulp on this host) with the correctly rounded cosine, computed here from
// libquadmath's cosq rounded once to double.  mode 1 is the build's
// definition of delta and the one the GPU must reproduce bit for bit.
//
// Pinning status (checked by tests/test_oracle_pins.py):
//   delta (mode 0), interp3, restriction3, prolong3, laplacian SpMV,
//   lexicographic smooth, CG, MG solve (lex), helmholtz_periodic3 CSR,
//   spread3 (up to summation order)  -- against the reference build;
//   2-D laplacian, smooth, CG, div, grad -- against the reference build with
//   ny == 1;  prolong2 -- against the wall rows of the reference's prolong;
//   restriction2, interp2/spread2, RB smoother, helmholtz2, advection2,
//   ns_step2, ib_step2 -- by closed forms and invariants (see the tests).
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>
#include <chrono>
#include <quadmath.h>

namespace {

typedef std::vector<double> vec;

const long double PI_L = 3.141592653589793238462643383279502884L;  // ib/grid.h:31

// ---------------------------------------------------------------- delta ----
// ib/grid.h:51-55.  The argument is formed in x87 long double and rounded to
// double on the way into cos().
double delta_arg(double r, unsigned n)
{
	long double a = PI_L * r;
	a = a * n;
	a = a / 2;
	return (double) a;
}

double cos_cr(double x) { return (double) cosq((__float128) x); }

double delta(double r, unsigned n, int mode)
{
	if (std::abs((int) (n * r)) > 2.)   // int abs, as resolved in the reference
		return 0;
	double a = delta_arg(r, n);
	double c = mode == 0 ? ::cos(a) : cos_cr(a);
	return 0.25 * n * (1 + c);
}

// ------------------------------------------------------------ vectors ------
// dt/vector.h:75-94: sequential dot product, RMS norm sqrt(sum)/sqrt(Size).
double dot(const vec& a, const vec& b)
{
	double r = 0.0;
	for (size_t i = 0; i < a.size(); ++i) r += a[i] * b[i];
	return r;
}

double rms(const vec& v)
{
	double mod = 0.0;
	for (double e : v) mod += e * e;
	return sqrt(mod) / sqrt((double) v.size());
}

// ------------------------------------------------------------ CSR ----------
struct Csr {
	size_t N = 0;
	std::vector<size_t> rs, col;
	vec val;
};

// dt/matrix.h:81-94: y[row] += a_ij * x[j] in stored order, y starting at 0.
vec spmv(const Csr& A, const vec& x)
{
	vec y(A.N, 0.0);
	for (size_t r = 0; r < A.N; ++r)
		for (size_t k = A.rs[r]; k < A.rs[r + 1]; ++k)
			y[r] += A.val[k] * x[A.col[k]];
	return y;
}

// alg/smooth.h:35-51: one in-place sweep of one row.
inline void gs_row(vec& x, const Csr& A, const vec& b, size_t row)
{
	double value = 0.0, diagonal = 0.0;
	for (size_t k = A.rs[row]; k < A.rs[row + 1]; ++k) {
		if (A.col[k] == row) { diagonal = A.val[k]; continue; }
		value += A.val[k] * x[A.col[k]];
	}
	x[row] = (b[row] - value) / diagonal;
}

void smooth_lex(vec& x, const Csr& A, const vec& b)
{
	for (size_t row = 0; row < A.N; ++row) gs_row(x, A, b, row);
}

// Red-black ordering of the same row update: red = (x + y + z) even, then
// black.  The per-row arithmetic is identical to alg/smooth.h:39-50; only the
// visiting order differs (the product's GPU smoother).
void smooth_rb(vec& x, const Csr& A, const vec& b, unsigned nx, unsigned ny, unsigned nz)
{
	for (unsigned color = 0; color < 2; ++color)
		for (size_t row = 0; row < A.N; ++row) {
			unsigned ix = row % nx, iy = (row % ((size_t) nx * ny)) / nx, iz = row / ((size_t) nx * ny);
			if (((ix + iy + iz) & 1u) == color) gs_row(x, A, b, row);
		}
}

// alg/cg.h:32-56.  x and b may alias (mg/solver.h:270 passes the same vector).
unsigned cg(vec& x, const Csr& A, const vec& bin, double tol, unsigned max_iter, double* final_norm)
{
	vec b = bin;
	vec Ax = spmv(A, x);
	vec r(A.N), p(A.N), z;
	for (size_t i = 0; i < A.N; ++i) r[i] = b[i] - Ax[i];
	p = r;
	double n0 = dot(r, r);
	unsigned count = 0;
	while (sqrt(n0) > tol && count < max_iter) {
		z = spmv(A, p);
		double nu = n0 / dot(p, z);
		for (size_t i = 0; i < A.N; ++i) x[i] += p[i] * nu;
		for (size_t i = 0; i < A.N; ++i) z[i] *= nu;
		for (size_t i = 0; i < A.N; ++i) r[i] -= z[i];
		double n1 = dot(r, r);
		double mu = n1 / n0;
		n0 = n1;
		for (size_t i = 0; i < A.N; ++i) p[i] *= mu;
		for (size_t i = 0; i < A.N; ++i) p[i] += r[i];
		++count;
	}
	if (final_norm) *final_norm = sqrt(n0);
	return count;
}

// ------------------------------------------------------- laplacian ---------
// mg/laplacian.h:46-153 (and, with a unit diagonal shift and scale, the
// periodic Helmholtz ns/helmholtz.h:61-175 when wall == false).  Entries are
// emitted in the reference's order, which is ascending column index.
// kind 0: laplacian; kind 1: helmholtz_periodic (3-D walls, ny > 1 closure);
// kind 2: the 2-D doubly periodic Helmholtz I + s*L5 (ny == 1, no wall);
// kind 3: helmholtz_dirichlet (ns/helmholtz.h:218-319) of the v grid, called
// with ny = the cell grid's ny - 1 (its rows are the interior y faces).
Csr stencil_csr(unsigned nx, unsigned ny, unsigned nz, unsigned n, int kind, double scale)
{
	Csr A;
	A.N = (size_t) nx * ny * nz;
	A.rs.assign(A.N + 1, 0);
	A.col.reserve(A.N * 7);
	A.val.reserve(A.N * 7);
	for (size_t i = 0; i < A.N; ++i) {
		A.rs[i] = A.val.size();
		unsigned x = i % nx, y = (i % ((size_t) nx * ny)) / nx, z = i / ((size_t) nx * ny);
		double off, diagonal;
		if (kind == 0) { off = n * n; diagonal = -6.0 * off; }
		else if (kind == 3) { off = scale * n * n; diagonal = 1. - scale * 6. * (double) (n * n); }
		else { off = scale * n * n; diagonal = kind == 2 ? 1. - 4. * off : 1. - 6. * off; }
		double left = off, right = off, top = off, bottom = off, front = off, back = off;
		if (nx == 1) { diagonal += left + right; left = right = 0; }
		else if (nx == 2) {
			if (x == 0) { right += left; left = 0; } else { left += right; right = 0; }
		}
		if (kind == 0) {
			if (y == 0) { diagonal += bottom; bottom = 0; }
			if (y == ny - 1) { diagonal += top; top = 0; }
		} else if (kind == 1) {
			if (ny == 1) { diagonal -= 3 * (top + bottom); top = bottom = 0; }
			else if (y == 0) { diagonal -= 2 * bottom; top += off / 3; bottom = 0; }
			else if (y == ny - 1) { diagonal -= 2 * top; bottom += off / 3; top = 0; }
		} else if (kind == 3) {
			if (y == 0) bottom = 0;          // the wall values enter through boundary()
			if (y == ny - 1) top = 0;
		} else {
			top = bottom = 0;   // 2-D: the y axis does not exist
		}
		if (nz == 1) { diagonal += back + front; back = front = 0; }
		else if (nz == 2) {
			if (z == 0) { front += back; back = 0; } else { back += front; front = 0; }
		}
		auto put = [&](double v, size_t c) { A.val.push_back(v); A.col.push_back(c); };
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
	A.rs[A.N] = A.val.size();
	return A;
}

// ------------------------------------------------- transfer operators ------
// mg/solver.h:153-181 (3-D, y clamped) -- output accumulates in tap order.
void restrict3(vec& out, const vec& in, unsigned nx, unsigned ny, unsigned nz)
{
	unsigned cx = nx / 2, cy = ny / 2, cz = nz / 2;
	out.assign((size_t) cx * cy * cz, 0.0);
	for (unsigned xcc = 0; xcc < cx; ++xcc)
		for (unsigned ycc = 0; ycc < cy; ++ycc)
			for (unsigned zcc = 0; zcc < cz; ++zcc) {
				unsigned xcf = 2 * xcc + 1, ycf = 2 * ycc + 1, zcf = 2 * zcc + 1;
				double& o = out[xcc + (size_t) cx * (ycc + (size_t) cy * zcc)];
				for (int i = 0; i < 64; ++i) {
					int dx = (i % 4) - 2, dy = (i % 16) / 4 - 2, dz = i / 16 - 2;
					unsigned x = (xcf + dx + nx) % nx;
					unsigned y = std::min(ycf + std::max(dy, -((int) ycf)), ny - 1);
					unsigned z = (zcf + dz + nz) % nz;
					int w = (4 - abs(2 * dx + 1)) * (4 - abs(2 * dy + 1)) * (4 - abs(2 * dz + 1));
					o += w / 512. * in[x + (size_t) nx * (y + (size_t) ny * z)];
				}
			}
}

// 2-D analogue of mg/solver.h:166-176: 16 taps over the periodic (x, z) plane,
// weights (1,3,3,1)x(1,3,3,1)/64, dx fastest.
void restrict2(vec& out, const vec& in, unsigned nx, unsigned nz)
{
	unsigned cx = nx / 2, cz = nz / 2;
	out.assign((size_t) cx * cz, 0.0);
	for (unsigned xcc = 0; xcc < cx; ++xcc)
		for (unsigned zcc = 0; zcc < cz; ++zcc) {
			unsigned xcf = 2 * xcc + 1, zcf = 2 * zcc + 1;
			double& o = out[xcc + (size_t) cx * zcc];
			for (int i = 0; i < 16; ++i) {
				int dx = (i % 4) - 2, dz = i / 4 - 2;
				unsigned x = (xcf + dx + nx) % nx, z = (zcf + dz + nz) % nz;
				int w = (4 - abs(2 * dx + 1)) * (4 - abs(2 * dz + 1));
				o += w / 64. * in[x + (size_t) nx * z];
			}
		}
}

// mg/solver.h:71-151: trilinear interior rows, bilinear wall rows (including
// the reference's ny-for-nx index in the top wall row, mg/solver.h:136-139).
// Overwrites the rows it visits.
void prolong3(vec& out, const vec& in, unsigned nx, unsigned ny, unsigned nz)
{
	unsigned hx = nx / 2, hy = ny / 2, hz = nz / 2;
	for (unsigned xcc = 0; xcc < hx; ++xcc) {
		for (unsigned ycc = 0; ycc + 1 < hy; ++ycc)
			for (unsigned zcc = 0; zcc < hz; ++zcc) {
				unsigned xnc = (xcc + 1) % hx, ync = ycc + 1, znc = (zcc + 1) % hz;
				unsigned xcf = 2 * xcc + 1, ycf = 2 * ycc + 1, zcf = 2 * zcc + 1;
				unsigned xnf = (xcf + 1) % nx, ynf = ycf + 1, znf = (zcf + 1) % nz;
				auto C = [&](unsigned a, unsigned b, unsigned c) { return in[a + (size_t) hx * (b + (size_t) hy * c)]; };
				auto F = [&](unsigned a, unsigned b, unsigned c) -> double& { return out[a + (size_t) nx * (b + (size_t) ny * c)]; };
				double blbc = C(xcc, ycc, zcc), brbc = C(xnc, ycc, zcc), ulbc = C(xcc, ync, zcc), urbc = C(xnc, ync, zcc),
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
		for (unsigned zcc = 0; zcc < hz; ++zcc) {
			unsigned xnc = (xcc + 1) % hx, znc = (zcc + 1) % hz;
			unsigned xcf = 2 * xcc + 1, zcf = 2 * zcc + 1;
			unsigned xnf = (xcf + 1) % nx, znf = (zcf + 1) % nz;
			double blbc = in[xcc + (size_t) hx * hy * zcc], brbc = in[xnc + (size_t) hx * hy * zcc],
			       blfc = in[xcc + (size_t) hx * hy * znc], brfc = in[xnc + (size_t) hx * hy * znc],
			       ulbc = in[xcc + (size_t) hx * (hy - 1 + (size_t) hy * zcc)], urbc = in[xnc + (size_t) hx * (hy - 1 + (size_t) hy * zcc)],
			       ulfc = in[xcc + (size_t) hx * (hy - 1 + (size_t) hy * znc)], urfc = in[xnc + (size_t) hx * (hy - 1 + (size_t) hy * znc)];
			out[xcf + (size_t) nx * ny * zcf] = 9. / 16. * blbc + 3. / 16. * (brbc + blfc) + 1. / 16. * brfc;
			out[xnf + (size_t) nx * ny * zcf] = 9. / 16. * brbc + 3. / 16. * (blbc + brfc) + 1. / 16. * blfc;
			out[xcf + (size_t) nx * ny * znf] = 9. / 16. * blfc + 3. / 16. * (brfc + blbc) + 1. / 16. * brbc;
			out[xnf + (size_t) nx * ny * znf] = 9. / 16. * brfc + 3. / 16. * (blfc + brbc) + 1. / 16. * blbc;
			out[xcf + (size_t) ny * (ny - 1 + (size_t) ny * zcf)] = 9. / 16. * ulbc + 3. / 16. * (urbc + ulfc) + 1. / 16. * urfc;
			out[xnf + (size_t) ny * (ny - 1 + (size_t) ny * zcf)] = 9. / 16. * urbc + 3. / 16. * (ulbc + urfc) + 1. / 16. * ulfc;
			out[xcf + (size_t) ny * (ny - 1 + (size_t) ny * znf)] = 9. / 16. * ulfc + 3. / 16. * (urfc + ulbc) + 1. / 16. * urbc;
			out[xnf + (size_t) ny * (ny - 1 + (size_t) ny * znf)] = 9. / 16. * urfc + 3. / 16. * (ulfc + urbc) + 1. / 16. * ulbc;
		}
	}
}

// 2-D analogue: exactly the wall-row bilinear formula of mg/solver.h:141-148
// on the periodic (x, z) plane.
void prolong2(vec& out, const vec& in, unsigned nx, unsigned nz)
{
	unsigned hx = nx / 2, hz = nz / 2;
	for (unsigned xcc = 0; xcc < hx; ++xcc)
		for (unsigned zcc = 0; zcc < hz; ++zcc) {
			unsigned xnc = (xcc + 1) % hx, znc = (zcc + 1) % hz;
			unsigned xcf = 2 * xcc + 1, zcf = 2 * zcc + 1;
			unsigned xnf = (xcf + 1) % nx, znf = (zcf + 1) % nz;
			double blbc = in[xcc + (size_t) hx * zcc], brbc = in[xnc + (size_t) hx * zcc],
			       blfc = in[xcc + (size_t) hx * znc], brfc = in[xnc + (size_t) hx * znc];
			out[xcf + (size_t) nx * zcf] = 9. / 16. * blbc + 3. / 16. * (brbc + blfc) + 1. / 16. * brfc;
			out[xnf + (size_t) nx * zcf] = 9. / 16. * brbc + 3. / 16. * (blbc + brfc) + 1. / 16. * blfc;
			out[xcf + (size_t) nx * znf] = 9. / 16. * blfc + 3. / 16. * (brfc + blbc) + 1. / 16. * brbc;
			out[xnf + (size_t) nx * znf] = 9. / 16. * brfc + 3. / 16. * (blfc + brbc) + 1. / 16. * blbc;
		}
}

// ------------------------------------------------------------ multigrid ----
// mg/solver.h:35-271, with the level hierarchy built at run time.
struct Level { unsigned n, nx, ny, nz; Csr L; };

struct MG {
	std::vector<Level> lv;   // lv[0] coarsest
	double tol;
	unsigned nu1, nu2;
	int dim, smoother;       // smoother 0 lexicographic (reference), 1 red-black
	unsigned cg_max = 100000;

	MG(int dim_, unsigned levels, unsigned W, unsigned H, unsigned D, double tol_, unsigned nu1_, unsigned nu2_, int sm)
		: tol(tol_), nu1(nu1_), nu2(nu2_), dim(dim_), smoother(sm)
	{
		for (unsigned l = 0; l <= levels; ++l) {
			Level L;
			L.n = 1u << l;
			L.nx = W * L.n; L.ny = dim == 3 ? H * L.n : 1; L.nz = D * L.n;
			L.L = stencil_csr(L.nx, L.ny, L.nz, L.n, 0, 0.0);
			lv.push_back(std::move(L));
		}
	}
	void smooth(unsigned l, vec& x, const vec& b)
	{
		const Level& L = lv[l];
		if (smoother == 0) smooth_lex(x, L.L, b);
		else smooth_rb(x, L.L, b, L.nx, L.ny, L.nz);
	}
	void restrict_(unsigned l, vec& coarse, const vec& fine)
	{
		const Level& L = lv[l];
		if (dim == 3) restrict3(coarse, fine, L.nx, L.ny, L.nz); else restrict2(coarse, fine, L.nx, L.nz);
	}
	void prolong(unsigned l, vec& fine, const vec& coarse)
	{
		const Level& L = lv[l];
		if (dim == 3) prolong3(fine, coarse, L.nx, L.ny, L.nz); else prolong2(fine, coarse, L.nx, L.nz);
	}
	// mg/solver.h:264-271
	unsigned solve0(vec& input)
	{
		const Level& L = lv[0];
		vec ones(input.size(), 1.0);
		double c = dot(ones, input) / (L.nx * L.ny * L.nz);
		for (double& e : input) e -= 1.0 * c;
		return cg(input, L.L, input, tol, cg_max, nullptr);
	}
	// mg/solver.h:183-197
	void nested(unsigned l, vec& input)
	{
		if (l == 0) return;
		vec copy(input), coarse;
		restrict_(l, coarse, input);
		nested(l - 1, coarse);
		recurse(l - 1, coarse);
		prolong(l, input, coarse);
		for (unsigned j = 0; j < nu2; ++j) smooth(l, input, copy);
	}
	// mg/solver.h:199-216
	void recurse(unsigned l, vec& rhs)
	{
		if (l == 0) { solve0(rhs); return; }
		vec x(rhs.size(), 0.0), residualc;
		for (unsigned j = 0; j < nu1; ++j) smooth(l, x, rhs);
		vec Lx = spmv(lv[l].L, x);
		vec residualf(rhs.size());
		for (size_t i = 0; i < rhs.size(); ++i) residualf[i] = rhs[i] - Lx[i];
		restrict_(l, residualc, residualf);
		recurse(l - 1, residualc);
		prolong(l, residualf, residualc);
		for (size_t i = 0; i < x.size(); ++i) x[i] += residualf[i];
		for (unsigned j = 0; j < nu2; ++j) smooth(l, x, rhs);
		rhs = x;
	}
	// mg/solver.h:218-237; max_cycles bounds the loop (the reference has none).
	unsigned solve(vec& rhs, unsigned max_cycles, double* final_rms)
	{
		unsigned top = lv.size() - 1;
		if (top == 0) return solve0(rhs);
		vec x(rhs);
		nested(top, x);
		vec ones(rhs.size(), 1.0);
		double c = dot(rhs, ones) / (double) rhs.size();
		for (double& e : rhs) e -= c * 1.0;
		vec residual(rhs.size());
		auto resid = [&]() {
			vec Lx = spmv(lv[top].L, x);
			for (size_t i = 0; i < rhs.size(); ++i) residual[i] = rhs[i] - Lx[i];
		};
		resid();
		unsigned count = 0;
		double r = rms(residual);
		while (r > tol && count < max_cycles) {
			recurse(top, residual);
			for (size_t i = 0; i < x.size(); ++i) x[i] += residual[i];
			resid();
			r = rms(residual);
			++count;
		}
		if (final_rms) *final_rms = r;
		rhs = x;
		return count;
	}
};

// ---------------------------------------------------------- IB windows -----
// 2-D window of ib/eulerian_grid.h:64-93 as used through the generic
// ib/mac_grid.h:49-73 (the reference's x and z components).  comp 0 (u,
// x-faces) is shifted in the 2-D y axis, comp 1 (v, y-faces) in x.
// Returns the 16 (index, weight) pairs in tap order.
void window2(int comp, double cx, double cy, unsigned n, unsigned W, unsigned D, unsigned nx, unsigned ny,
             int mode, size_t idx[16], double w[16])
{
	double q0 = cx, q1 = cy;
	if (comp == 0) q1 -= 0.5 / n; else q0 -= 0.5 / n;
	q0 = std::fmod(std::fmod(q0, (double) W) + W, (double) W);
	q1 = std::fmod(std::fmod(q1, (double) D) + D, (double) D);
	const double fx = std::floor(q0 * n), fy = std::floor(q1 * n);
	// delta depends on the tap only through dix (x factor) or diy (y factor);
	// evaluating the 4 + 4 distinct factors once gives the same bits as the
	// reference's per-tap evaluation.
	double dxs[4], dys[4];
	for (int d = 0; d < 4; ++d) {
		double x = (fx + d - 1.0) / n, y = (fy + d - 1.0) / n;
		if (comp == 0) y += 0.5 / n; else x += 0.5 / n;
		dxs[d] = delta(x - cx, n, mode);
		dys[d] = delta(y - cy, n, mode);
	}
	for (int i = 0; i < 16; ++i) {
		const int dix = i % 4, diy = i / 4;
		const size_t ix = ((size_t) fx + dix + nx - 1) % nx, iy = ((size_t) fy + diy + ny - 1) % ny;
		idx[i] = ix + (size_t) nx * iy;
		w[i] = dxs[dix] * dys[diy];
	}
}

// 3-D windows: generic components ib/mac_grid.h:49-73 + eulerian search
// ib/eulerian_grid.h:64-93; v component ib/mac_grid.h:97-136.  Entries
// skipped by the reference's y filter get idx = SIZE_MAX.
size_t sz_of(double f) { return (size_t) (int64_t) f; }

void window3(int comp, const double c[3], unsigned n, unsigned W, unsigned H, unsigned D,
             unsigned nx, unsigned ny, unsigned nz, int mode, size_t idx[64], double w[64])
{
	double q[3] = {c[0], c[1], c[2]};
	if (comp == 0 || comp == 2) {      // comp 3: the node grid, unshifted
		q[(comp + 1) % 3] -= 0.5 / n;
		q[(comp + 2) % 3] -= 0.5 / n;
	}
	q[0] = std::fmod(std::fmod(q[0], (double) W) + W, (double) W);
	q[2] = std::fmod(std::fmod(q[2], (double) D) + D, (double) D);
	double fx, fy, fz;
	if (comp == 1) { fx = std::floor(q[0] * n - 0.5); fy = std::floor(q[1] * n - 1.0); fz = std::floor(q[2] * n - 0.5); }
	else { fx = std::floor(q[0] * n); fy = std::floor(q[1] * n); fz = std::floor(q[2] * n); }
	for (int i = 0; i < 64; ++i) {
		const int dix = i % 4, diy = (i / 4) % 4, diz = (i / 16) % 4;
		const size_t ix = (sz_of(fx) + dix + nx - 1) % nx, iy = sz_of(fy) + diy, iz = (sz_of(fz) + diz + nz - 1) % nz;
		double p[3];
		p[0] = (fx + dix - 1.0) / n;
		p[1] = comp == 1 ? (fy + diy) / n : (fy + diy - 1.0) / n;
		p[2] = (fz + diz - 1.0) / n;
		if (p[1] < 0 || p[1] > H) { idx[i] = SIZE_MAX; w[i] = 0; continue; }
		// Taps the grid does not store, which the reference indexes out of
		// range (undefined; its own tests keep markers 3.5 cells from the
		// walls): the v window's wall faces y = 0 and y = H (rows -1 and
		// ny - 1 of the (ny-1)-row array), and the u / w windows' ghost row
		// above the top wall (y = H + h/2, row ny; the test above keeps
		// p[1] == H, while the bottom ghost row is dropped by p[1] < 0).
		// Here such taps read nothing and receive nothing (zero wall velocity
		// and no ghost cells).
		if (comp == 1 ? (iy == 0 || iy == ny) : iy > ny) { idx[i] = SIZE_MAX; w[i] = 0; continue; }
		if (comp == 1) {
			p[0] = p[0] + 0.5 / n; p[2] = p[2] + 0.5 / n;
			idx[i] = ix + (size_t) nx * (iy - 1 + (size_t) (ny - 1) * iz);
		} else {
			idx[i] = ix + (size_t) nx * (iy - 1 + (size_t) ny * iz);
			if (comp != 3) {
				p[(comp + 1) % 3] += 0.5 / n;
				p[(comp + 2) % 3] += 0.5 / n;
			}
		}
		w[i] = delta(p[0] - c[0], n, mode) * delta(p[1] - c[1], n, mode) * delta(p[2] - c[2], n, mode);
	}
}

// ------------------------------------------------------ 2-D NS pieces -----
struct G2 { unsigned n, nx, ny, W, D; };   // ny here is the 2-D y (reference z)

// alg/div.h:29-60 with ny == 1 (dv = top - bottom = 0 - 0).
void div2(vec& d, const vec& u, const vec& w, const G2& g)
{
	const double dv = 0.0 - 0.0;
	for (size_t i = 0; i < d.size(); ++i) {
		unsigned x = i % g.nx, z = i / g.nx;
		size_t ixp = (x + 1) % g.nx + (size_t) g.nx * z, izp = x + (size_t) g.nx * ((z + 1) % g.ny);
		d[i] = g.n * ((u[ixp] - u[i]) + dv + (w[izp] - w[i]));
	}
}

// alg/grad.h:29-44 with ny == 1.
void grad2(vec& px, vec& pz, const vec& p, const G2& g)
{
	for (size_t i = 0; i < p.size(); ++i) {
		unsigned x = i % g.nx, z = i / g.nx;
		size_t ixm = (x + g.nx - 1) % g.nx + (size_t) g.nx * z, izm = x + (size_t) g.nx * ((z + g.ny - 1) % g.ny);
		px[i] = g.n * (p[i] - p[ixm]);
		pz[i] = g.n * (p[i] - p[izm]);
	}
}

// ns/solver.h:204-234 (the ny == 1 branch), w renamed to the 2-D v.
void advect2(vec& uadv, vec& wadv, const vec& u, const vec& w, const G2& g)
{
	size_t N = u.size();
	vec uu(N), ww(N), uw(N);
	for (size_t i = 0; i < N; ++i) {
		unsigned x = i % g.nx, z = i / g.nx;
		size_t ixm = (x + g.nx - 1) % g.nx + (size_t) g.nx * z, izm = x + (size_t) g.nx * ((z + g.ny - 1) % g.ny);
		double q = 0.5 * (u[i] + u[ixm]);
		uu[i] = q * q;
		q = 0.5 * (w[i] + w[izm]);
		ww[i] = q * q;
		q = 0.5 * (u[izm] + u[i]);
		double r = 0.5 * (w[ixm] + w[i]);
		uw[i] = q * r;
	}
	div2(uadv, uu, uw, g);
	div2(wadv, uw, ww, g);
}

// ns/solver.h:37-140 restricted to the doubly periodic plane.  One dt is two
// calls: backward-Euler half step, then Crank-Nicolson full step from u_old.
struct NS2 {
	G2 g;
	double mu, rho, dt, tol;
	int tol_mode;              // 0 absolute (reference), 1 relative
	unsigned nu1 = 2, nu2 = 2; // ns/projection.h:44
	int smoother;
	Csr phhm, phhs;
	vec *u, *w, u_old, w_old;
	bool half = false;
	unsigned levels;
	unsigned last_cycles = 0, last_cg = 0;
	MG* mg = nullptr;   // per-level laplacians built once (mg/solver.h:68-69)

	NS2(const G2& g_, double mu_, double rho_, double k, double tol_, int tm, int sm, vec* u0, vec* w0)
		: g(g_), mu(mu_), rho(rho_), dt(k), tol(tol_), tol_mode(tm), smoother(sm), u(u0), w(w0)
	{
		phhm = stencil_csr(g.nx, 1, g.ny, g.n, 2, mu * k / (2 * rho));
		phhs = stencil_csr(g.nx, 1, g.ny, g.n, 2, -mu * k / (2 * rho));
		levels = 0;
		while ((1u << levels) < g.n) ++levels;
		mg = new MG(2, levels, g.W, 1, g.D, tol, nu1, nu2, smoother);
		project(*u, *w);
	}
	~NS2() { delete mg; }
	// ns/projection.h:60-73
	void project(vec& uu, vec& ww)
	{
		vec phi(uu.size());
		div2(phi, uu, ww, g);
		double t = tol;
		if (tol_mode == 1) {
			vec ones(phi.size(), 1.0), tmp(phi);
			double c = dot(tmp, ones) / (double) tmp.size();
			for (double& e : tmp) e -= c * 1.0;
			t = tol * rms(tmp);
		}
		mg->tol = t;
		last_cycles = mg->solve(phi, 10000, nullptr);
		vec px(phi.size()), pz(phi.size());
		grad2(px, pz, phi, g);
		for (size_t i = 0; i < uu.size(); ++i) uu[i] -= px[i];
		for (size_t i = 0; i < ww.size(); ++i) ww[i] -= pz[i];
		p_last = phi;
	}
	vec p_last;
	unsigned cg_solve(vec& x, const vec& b)
	{
		double t = tol;
		if (tol_mode == 1) t = tol * sqrt(dot(b, b));
		return cg(x, phhs, b, t, 100000, nullptr);
	}
	void step(const vec& fx, const vec& fz)
	{
		size_t N = u->size();
		vec uadv(N), wadv(N), rx(N), rz(N);
		advect2(uadv, wadv, *u, *w, g);
		if (!half) {
			u_old = *u; w_old = *w;
			for (size_t i = 0; i < N; ++i) {
				rx[i] = u_old[i] + (dt / 2) * (1 / rho * fx[i] - uadv[i]);
				rz[i] = w_old[i] + (dt / 2) * (1 / rho * fz[i] - wadv[i]);
			}
		} else {
			vec hu = spmv(phhm, u_old), hw = spmv(phhm, w_old);
			for (size_t i = 0; i < N; ++i) {
				rx[i] = hu[i] + dt * (1 / rho * fx[i] - uadv[i]);
				rz[i] = hw[i] + dt * (1 / rho * fz[i] - wadv[i]);
			}
		}
		last_cg = cg_solve(*u, rx);
		last_cg += cg_solve(*w, rz);
		project(*u, *w);
		half = !half;
	}
};

// ------------------------------------------- 3-D wall-bounded mode ---------
// The reference's native geometry: nx x ny x nz cells, periodic in x and z,
// walls in y; u, w on the nx*ny*nz x- and z-faces, v on the nx*(ny-1)*nz
// interior y-faces, wall values in nx*nz vectors (index x + nx*z).

// alg/div.h:29-58 (v has nx*(ny-1)*nz entries, top/bottom nx*nz).
void div3(vec& d, const vec& u, const vec& v, const vec& w, const vec& top, const vec& bottom,
          unsigned nx, unsigned ny, unsigned nz, unsigned n)
{
	const size_t N = (size_t) nx * ny * nz;
	d.assign(N, 0.0);
	for (size_t i = 0; i < N; ++i) {
		unsigned x = i % nx, y = (i % ((size_t) nx * ny)) / nx, z = i / ((size_t) nx * ny);
		size_t ixz = x + (size_t) nx * z;
		double dv;
		if (ny > 1) {
			if (y == 0) dv = v[i - (size_t) nx * z] - bottom[ixz];
			else if (y == ny - 1) dv = top[ixz] - v[i - (size_t) nx * (z + 1)];
			else dv = v[i - (size_t) nx * z] - v[i - (size_t) nx * (z + 1)];
		} else {
			dv = top[ixz] - bottom[ixz];
		}
		size_t ixp = (x + 1) % nx + (size_t) nx * (y + (size_t) ny * z);
		size_t izp = x + (size_t) nx * (y + (size_t) ny * ((z + 1) % nz));
		d[i] = n * ((u[ixp] - u[i]) + dv + (w[izp] - w[i]));
	}
}

// alg/grad.h:32-43.
void grad3(vec& px, vec& py, vec& pz, const vec& p, unsigned nx, unsigned ny, unsigned nz, unsigned n)
{
	const size_t N = (size_t) nx * ny * nz;
	px.assign(N, 0.0);
	pz.assign(N, 0.0);
	py.assign((size_t) nx * (ny - 1) * nz, 0.0);
	for (size_t i = 0; i < N; ++i) {
		unsigned x = i % nx, y = (i % ((size_t) nx * ny)) / nx, z = i / ((size_t) nx * ny);
		size_t ixm = (x + nx - 1) % nx + (size_t) nx * (y + (size_t) ny * z);
		size_t izm = x + (size_t) nx * (y + (size_t) ny * ((z + nz - 1) % nz));
		px[i] = n * (p[i] - p[ixm]);
		pz[i] = n * (p[i] - p[izm]);
		if (y > 0) py[i - (size_t) nx * (z + 1)] = n * (p[i] - p[i - nx]);
	}
}

// ns/solver.h:142-202 (the ny > 1 branch): centred products on the cell
// grid and the edges, then three divergences; the v equation's divergence
// runs on the (ny-1)-row grid with the wall v^2 values as top and bottom.
void advect3(vec& uadv, vec& vadv, vec& wadv, const vec& u, const vec& v, const vec& w,
             unsigned nx, unsigned ny, unsigned nz, unsigned n)
{
	const size_t N = (size_t) nx * ny * nz, Nv = (size_t) nx * (ny - 1) * nz, Nxz = (size_t) nx * nz;
	vec uu(N), ww(N), uw(N), uv(Nv), vw(Nv), vv((size_t) nx * (ny - 2) * nz), vvt(Nxz), vvb(Nxz), zero(Nxz, 0.0);
	for (size_t i = 0; i < N; ++i) {
		unsigned x = i % nx, y = (i % ((size_t) nx * ny)) / nx, z = i / ((size_t) nx * ny);
		size_t ixz = x + (size_t) nx * z, iv = i - (size_t) nx * z;
		size_t ixm = (x + nx - 1) % nx + (size_t) nx * (y + (size_t) ny * z);
		size_t izm = x + (size_t) nx * (y + (size_t) ny * ((z + nz - 1) % nz));
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
			size_t ivxm = (x + nx - 1) % nx + (size_t) nx * (y + (size_t) (ny - 1) * z);
			size_t ivzm = x + (size_t) nx * (y + (size_t) (ny - 1) * ((z + nz - 1) % nz));
			q = 0.5 * (u[i + nx] + u[i]);
			double r = 0.5 * (v[ivxm] + v[iv]);
			uv[iv] = q * r;
			q = 0.5 * (v[ivzm] + v[iv]);
			r = 0.5 * (w[i + nx] + w[i]);
			vw[iv] = q * r;
		}
		q = 0.5 * (u[izm] + u[i]);
		double r = 0.5 * (w[ixm] + w[i]);
		uw[i] = q * r;
	}
	div3(uadv, uu, uv, uw, zero, zero, nx, ny, nz, n);
	div3(vadv, uv, vv, vw, vvt, vvb, nx, ny - 1, nz, n);
	div3(wadv, uw, vw, ww, zero, zero, nx, ny, nz, n);
}

// helmholtz_periodic::boundary (ns/helmholtz.h:177-196, ny > 1) and
// helmholtz_dirichlet::boundary (ns/helmholtz.h:321-334): the wall values
// into the first and last rows of a right-hand side.
vec hbound3(int kind, const vec& top, const vec& bottom, double scale, unsigned nx, unsigned ny, unsigned nz, unsigned n)
{
	if (kind == 1) {
		vec b((size_t) nx * ny * nz, 0.0);
		for (size_t i = 0; i < (size_t) nx * nz; ++i) {
			unsigned x = i % nx, z = i / nx;
			b[x + (size_t) nx * ny * z] = 8. / 3. * scale * n * n * bottom[i];
			b[x + (size_t) nx * (ny - 1 + (size_t) ny * z)] = 8. / 3. * scale * n * n * top[i];
		}
		return b;
	}
	vec b((size_t) nx * (ny - 1) * nz, 0.0);
	for (size_t i = 0; i < (size_t) nx * nz; ++i) {
		unsigned x = i % nx, z = i / nx;
		b[x + (size_t) nx * (ny - 1) * z] = scale * n * n * bottom[i];
		b[x + (size_t) nx * (ny - 2 + (size_t) (ny - 1) * z)] = scale * n * n * top[i];
	}
	return b;
}

// projection::boundary (ns/projection.h:75-106, ny > 1), as written: the
// loop runs over nx*ny wall points and the z-component terms read phidtx at
// the second row (pzb2, pzt2); only cubes (ny == nz) keep it in bounds.
void proj_boundary3(vec& ut, vec& ub, vec& wt, vec& wb, const vec& phx, const vec& phz, unsigned nx, unsigned ny)
{
	for (size_t i = 0; i < (size_t) nx * ny; ++i) {
		unsigned x = i % nx, z = i / nx;
		size_t ixzb = x + (size_t) nx * ny * z, ixzt = x + (size_t) nx * ((size_t) ny * (z + 1) - 1);
		double pxb1 = phx[ixzb], pxb2 = phx[ixzb + nx], pxt1 = phx[ixzt], pxt2 = phx[ixzt - nx],
		       pzb1 = phz[ixzb], pzb2 = phx[ixzb + nx], pzt1 = phz[ixzt], pzt2 = phx[ixzt - nx];
		ut[i] += (9 * pxt1 - pxt2) / 8;
		ub[i] += (9 * pxb1 - pxb2) / 8;
		wt[i] += (9 * pzt1 - pzt2) / 8;
		wb[i] += (9 * pzb1 - pzb2) / 8;
	}
}

// ns::solver<geometry<L,1,1,1>> (ns/solver.h:37-140, ns/projection.h:60-73):
// n^3 cells; the multigrid and the CG stop on absolute tolerances as in the
// reference.  smoother 0 is the reference's lexicographic sweep, 1 red-black.
struct NS3 {
	unsigned n, levels;
	size_t N, Nv;
	double mu, rho, dt, tol;
	Csr phhm, phhs, dhhm, dhhs;
	vec *u, *v, *w, u_old, v_old, w_old, phx, phy, phz, phi;
	bool half = false;
	unsigned last_cycles = 0, last_cg = 0;
	MG* mg = nullptr;

	NS3(unsigned L, double mu_, double rho_, double k, double tol_, int smoother, vec* u0, vec* v0, vec* w0)
		: n(1u << L), levels(L), mu(mu_), rho(rho_), dt(k), tol(tol_), u(u0), v(v0), w(w0)
	{
		N = (size_t) n * n * n;
		Nv = (size_t) n * (n - 1) * n;
		phhm = stencil_csr(n, n, n, n, 1, mu * k / (2 * rho));
		phhs = stencil_csr(n, n, n, n, 1, -mu * k / (2 * rho));
		dhhm = stencil_csr(n, n - 1, n, n, 3, mu * k / (2 * rho));
		dhhs = stencil_csr(n, n - 1, n, n, 3, -mu * k / (2 * rho));
		phx.assign(N, 0.0); phz.assign(N, 0.0); phi.assign(N, 0.0); phy.assign(Nv, 0.0);
		mg = new MG(3, L, 1, 1, 1, tol, 2, 2, smoother);
		project(*u, *v, *w);
	}
	~NS3() { delete mg; }
	NS3(const NS3&) = delete;
	NS3& operator=(const NS3&) = delete;
	void project(vec& uu, vec& vv, vec& ww)
	{
		vec zero((size_t) n * n, 0.0);
		div3(phi, uu, vv, ww, zero, zero, n, n, n, n);
		last_cycles = mg->solve(phi, 100000, nullptr);
		grad3(phx, phy, phz, phi, n, n, n, n);
		for (size_t i = 0; i < N; ++i) uu[i] -= phx[i];
		for (size_t i = 0; i < Nv; ++i) vv[i] -= phy[i];
		for (size_t i = 0; i < N; ++i) ww[i] -= phz[i];
	}
	// ns/solver.h:90-140
	void step(vec& ut, vec& ub, vec& wt, vec& wb, const vec& fx, const vec& fy, const vec& fz)
	{
		vec uadv, vadv, wadv, rx(N), ry(Nv), rz(N);
		advect3(uadv, vadv, wadv, *u, *v, *w, n, n, n, n);
		const double sm = mu * dt / (2 * rho), ss = -mu * dt / (2 * rho);
		if (!half) {
			u_old = *u; v_old = *v; w_old = *w;
			proj_boundary3(ut, ub, wt, wb, phx, phz, n, n);
			vec bu = hbound3(1, ut, ub, ss, n, n, n, n), bw = hbound3(1, wt, wb, ss, n, n, n, n);
			for (double& e : bu) e *= -1;
			for (double& e : bw) e *= -1;
			for (size_t i = 0; i < N; ++i) {
				rx[i] = u_old[i] + bu[i] + (dt / 2) * (1 / rho * fx[i] - uadv[i]);
				rz[i] = w_old[i] + bw[i] + (dt / 2) * (1 / rho * fz[i] - wadv[i]);
			}
			for (size_t i = 0; i < Nv; ++i) ry[i] = v_old[i] + (dt / 2) * (1 / rho * fy[i] - vadv[i]);
		} else {
			proj_boundary3(ut, ub, wt, wb, phx, phz, n, n);
			vec bu = hbound3(1, ut, ub, sm, n, n, n, n), bw = hbound3(1, wt, wb, sm, n, n, n, n);
			vec su = hbound3(1, ut, ub, ss, n, n, n, n), sw = hbound3(1, wt, wb, ss, n, n, n, n);
			for (size_t i = 0; i < N; ++i) { bu[i] -= su[i]; bw[i] -= sw[i]; }
			vec hu = spmv(phhm, u_old), hv = spmv(dhhm, v_old), hw = spmv(phhm, w_old);
			for (size_t i = 0; i < N; ++i) {
				rx[i] = hu[i] + bu[i] + dt * (1 / rho * fx[i] - uadv[i]);
				rz[i] = hw[i] + bw[i] + dt * (1 / rho * fz[i] - wadv[i]);
			}
			for (size_t i = 0; i < Nv; ++i) ry[i] = hv[i] + dt * (1 / rho * fy[i] - vadv[i]);
		}
		last_cg = cg(*u, phhs, rx, tol, 1000000, nullptr);
		last_cg += cg(*v, dhhs, ry, tol, 1000000, nullptr);
		last_cg += cg(*w, phhs, rz, tol, 1000000, nullptr);
		project(*u, *v, *w);
		half = !half;
	}
	// ns/solver.h:83: p = phhs * phidt
	vec pressure() const { return spmv(phhs, phi); }
};

}  // namespace

// =================================================================== C ABI
// Spreading as the reference computes it: every face gathers the markers in
// the box of half-width 2/n around it from a kd-tree (ib/grid.h:68-84 over
// ib/lagrangian_grid.h:58-67 and dt/kdtree.h:39-146: inclusive box, one
// axis per tree level, periodic images of markers within 2/n of a seam at
// +-1), and sums delta_x delta_y ds F over them in traversal order.  The
// same terms as or_spread2 in another order; or_spread2's scatter is what
// the parity tests use, this gather what the CPU baseline times.
struct KdTree2 {
	struct Node { double p[2]; long k; int left = -1, right = -1, axis = 0; };
	std::vector<Node> nodes;
	int build(std::vector<long>& ids, long lo, long hi, int axis, const double* X)
	{
		if (lo >= hi) return -1;
		const long mid = lo + (hi - lo) / 2;
		std::nth_element(ids.begin() + lo, ids.begin() + mid, ids.begin() + hi,
		                 [&](long a, long b) { return X[2 * a + axis] < X[2 * b + axis]; });
		const int me = (int) nodes.size();
		nodes.push_back(Node());
		nodes[me].k = ids[mid];
		nodes[me].p[0] = X[2 * ids[mid]];
		nodes[me].p[1] = X[2 * ids[mid] + 1];
		nodes[me].axis = axis;
		const int l = build(ids, lo, mid, axis ^ 1, X);
		const int r = build(ids, mid + 1, hi, axis ^ 1, X);
		nodes[me].left = l;
		nodes[me].right = r;
		return me;
	}
	// visit(k, sx, sy) for every marker k whose image (X + (sx, sy)) lies in
	// the box of half-width h around q
	template<class Visit>
	void query(int node, const double q[2], double h, double sx, double sy, Visit&& visit) const
	{
		while (node >= 0) {
			const Node& nd = nodes[node];
			const double c = nd.p[nd.axis] + (nd.axis ? sy : sx);
			if (q[nd.axis] + h < c) { node = nd.left; continue; }
			if (q[nd.axis] - h > c) { node = nd.right; continue; }
			query(nd.right, q, h, sx, sy, visit);
			if (std::fabs(nd.p[0] + sx - q[0]) <= h && std::fabs(nd.p[1] + sy - q[1]) <= h) visit(nd.k, sx, sy);
			node = nd.left;
		}
	}
};

extern "C" {

double or_delta(double r, unsigned n, int mode) { return delta(r, n, mode); }
double or_delta_arg(double r, unsigned n) { return delta_arg(r, n); }
double or_cos_glibc(double x) { return ::cos(x); }
double or_cos_cr(double x) { return cos_cr(x); }

void or_delta_many(const double* r, long count, unsigned n, int mode, double* out)
{
	for (long i = 0; i < count; ++i) out[i] = delta(r[i], n, mode);
}
void or_delta_arg_many(const double* r, long count, unsigned n, double* out)
{
	for (long i = 0; i < count; ++i) out[i] = delta_arg(r[i], n);
}

// ib/grid.h:68-84 (marker targets): U_k = sum_taps (w * vol) * field[idx].
void or_interp2(int comp, const double* field, unsigned nx, unsigned ny, unsigned n, unsigned W, unsigned D,
                const double* X, long m, double* U, int mode)
{
	const double vol = 1. / (n * n);
	for (long k = 0; k < m; ++k) {
		size_t idx[16]; double w[16];
		window2(comp, X[2 * k], X[2 * k + 1], n, W, D, nx, ny, mode, idx, w);
		double value = 0.0;
		for (int t = 0; t < 16; ++t) value += w[t] * vol * field[idx[t]];
		U[k] = value;
	}
}

// Spreading: the same (idx, w) pairs, summed per cell in ascending marker
// order: f[idx] += (w * ds_k) * F_k (ib/grid.h:80 with the source volume ds).
void or_spread2(int comp, const double* F, const double* ds, const double* X, long m,
                unsigned nx, unsigned ny, unsigned n, unsigned W, unsigned D, double* f, int mode)
{
	std::fill(f, f + (size_t) nx * ny, 0.0);
	for (long k = 0; k < m; ++k) {
		size_t idx[16]; double w[16];
		window2(comp, X[2 * k], X[2 * k + 1], n, W, D, nx, ny, mode, idx, w);
		for (int t = 0; t < 16; ++t) f[idx[t]] += w[t] * ds[k] * F[k];
	}
}

void or_spread2_gather(int comp, const double* F, const double* ds, const double* X, long m,
                       unsigned nx, unsigned ny, unsigned n, unsigned W, unsigned D, double* f, int mode)
{
	KdTree2 t;
	std::vector<long> ids(m);
	for (long k = 0; k < m; ++k) ids[k] = k;
	t.nodes.reserve(m);
	const int root = t.build(ids, 0, m, 0, X);
	const double h = 2.0 / n;
	for (unsigned j = 0; j < ny; ++j)
		for (unsigned i = 0; i < nx; ++i) {
			double q[2] = {(double) i / n, (double) j / n};
			if (comp == 0) q[1] += 0.5 / n; else q[0] += 0.5 / n;
			double value = 0.0;
			auto add = [&](long k, double sx, double sy) {
				const double w = delta(X[2 * k] + sx - q[0], n, mode) * delta(X[2 * k + 1] + sy - q[1], n, mode);
				value += w * ds[k] * F[k];
			};
			for (int a = -1; a <= 1; ++a)
				for (int b = -1; b <= 1; ++b) {
					const double sx = a * (double) W, sy = b * (double) D;
					// images one period left (right) only matter for faces within h of
					// the left (right) seam
					if ((a == -1 && q[0] > h) || (a == 1 && q[0] < W - h)) continue;
					if ((b == -1 && q[1] > h) || (b == 1 && q[1] < D - h)) continue;
					t.query(root, q, h, sx, sy, add);
				}
			f[i + (size_t) nx * j] = value;
		}
}

// Per-marker window of one component: base cell, 16 indices, 16 weights.
void or_window2(int comp, double cx, double cy, unsigned nx, unsigned ny, unsigned n, unsigned W, unsigned D,
                int mode, long* idx_out, double* w_out)
{
	size_t idx[16]; double w[16];
	window2(comp, cx, cy, n, W, D, nx, ny, mode, idx, w);
	for (int t = 0; t < 16; ++t) { idx_out[t] = (long) idx[t]; w_out[t] = w[t]; }
}

void or_interp3(int comp, const double* field, unsigned nx, unsigned ny, unsigned nz, unsigned n,
                unsigned W, unsigned H, unsigned D, const double* X, long m, double* U, int mode)
{
	const double vol = 1. / (n * n * n);
	for (long k = 0; k < m; ++k) {
		size_t idx[64]; double w[64];
		window3(comp, X + 3 * k, n, W, H, D, nx, ny, nz, mode, idx, w);
		double value = 0.0;
		for (int t = 0; t < 64; ++t)
			if (idx[t] != SIZE_MAX) value += w[t] * vol * field[idx[t]];
		U[k] = value;
	}
}

void or_spread3(int comp, const double* F, const double* ds, const double* X, long m,
                unsigned nx, unsigned ny, unsigned nz, unsigned n, unsigned W, unsigned H, unsigned D,
                double* f, int mode)
{
	size_t N = (size_t) nx * (comp == 1 ? ny - 1 : ny) * nz;
	std::fill(f, f + N, 0.0);
	for (long k = 0; k < m; ++k) {
		size_t idx[64]; double w[64];
		window3(comp, X + 3 * k, n, W, H, D, nx, ny, nz, mode, idx, w);
		for (int t = 0; t < 64; ++t)
			if (idx[t] != SIZE_MAX) f[idx[t]] += w[t] * ds[k] * F[k];
	}
}

// CSR export: kind 0 laplacian, 1 helmholtz_periodic (3-D), 2 helmholtz 2-D.
long or_stencil_nnz(unsigned nx, unsigned ny, unsigned nz, unsigned n, int kind, double scale)
{
	return (long) stencil_csr(nx, ny, nz, n, kind, scale).val.size();
}
void or_stencil_csr(unsigned nx, unsigned ny, unsigned nz, unsigned n, int kind, double scale,
                    long* rs, long* col, double* val)
{
	Csr A = stencil_csr(nx, ny, nz, n, kind, scale);
	for (size_t i = 0; i <= A.N; ++i) rs[i] = (long) A.rs[i];
	for (size_t k = 0; k < A.val.size(); ++k) { col[k] = (long) A.col[k]; val[k] = A.val[k]; }
}
void or_stencil_apply(unsigned nx, unsigned ny, unsigned nz, unsigned n, int kind, double scale,
                      const double* x, double* y)
{
	Csr A = stencil_csr(nx, ny, nz, n, kind, scale);
	vec xv(x, x + A.N);
	vec yv = spmv(A, xv);
	std::copy(yv.begin(), yv.end(), y);
}
// One smoothing sweep (lexicographic: smoother 0; red-black: 1) of the
// laplacian on the given grid.
void or_smooth(unsigned nx, unsigned ny, unsigned nz, unsigned n, int smoother, double* x, const double* b, int sweeps)
{
	Csr A = stencil_csr(nx, ny, nz, n, 0, 0.0);
	vec xv(x, x + A.N), bv(b, b + A.N);
	for (int s = 0; s < sweeps; ++s) {
		if (smoother == 0) smooth_lex(xv, A, bv); else smooth_rb(xv, A, bv, nx, ny, nz);
	}
	std::copy(xv.begin(), xv.end(), x);
}
unsigned or_cg(unsigned nx, unsigned ny, unsigned nz, unsigned n, int kind, double scale,
               double* x, const double* b, double tol, unsigned max_iter, double* final_norm)
{
	Csr A = stencil_csr(nx, ny, nz, n, kind, scale);
	vec xv(x, x + A.N), bv(b, b + A.N);
	unsigned it = cg(xv, A, bv, tol, max_iter, final_norm);
	std::copy(xv.begin(), xv.end(), x);
	return it;
}
void or_restrict3(unsigned nx, unsigned ny, unsigned nz, const double* fine, double* coarse)
{
	vec in(fine, fine + (size_t) nx * ny * nz), out;
	restrict3(out, in, nx, ny, nz);
	std::copy(out.begin(), out.end(), coarse);
}
void or_restrict2(unsigned nx, unsigned ny, const double* fine, double* coarse)
{
	vec in(fine, fine + (size_t) nx * ny), out;
	restrict2(out, in, nx, ny);
	std::copy(out.begin(), out.end(), coarse);
}
// out is read-modify-write: rows the reference does not visit keep their value.
void or_prolong3(unsigned nx, unsigned ny, unsigned nz, const double* coarse, double* fine)
{
	vec in(coarse, coarse + (size_t) nx * ny * nz / 8), out(fine, fine + (size_t) nx * ny * nz);
	prolong3(out, in, nx, ny, nz);
	std::copy(out.begin(), out.end(), fine);
}
void or_prolong2(unsigned nx, unsigned ny, const double* coarse, double* fine)
{
	vec in(coarse, coarse + (size_t) nx * ny / 4), out(fine, fine + (size_t) nx * ny);
	prolong2(out, in, nx, ny);
	std::copy(out.begin(), out.end(), fine);
}
// mg::solver<G>(tol, nu1, nu2).solve(rhs): rhs is overwritten by the
// solution; returns the number of V-cycles.  dim 2: (W, D) plane, H ignored.
unsigned or_mg_solve(int dim, unsigned levels, unsigned W, unsigned H, unsigned D, double tol, int tol_mode,
                     unsigned nu1, unsigned nu2, int smoother, unsigned max_cycles, double* rhs, double* final_rms)
{
	unsigned n = 1u << levels;
	size_t N = (size_t) W * n * (dim == 3 ? H * n : 1) * D * n;
	vec r(rhs, rhs + N);
	double t = tol;
	if (tol_mode == 1) {
		vec tmp(r), ones(N, 1.0);
		double c = dot(tmp, ones) / (double) N;
		for (double& e : tmp) e -= c * 1.0;
		t = tol * rms(tmp);
	}
	MG mg(dim, levels, W, H, D, t, nu1, nu2, smoother);
	unsigned cyc = mg.solve(r, max_cycles, final_rms);
	std::copy(r.begin(), r.end(), rhs);
	return cyc;
}
// One V-cycle (mg/solver.h:199-216) on the residual equation at the top level.
void or_mg_vcycle(int dim, unsigned levels, unsigned W, unsigned H, unsigned D, double tol,
                  unsigned nu1, unsigned nu2, int smoother, double* rhs_inout)
{
	unsigned n = 1u << levels;
	size_t N = (size_t) W * n * (dim == 3 ? H * n : 1) * D * n;
	MG mg(dim, levels, W, H, D, tol, nu1, nu2, smoother);
	vec r(rhs_inout, rhs_inout + N);
	mg.recurse(levels, r);
	std::copy(r.begin(), r.end(), rhs_inout);
}
void or_div2(unsigned nx, unsigned ny, unsigned n, const double* u, const double* v, double* d)
{
	G2 g{n, nx, ny, nx / n, ny / n};
	size_t N = (size_t) nx * ny;
	vec uv(u, u + N), vv(v, v + N), dv(N);
	div2(dv, uv, vv, g);
	std::copy(dv.begin(), dv.end(), d);
}
void or_grad2(unsigned nx, unsigned ny, unsigned n, const double* p, double* px, double* py)
{
	G2 g{n, nx, ny, nx / n, ny / n};
	size_t N = (size_t) nx * ny;
	vec pv(p, p + N), a(N), b(N);
	grad2(a, b, pv, g);
	std::copy(a.begin(), a.end(), px);
	std::copy(b.begin(), b.end(), py);
}
void or_advect2(unsigned nx, unsigned ny, unsigned n, const double* u, const double* v, double* ua, double* va)
{
	G2 g{n, nx, ny, nx / n, ny / n};
	size_t N = (size_t) nx * ny;
	vec uv(u, u + N), vv(v, v + N), a(N), b(N);
	advect2(a, b, uv, vv, g);
	std::copy(a.begin(), a.end(), ua);
	std::copy(b.begin(), b.end(), va);
}

// Wall time of the step loop of the last or_ib_run (excludes the initial
// projection and setup), for the CPU baseline.
static double g_last_step_seconds = 0.0;
double or_last_step_seconds(void) { return g_last_step_seconds; }

// ------------------------------------------------------- IB time stepping
// Build-defined driver (the reference has none; its intended coupling is
// the comment at test/ib_ns_test.cc:80-90).  Per dt:
//   F = -k (X - R);  f = S(X) F;  BE half step  -> u^{n+1/2}
//   X_h = X + (dt/2) J(X) u^{n+1/2}
//   F = -k (X_h - R);  f = S(X_h) F;  CN full step -> u^{n+1}
//   X = X + dt * (0.5 * (J(X_h) u^n + J(X_h) u^{n+1}))
struct IB2 {
	G2 g;
	NS2* ns;
	long m;
	vec X, R, k, ds;
	vec u, v;
	int mode = 1;
	unsigned cycles = 0, cg_iters = 0;
};

void spring(const IB2& s, const vec& X, vec& Fx, vec& Fy)
{
	for (long i = 0; i < s.m; ++i) {
		Fx[i] = -s.k[i] * (X[2 * i] - s.R[2 * i]);
		Fy[i] = -s.k[i] * (X[2 * i + 1] - s.R[2 * i + 1]);
	}
}

static int g_spread_gather = 0;   // 1: the IB step spreads by the kd-tree gather
void or_set_spread_gather(int on) { g_spread_gather = on; }

void spread_both(const IB2& s, const vec& X, const vec& Fx, const vec& Fy, vec& fx, vec& fy)
{
	auto sp = g_spread_gather ? or_spread2_gather : or_spread2;
	sp(0, Fx.data(), s.ds.data(), X.data(), s.m, s.g.nx, s.g.ny, s.g.n, s.g.W, s.g.D, fx.data(), s.mode);
	sp(1, Fy.data(), s.ds.data(), X.data(), s.m, s.g.nx, s.g.ny, s.g.n, s.g.W, s.g.D, fy.data(), s.mode);
}

void interp_both(const IB2& s, const vec& u, const vec& v, const vec& X, vec& Ux, vec& Uy)
{
	or_interp2(0, u.data(), s.g.nx, s.g.ny, s.g.n, s.g.W, s.g.D, X.data(), s.m, Ux.data(), s.mode);
	or_interp2(1, v.data(), s.g.nx, s.g.ny, s.g.n, s.g.W, s.g.D, X.data(), s.m, Uy.data(), s.mode);
}

void ib_step(IB2& s)
{
	size_t N = (size_t) s.g.nx * s.g.ny;
	const double dt = s.ns->dt;
	vec Fx(s.m), Fy(s.m), fx(N), fy(N), Ux(s.m), Uy(s.m), Xh(2 * s.m);
	spring(s, s.X, Fx, Fy);
	spread_both(s, s.X, Fx, Fy, fx, fy);
	s.ns->step(fx, fy);
	s.cycles += s.ns->last_cycles; s.cg_iters += s.ns->last_cg;
	interp_both(s, s.u, s.v, s.X, Ux, Uy);
	const double half = 0.5 * dt;
	for (long i = 0; i < s.m; ++i) {
		Xh[2 * i] = s.X[2 * i] + half * Ux[i];
		Xh[2 * i + 1] = s.X[2 * i + 1] + half * Uy[i];
	}
	spring(s, Xh, Fx, Fy);
	spread_both(s, Xh, Fx, Fy, fx, fy);
	s.ns->step(fx, fy);
	s.cycles += s.ns->last_cycles; s.cg_iters += s.ns->last_cg;
	vec U0x(s.m), U0y(s.m);
	interp_both(s, s.ns->u_old, s.ns->w_old, Xh, U0x, U0y);
	interp_both(s, s.u, s.v, Xh, Ux, Uy);
	for (long i = 0; i < s.m; ++i) {
		s.X[2 * i] = s.X[2 * i] + dt * (0.5 * (U0x[i] + Ux[i]));
		s.X[2 * i + 1] = s.X[2 * i + 1] + dt * (0.5 * (U0y[i] + Uy[i]));
	}
}

// Runs `steps` IB steps in place.  u, v: nx*ny each (projected on entry, as
// the ns::solver constructor does, ns/solver.h:71); X, R: 2m; k, ds: m.
// Returns total V-cycles; *cg_iters receives total CG iterations.
// p_out (optional): the potential phi of the last projection (ns/projection.h:69).
unsigned or_ib_run(unsigned nx, unsigned ny, unsigned n, double mu, double rho, double dt, double tol, int tol_mode,
                   int smoother, int delta_mode, double* u, double* v, long m, double* X, const double* R,
                   const double* k, const double* ds, int steps, unsigned* cg_iters, double* p_out)
{
	IB2 s;
	s.g = G2{n, nx, ny, nx / n, ny / n};
	size_t N = (size_t) nx * ny;
	s.u.assign(u, u + N); s.v.assign(v, v + N);
	s.m = m;
	s.X.assign(X, X + 2 * m); s.R.assign(R, R + 2 * m);
	s.k.assign(k, k + m); s.ds.assign(ds, ds + m);
	s.mode = delta_mode;
	NS2 ns(s.g, mu, rho, dt, tol, tol_mode, smoother, &s.u, &s.v);
	s.ns = &ns;
	auto t0 = std::chrono::steady_clock::now();
	for (int i = 0; i < steps; ++i) ib_step(s);
	g_last_step_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
	std::copy(s.u.begin(), s.u.end(), u);
	std::copy(s.v.begin(), s.v.end(), v);
	std::copy(s.X.begin(), s.X.end(), X);
	if (p_out) std::copy(ns.p_last.begin(), ns.p_last.end(), p_out);
	if (cg_iters) *cg_iters = s.cg_iters;
	return s.cycles;
}

// ns::solver steps without markers: `calls` step() calls with fixed forces.
// Returns the V-cycles of the last projection.
unsigned or_ns_run(unsigned nx, unsigned ny, unsigned n, double mu, double rho, double dt, double tol, int tol_mode,
                   int smoother, double* u, double* v, const double* fx, const double* fy, int calls, double* p_out)
{
	G2 g{n, nx, ny, nx / n, ny / n};
	size_t N = (size_t) nx * ny;
	vec uv(u, u + N), vv(v, v + N), fxv(fx, fx + N), fyv(fy, fy + N);
	NS2 ns(g, mu, rho, dt, tol, tol_mode, smoother, &uv, &vv);
	for (int i = 0; i < calls; ++i) ns.step(fxv, fyv);
	std::copy(uv.begin(), uv.end(), u);
	std::copy(vv.begin(), vv.end(), v);
	if (p_out) std::copy(ns.p_last.begin(), ns.p_last.end(), p_out);
	return ns.last_cycles;
}

// --- 3-D wall-bounded mode (n^3 cells; v, py have n*(n-1)*n entries) ---
void or_div3(unsigned nx, unsigned ny, unsigned nz, unsigned n, const double* u, const double* v, const double* w,
             const double* top, const double* bottom, double* d)
{
	size_t N = (size_t) nx * ny * nz, Nv = (size_t) nx * (ny - 1) * nz, Nxz = (size_t) nx * nz;
	vec dv, uv(u, u + N), vv(v, v + Nv), wv(w, w + N), tv(top, top + Nxz), bv(bottom, bottom + Nxz);
	div3(dv, uv, vv, wv, tv, bv, nx, ny, nz, n);
	std::copy(dv.begin(), dv.end(), d);
}

void or_grad3(unsigned n, const double* p, double* px, double* py, double* pz)
{
	size_t N = (size_t) n * n * n;
	vec pv(p, p + N), a, b, c;
	grad3(a, b, c, pv, n, n, n, n);
	std::copy(a.begin(), a.end(), px);
	std::copy(b.begin(), b.end(), py);
	std::copy(c.begin(), c.end(), pz);
}

void or_advect3(unsigned n, const double* u, const double* v, const double* w, double* ua, double* va, double* wa)
{
	size_t N = (size_t) n * n * n, Nv = (size_t) n * (n - 1) * n;
	vec uv(u, u + N), vv(v, v + Nv), wv(w, w + N), a, b, c;
	advect3(a, b, c, uv, vv, wv, n, n, n, n);
	std::copy(a.begin(), a.end(), ua);
	std::copy(b.begin(), b.end(), va);
	std::copy(c.begin(), c.end(), wa);
}

// kind 1 periodic-cell (n^3 rows), 3 Dirichlet (n*(n-1)*n rows)
void or_hbound3(int kind, unsigned n, double scale, const double* top, const double* bottom, double* b)
{
	size_t Nxz = (size_t) n * n;
	vec tv(top, top + Nxz), bv(bottom, bottom + Nxz);
	vec r = hbound3(kind, tv, bv, scale, n, n, n, n);
	std::copy(r.begin(), r.end(), b);
}

void or_proj_boundary3(unsigned n, const double* phx, const double* phz, double* ut, double* ub, double* wt,
                       double* wb)
{
	size_t N = (size_t) n * n * n, Nxz = (size_t) n * n;
	vec a(phx, phx + N), c(phz, phz + N), t1(ut, ut + Nxz), b1(ub, ub + Nxz), t2(wt, wt + Nxz), b2(wb, wb + Nxz);
	proj_boundary3(t1, b1, t2, b2, a, c, n, n);
	std::copy(t1.begin(), t1.end(), ut);
	std::copy(b1.begin(), b1.end(), ub);
	std::copy(t2.begin(), t2.end(), wt);
	std::copy(b2.begin(), b2.end(), wb);
}

// ns::solver<geometry<L,1,1,1>>: construct (projects u, v, w), then `steps`
// step() calls with fixed forces; the wall vectors are updated in place as
// the reference's step does.  p_out (n^3) gets get_p(); stats[0] the last
// projection's V-cycles, stats[1] the last step's CG iterations.
void or_ns3_run(unsigned L, double mu, double rho, double dt, double tol, int smoother, double* u, double* v,
                double* w, double* ut, double* ub, double* wt, double* wb, const double* fx, const double* fy,
                const double* fz, int steps, double* p_out, unsigned* stats)
{
	unsigned n = 1u << L;
	size_t N = (size_t) n * n * n, Nv = (size_t) n * (n - 1) * n, Nxz = (size_t) n * n;
	vec uv(u, u + N), vv(v, v + Nv), wv(w, w + N), fxv(fx, fx + N), fyv(fy, fy + Nv), fzv(fz, fz + N);
	vec t1(ut, ut + Nxz), b1(ub, ub + Nxz), t2(wt, wt + Nxz), b2(wb, wb + Nxz);
	NS3 ns(L, mu, rho, dt, tol, smoother, &uv, &vv, &wv);
	for (int s = 0; s < steps; ++s) ns.step(t1, b1, t2, b2, fxv, fyv, fzv);
	std::copy(uv.begin(), uv.end(), u);
	std::copy(vv.begin(), vv.end(), v);
	std::copy(wv.begin(), wv.end(), w);
	std::copy(t1.begin(), t1.end(), ut);
	std::copy(b1.begin(), b1.end(), ub);
	std::copy(t2.begin(), t2.end(), wt);
	std::copy(b2.begin(), b2.end(), wb);
	if (p_out) { vec p = ns.pressure(); std::copy(p.begin(), p.end(), p_out); }
	if (stats) { stats[0] = ns.last_cycles; stats[1] = ns.last_cg; }
}

}  // extern "C"

// ===================================================================
// Discrete adjoint of the IB time step (build-only; no reference
// counterpart).  Objective at the final time T:
//   J = 1/2 sum (u^2 + v^2) h^2 + sum_m |c_m(T) - c_m(0)|^2,
// c_m the centroid of membrane m (consecutive blocks of `msize` markers).
// Solves are differentiated as exact linear solves (implicit
// differentiation): the Helmholtz operators and the projection are
// symmetric on the periodic grid, so their adjoints are the same solves.
namespace {

// d delta / dr for the window's arguments (|n r| <= 2), with the forward
// argument theta: -(0.25 n)(pi n / 2) sin(theta).
double ddelta(double r, unsigned n)
{
	if (std::abs((int) (n * r)) > 2.) return 0;
	double a = delta_arg(r, n);
	return -(0.25 * n) * (M_PI * n / 2) * ::sin(a);
}

// window weights and their derivatives with respect to the marker position
void window2_d(int comp, double cx, double cy, const G2& g, int mode, size_t idx[16], double w[16],
               double wx[16], double wy[16])
{
	double q0 = cx, q1 = cy;
	const unsigned n = g.n;
	if (comp == 0) q1 -= 0.5 / n; else q0 -= 0.5 / n;
	q0 = std::fmod(std::fmod(q0, (double) g.W) + g.W, (double) g.W);
	q1 = std::fmod(std::fmod(q1, (double) g.D) + g.D, (double) g.D);
	const double fx = std::floor(q0 * n), fy = std::floor(q1 * n);
	double dx[4], dy[4], ddx[4], ddy[4];
	for (int d = 0; d < 4; ++d) {
		double x = (fx + d - 1.0) / n, y = (fy + d - 1.0) / n;
		if (comp == 0) y += 0.5 / n; else x += 0.5 / n;
		dx[d] = delta(x - cx, n, mode); dy[d] = delta(y - cy, n, mode);
		ddx[d] = ddelta(x - cx, n); ddy[d] = ddelta(y - cy, n);
	}
	for (int i = 0; i < 16; ++i) {
		const int a = i % 4, b = i / 4;
		idx[i] = ((size_t) fx + a + g.nx - 1) % g.nx + (size_t) g.nx * (((size_t) fy + b + g.ny - 1) % g.ny);
		w[i] = dx[a] * dy[b];
		wx[i] = -ddx[a] * dy[b];
		wy[i] = -dx[a] * ddy[b];
	}
}

// out += scale * J(X)^T Ub   (J = interpolation with volume h^2)
void interp_T(int comp, const vec& Ub, const vec& X, double scale, vec& out, const G2& g, int mode)
{
	const double vol = 1. / (g.n * g.n);
	size_t idx[16]; double w[16], wx[16], wy[16];
	for (size_t k = 0; k < Ub.size(); ++k) {
		window2_d(comp, X[2 * k], X[2 * k + 1], g, mode, idx, w, wx, wy);
		for (int t = 0; t < 16; ++t) out[idx[t]] += scale * (w[t] * vol * Ub[k]);
	}
}

// Xb[k] += coef_k * d(J f)_k / dX_k  (both position components)
void interp_dX(int comp, const vec& f, const vec& X, const vec& coef, vec& Xb, const G2& g, int mode)
{
	const double vol = 1. / (g.n * g.n);
	size_t idx[16]; double w[16], wx[16], wy[16];
	for (size_t k = 0; k < coef.size(); ++k) {
		window2_d(comp, X[2 * k], X[2 * k + 1], g, mode, idx, w, wx, wy);
		double gx = 0.0, gy = 0.0;
		for (int t = 0; t < 16; ++t) { gx += wx[t] * vol * f[idx[t]]; gy += wy[t] * vol * f[idx[t]]; }
		Xb[2 * k] += coef[k] * gx;
		Xb[2 * k + 1] += coef[k] * gy;
	}
}

// (ub, vb) += A'(u, v)^T (ab, bb), A the advection of advect2
void advect2_T(const vec& ab, const vec& bb, const vec& u, const vec& w, vec& ub, vec& wb, const G2& g)
{
	const size_t N = u.size();
	vec uub(N, 0.0), uwb(N, 0.0), wwb(N, 0.0);
	for (size_t i = 0; i < N; ++i) {
		unsigned x = i % g.nx, z = i / g.nx;
		size_t ixp = (x + 1) % g.nx + (size_t) g.nx * z, izp = x + (size_t) g.nx * ((z + 1) % g.ny);
		// ua[i] = n ((uu[ixp] - uu[i]) + (uw[izp] - uw[i])); wa[i] = n ((uw[ixp] - uw[i]) + (ww[izp] - ww[i]))
		uub[ixp] += g.n * ab[i]; uub[i] -= g.n * ab[i];
		uwb[izp] += g.n * ab[i]; uwb[i] -= g.n * ab[i];
		uwb[ixp] += g.n * bb[i]; uwb[i] -= g.n * bb[i];
		wwb[izp] += g.n * bb[i]; wwb[i] -= g.n * bb[i];
	}
	for (size_t i = 0; i < N; ++i) {
		unsigned x = i % g.nx, z = i / g.nx;
		size_t ixm = (x + g.nx - 1) % g.nx + (size_t) g.nx * z, izm = x + (size_t) g.nx * ((z + g.ny - 1) % g.ny);
		double q = 0.5 * (u[i] + u[ixm]);          // uu = q^2
		double qb = 2 * q * uub[i];
		ub[i] += 0.5 * qb; ub[ixm] += 0.5 * qb;
		q = 0.5 * (w[i] + w[izm]);                 // ww = q^2
		qb = 2 * q * wwb[i];
		wb[i] += 0.5 * qb; wb[izm] += 0.5 * qb;
		q = 0.5 * (u[izm] + u[i]);                 // uw = q r
		double r = 0.5 * (w[ixm] + w[i]);
		ub[izm] += 0.5 * r * uwb[i]; ub[i] += 0.5 * r * uwb[i];
		wb[ixm] += 0.5 * q * uwb[i]; wb[i] += 0.5 * q * uwb[i];
	}
}

struct Tape { vec u, v, X, uh, vh, Xh; };

double objective(const IB2& s, const vec& X0c)
{
	const double h2 = 1.0 / ((double) s.g.n * s.g.n);
	double J = 0.0;
	for (size_t i = 0; i < s.u.size(); ++i) J += 0.5 * (s.u[i] * s.u[i] + s.v[i] * s.v[i]) * h2;
	size_t nm = X0c.size() / 2;
	long msize = s.m / (long) nm;
	for (size_t mm = 0; mm < nm; ++mm) {
		double cx = 0, cy = 0;
		for (long j = 0; j < msize; ++j) { cx += s.X[2 * (mm * msize + j)]; cy += s.X[2 * (mm * msize + j) + 1]; }
		cx /= msize; cy /= msize;
		J += (cx - X0c[2 * mm]) * (cx - X0c[2 * mm]) + (cy - X0c[2 * mm + 1]) * (cy - X0c[2 * mm + 1]);
	}
	return J;
}

vec centroids(const vec& X, long m, long msize)
{
	long nm = m / msize;
	vec c(2 * nm, 0.0);
	for (long mm = 0; mm < nm; ++mm) {
		for (long j = 0; j < msize; ++j) { c[2 * mm] += X[2 * (mm * msize + j)]; c[2 * mm + 1] += X[2 * (mm * msize + j) + 1]; }
		c[2 * mm] /= msize; c[2 * mm + 1] /= msize;
	}
	return c;
}

// P applied to an adjoint velocity (the projection is symmetric)
void project_adj(NS2& ns, vec& ub, vec& vb) { ns.project(ub, vb); }

// H_s^{-1} b (H_s symmetric), initial guess b
vec helm_inv(NS2& ns, const vec& b)
{
	vec x(b);
	double t = ns.tol_mode == 1 ? ns.tol * sqrt(dot(b, b)) : ns.tol;
	cg(x, ns.phhs, b, t, 100000, nullptr);
	return x;
}

}  // namespace

extern "C" {

// Runs `steps` IB steps and returns J; with grad_k / grad_R non-null also
// the adjoint gradients dJ/dk (m) and dJ/dR (2m).
double or_ib_grad(unsigned nx, unsigned ny, unsigned n, double mu, double rho, double dt, double tol, int tol_mode,
                  int smoother, int delta_mode, const double* u0, const double* v0, long m, const double* X0,
                  const double* R, const double* k, const double* ds, long msize, int steps,
                  double* grad_k, double* grad_R)
{
	IB2 s;
	s.g = G2{n, nx, ny, nx / n, ny / n};
	const G2& g = s.g;
	size_t N = (size_t) nx * ny;
	s.u.assign(u0, u0 + N); s.v.assign(v0, v0 + N);
	s.m = m;
	s.X.assign(X0, X0 + 2 * m); s.R.assign(R, R + 2 * m);
	s.k.assign(k, k + m); s.ds.assign(ds, ds + m);
	s.mode = delta_mode;
	NS2 ns(g, mu, rho, dt, tol, tol_mode, smoother, &s.u, &s.v);
	s.ns = &ns;
	vec c0 = centroids(s.X, m, msize);
	std::vector<Tape> tape(steps);
	const double half = 0.5 * dt;
	for (int st = 0; st < steps; ++st) {
		Tape& tp = tape[st];
		tp.u = s.u; tp.v = s.v; tp.X = s.X;
		vec Fx(m), Fy(m), fx(N), fy(N), Ux(m), Uy(m), Xh(2 * m);
		spring(s, s.X, Fx, Fy);
		spread_both(s, s.X, Fx, Fy, fx, fy);
		ns.step(fx, fy);
		tp.uh = s.u; tp.vh = s.v;
		interp_both(s, s.u, s.v, s.X, Ux, Uy);
		for (long i = 0; i < m; ++i) { Xh[2 * i] = s.X[2 * i] + half * Ux[i]; Xh[2 * i + 1] = s.X[2 * i + 1] + half * Uy[i]; }
		tp.Xh = Xh;
		spring(s, Xh, Fx, Fy);
		spread_both(s, Xh, Fx, Fy, fx, fy);
		ns.step(fx, fy);
		vec U0x(m), U0y(m);
		interp_both(s, ns.u_old, ns.w_old, Xh, U0x, U0y);
		interp_both(s, s.u, s.v, Xh, Ux, Uy);
		for (long i = 0; i < m; ++i) {
			s.X[2 * i] = s.X[2 * i] + dt * (0.5 * (U0x[i] + Ux[i]));
			s.X[2 * i + 1] = s.X[2 * i + 1] + dt * (0.5 * (U0y[i] + Uy[i]));
		}
	}
	double J = objective(s, c0);
	if (!grad_k && !grad_R) return J;

	// ---- reverse sweep
	const double h2 = 1.0 / ((double) n * n), vol = h2, sm = mu * dt / (2 * rho);
	Csr Hm = stencil_csr(nx, 1, ny, n, 2, sm);
	vec ub(N), vb(N), Xb(2 * m, 0.0), kb(m, 0.0), Rb(2 * m, 0.0);
	for (size_t i = 0; i < N; ++i) { ub[i] = s.u[i] * h2; vb[i] = s.v[i] * h2; }
	vec c1 = centroids(s.X, m, msize);
	for (long i = 0; i < m; ++i) {
		long mm = i / msize;
		Xb[2 * i] = 2 * (c1[2 * mm] - c0[2 * mm]) / msize;
		Xb[2 * i + 1] = 2 * (c1[2 * mm + 1] - c0[2 * mm + 1]) / msize;
	}
	const vec* unext_u = &s.u;   // u^{n+1}
	const vec* unext_v = &s.v;
	for (int st = steps - 1; st >= 0; --st) {
		const Tape& tp = tape[st];
		const vec& un = tp.u; const vec& vn = tp.v;
		// step 7: X' = X + dt * 0.5 (J(Xh) u^n + J(Xh) u^{n+1})
		vec Xn_b = Xb;
		vec Ubx(m), Uby(m);
		for (long i = 0; i < m; ++i) { Ubx[i] = 0.5 * dt * Xb[2 * i]; Uby[i] = 0.5 * dt * Xb[2 * i + 1]; }
		vec un_b(N, 0.0), vn_b(N, 0.0);
		interp_T(0, Ubx, tp.Xh, 1.0, ub, g, delta_mode);
		interp_T(1, Uby, tp.Xh, 1.0, vb, g, delta_mode);
		interp_T(0, Ubx, tp.Xh, 1.0, un_b, g, delta_mode);
		interp_T(1, Uby, tp.Xh, 1.0, vn_b, g, delta_mode);
		vec Xhb(2 * m, 0.0);
		interp_dX(0, un, tp.Xh, Ubx, Xhb, g, delta_mode);
		interp_dX(0, *unext_u, tp.Xh, Ubx, Xhb, g, delta_mode);
		interp_dX(1, vn, tp.Xh, Uby, Xhb, g, delta_mode);
		interp_dX(1, *unext_v, tp.Xh, Uby, Xhb, g, delta_mode);
		// step 6: CN full step
		vec us = ub, vs = vb;
		project_adj(ns, us, vs);
		vec rbu = helm_inv(ns, us), rbv = helm_inv(ns, vs);
		vec t1 = spmv(Hm, rbu), t2 = spmv(Hm, rbv);
		for (size_t i = 0; i < N; ++i) { un_b[i] += t1[i]; vn_b[i] += t2[i]; }
		vec f1bu(N), f1bv(N), aub(N), avb(N);
		for (size_t i = 0; i < N; ++i) {
			f1bu[i] = dt * (1 / rho) * rbu[i]; f1bv[i] = dt * (1 / rho) * rbv[i];
			aub[i] = -dt * rbu[i]; avb[i] = -dt * rbv[i];
		}
		vec uh_b(N, 0.0), vh_b(N, 0.0);
		advect2_T(aub, avb, tp.uh, tp.vh, uh_b, vh_b, g);
		// step 5: f1 = S(Xh) F1, F1 = -k (Xh - R)
		{
			vec Fx(m), Fy(m), Jx(m), Jy(m);
			spring(s, tp.Xh, Fx, Fy);
			interp_both(s, f1bu, f1bv, tp.Xh, Jx, Jy);
			vec cx(m), cy(m);
			for (long i = 0; i < m; ++i) {
				double Fbx = s.ds[i] / vol * Jx[i], Fby = s.ds[i] / vol * Jy[i];
				kb[i] += -(tp.Xh[2 * i] - s.R[2 * i]) * Fbx - (tp.Xh[2 * i + 1] - s.R[2 * i + 1]) * Fby;
				Rb[2 * i] += s.k[i] * Fbx; Rb[2 * i + 1] += s.k[i] * Fby;
				Xhb[2 * i] += -s.k[i] * Fbx; Xhb[2 * i + 1] += -s.k[i] * Fby;
				cx[i] = s.ds[i] * Fx[i] / vol; cy[i] = s.ds[i] * Fy[i] / vol;
			}
			interp_dX(0, f1bu, tp.Xh, cx, Xhb, g, delta_mode);
			interp_dX(1, f1bv, tp.Xh, cy, Xhb, g, delta_mode);
		}
		// step 4: Xh = X + half J(X) u_h
		for (long i = 0; i < 2 * m; ++i) Xn_b[i] += Xhb[i];
		{
			vec Uhx(m), Uhy(m);
			for (long i = 0; i < m; ++i) { Uhx[i] = half * Xhb[2 * i]; Uhy[i] = half * Xhb[2 * i + 1]; }
			interp_T(0, Uhx, tp.X, 1.0, uh_b, g, delta_mode);
			interp_T(1, Uhy, tp.X, 1.0, vh_b, g, delta_mode);
			interp_dX(0, tp.uh, tp.X, Uhx, Xn_b, g, delta_mode);
			interp_dX(1, tp.vh, tp.X, Uhy, Xn_b, g, delta_mode);
		}
		// step 3: BE half step
		project_adj(ns, uh_b, vh_b);
		rbu = helm_inv(ns, uh_b); rbv = helm_inv(ns, vh_b);
		vec f0bu(N), f0bv(N);
		for (size_t i = 0; i < N; ++i) {
			un_b[i] += rbu[i]; vn_b[i] += rbv[i];
			f0bu[i] = (dt / 2) * (1 / rho) * rbu[i]; f0bv[i] = (dt / 2) * (1 / rho) * rbv[i];
			aub[i] = -(dt / 2) * rbu[i]; avb[i] = -(dt / 2) * rbv[i];
		}
		advect2_T(aub, avb, un, vn, un_b, vn_b, g);
		// steps 1-2: f0 = S(X) F0, F0 = -k (X - R)
		{
			vec Fx(m), Fy(m), Jx(m), Jy(m);
			spring(s, tp.X, Fx, Fy);
			interp_both(s, f0bu, f0bv, tp.X, Jx, Jy);
			vec cx(m), cy(m);
			for (long i = 0; i < m; ++i) {
				double Fbx = s.ds[i] / vol * Jx[i], Fby = s.ds[i] / vol * Jy[i];
				kb[i] += -(tp.X[2 * i] - s.R[2 * i]) * Fbx - (tp.X[2 * i + 1] - s.R[2 * i + 1]) * Fby;
				Rb[2 * i] += s.k[i] * Fbx; Rb[2 * i + 1] += s.k[i] * Fby;
				Xn_b[2 * i] += -s.k[i] * Fbx; Xn_b[2 * i + 1] += -s.k[i] * Fby;
				cx[i] = s.ds[i] * Fx[i] / vol; cy[i] = s.ds[i] * Fy[i] / vol;
			}
			interp_dX(0, f0bu, tp.X, cx, Xn_b, g, delta_mode);
			interp_dX(1, f0bv, tp.X, cy, Xn_b, g, delta_mode);
		}
		ub = un_b; vb = vn_b; Xb = Xn_b;
		unext_u = &tp.u; unext_v = &tp.v;
	}
	if (grad_k) std::copy(kb.begin(), kb.end(), grad_k);
	if (grad_R) std::copy(Rb.begin(), Rb.end(), grad_R);
	return J;
}

}  // extern "C"
