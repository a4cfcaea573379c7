// Programs in the style of the reference's test/ directory, written against
// the C++ drop-in headers (include/afm/afm_dropin.hpp) instead of the
// reference headers.  Each prints `key value` lines that tests/test_gpu_dropin.py
// checks.
#include <cmath>
#include <cstdio>
#include <vector>

#include "afm/afm_dropin.hpp"

typedef dt::plane<5> geometry;        // 32 x 32
constexpr unsigned N = geometry::size;

// test/ib_test.cc:33-44: a single marker with volume h^2
class point : public ib::lagrangian_grid<geometry, 1u> {
public:
	point(double x, double y) : ib::lagrangian_grid<geometry, 1u>({dt::point(x, y, 0u)}) {}
	double volume(unsigned) const { return 1.0 / (geometry::n * geometry::n); }
};

// test/ib_test.cc:46-73: markers on a circle, volume 2 pi r / np
template<unsigned np>
class circle : public ib::lagrangian_grid<geometry, np> {
	double radius;
	static std::vector<dt::point> pts(double x, double y, double r)
	{
		std::vector<dt::point> p;
		for (unsigned i = 0; i < np; ++i) {
			double a = 2 * M_PI * i / np;
			p.push_back(dt::point(x + r * std::cos(a), y + r * std::sin(a), i));
		}
		return p;
	}
public:
	circle(double x, double y, double r) : ib::lagrangian_grid<geometry, np>(pts(x, y, r)), radius(r) {}
	double volume(unsigned) const { return radius * 2 * M_PI / np; }
};

// the reference's 3-D geometry (16^3, walls in y) and a cube of width 8 at
// level 1 (test/mg_test2.cc) and the single coarse cell (test/restr_test.cc)
typedef dt::geometry<4u, 1u, 1u, 1u> g3;
typedef dt::geometry<1u, 8u, 8u, 8u> g8;
typedef dt::geometry<1u, 1u, 1u, 1u> g1;

// test/ib_test.cc:33-44 in 3-D
class point3 : public ib::lagrangian_grid<g3, 1u> {
public:
	point3(double x, double y, double z) : ib::lagrangian_grid<g3, 1u>({dt::point(x, y, z, 0u)}) {}
	double volume(unsigned) const { return 1.0 / (g3::n * g3::n * g3::n); }
};

// test/ib_test.cc:46-73 in 3-D: a ring in the plane z = z0
class ring3 : public ib::lagrangian_grid<g3, 96u> {
	double radius;
	static std::vector<dt::point> pts(double x, double y, double z, double r)
	{
		std::vector<dt::point> p;
		for (unsigned i = 0; i < 96; ++i) {
			double a = 2 * pi * i / 96;
			p.push_back(dt::point(x + r * std::cos(a), y + r * std::sin(a), z, i));
		}
		return p;
	}
public:
	ring3(double x, double y, double z, double r) : ib::lagrangian_grid<g3, 96u>(pts(x, y, z, r)), radius(r) {}
	double volume(unsigned) const { return radius * 2 * pi / 96; }
};

// test/mg_test2.cc:28-35 and test/restr_test.cc:28-36: the protected
// operators reached through `using`
template<class T>
class open_solver : public mg::solver<T> {
public:
	using mg::solver<T>::L;
	using mg::solver<T>::interpolation;
	open_solver(double tol, unsigned nu1, unsigned nu2) : mg::solver<T>(tol, nu1, nu2) {}
};
template<class T>
class restriction_solver : mg::solver<T> {
public:
	using mg::solver<T>::restriction;
	restriction_solver(double tol, unsigned nu1, unsigned nu2) : mg::solver<T>(tol, nu1, nu2) {}
};

// a matrix a program fills itself: the 1-D Dirichlet second difference
// (2, -1) on 200 unknowns, symmetric positive definite
class tridiagonal : public dt::matrix<200> {
public:
	tridiagonal()
	{
		row_start.assign(1, 0);
		for (std::size_t r = 0; r < 200; ++r) {
			if (r > 0) { column.push_back(r - 1); values.push_back(-1.0); }
			column.push_back(r); values.push_back(2.0);
			if (r + 1 < 200) { column.push_back(r + 1); values.push_back(-1.0); }
			row_start.push_back(values.size());
		}
	}
};

// 7-point stencil on an n-per-unit grid of nx x ny x nz cells, periodic in
// x and z, walls in y folded into the diagonal (written out independently)
template<unsigned NX, unsigned NY, unsigned NZ>
static double stencil7(const dt::vector<NX * NY * NZ>& b, unsigned i, double n)
{
	const unsigned x = i % NX, y = (i / NX) % NY, z = i / (NX * NY);
	auto at = [&](unsigned a, unsigned bb, unsigned cc) { return b[a + NX * (bb + NY * cc)]; };
	double s = at((x + 1) % NX, y, z) + at((x + NX - 1) % NX, y, z) + at(x, y, (z + 1) % NZ) + at(x, y, (z + NZ - 1) % NZ);
	int cnt = 4;
	if (y > 0) { s += at(x, y - 1, z); ++cnt; }
	if (y < NY - 1) { s += at(x, y + 1, z); ++cnt; }
	return n * n * (s - cnt * b[i]);
}

int main()
{
	// ib_test: spread 1 from one marker, binarise, interpolate back.  The
	// result is the delta's partition of unity over the 4x4 window: 1.
	{
		ib::mac_grid<geometry, ib::mac::y> egrid;
		point lgrid(0.5, 0.5);
		lgrid[0] = 1.0;
		egrid.interpolate(lgrid);
		unsigned nz = 0;
		for (auto& e : egrid) if (e != 0.0) { e = 1; ++nz; }
		lgrid[0] = 0.0;
		lgrid.interpolate(egrid);
		std::printf("ib_nonzero %u\nib_value %.17g\n", nz, lgrid[0]);
	}
	// spreading conserves force: sum_grid f h^2 = sum_k F_k ds_k
	{
		circle<64> c(0.5, 0.5, 0.25);
		for (unsigned k = 0; k < 64; ++k) c[k] = std::cos(3.0 * k);
		ib::mac_grid<geometry, ib::mac::x> f;
		f.interpolate(c);
		double s = 0.0, t = 0.0;
		for (unsigned i = 0; i < N; ++i) s += f[i];
		for (unsigned k = 0; k < 64; ++k) t += c[k] * c.volume(k);
		std::printf("spread_momentum_error %.3e\n", std::fabs(s / N - t));
	}
	// mg_test: a Fourier-mode right-hand side has the exact discrete solution
	// f / lambda(k) on the cell-centred periodic grid
	{
		const unsigned n = geometry::n;
		dt::vector<N> b, exact;
		const double lam = -4.0 * n * n * (std::pow(std::sin(M_PI * 2 / n), 2) + std::pow(std::sin(M_PI * 3 / n), 2));
		for (unsigned i = 0; i < N; ++i) {
			double x = (i % n + 0.5) / n, y = (i / n + 0.5) / n;
			b[i] = std::cos(2 * M_PI * (2 * x + 3 * y));
			exact[i] = b[i] / lam;
		}
		mg::solver<geometry> s(1e-10, 2, 2);
		unsigned cycles = s.solve(b);
		double mean = 0.0, err = 0.0, mx = 0.0;
		for (unsigned i = 0; i < N; ++i) mean += b[i];
		mean /= N;
		for (unsigned i = 0; i < N; ++i) {
			err = std::fmax(err, std::fabs(b[i] - mean - exact[i]));
			mx = std::fmax(mx, std::fabs(exact[i]));
		}
		std::printf("mg_cycles %u\nmg_rel_error %.3e\n", cycles, err / mx);
	}
	// ns_test: zero data and zero forces stay zero through a BE + CN pair
	{
		dt::vector<N> u0, v0, fx, fy;
		ns::solver<geometry> s(1.0, 1.0, 5e-5, 1e-12, u0, v0);
		s.step(fx, fy);
		s.step(fx, fy);
		double m = 0.0;
		for (unsigned i = 0; i < N; ++i) m = std::fmax(m, std::fabs(u0[i]) + std::fabs(v0[i]));
		std::printf("ns_zero_max %.3e\n", m);
	}
	// ib_ns_test: a uniform body force accelerates a uniform flow at 1/rho
	// (the projection removes nothing from a constant field)
	{
		dt::vector<N> u0(0.5), v0, fx(1.0), fy;
		const double dt_ = 1e-3;
		ns::solver<geometry> s(1.0, 1.0, dt_, 1e-13, u0, v0);
		s.step(fx, fy);
		s.step(fx, fy);
		double e = 0.0;
		for (unsigned i = 0; i < N; ++i) e = std::fmax(e, std::fabs(u0[i] - (0.5 + dt_)) + std::fabs(v0[i]));
		std::printf("ns_uniform_error %.3e\n", e);
	}
	// mg_test3: the reference's 3-D geometry (16^3, walls in y), a U(0, 1)
	// right-hand side; the residual is checked here with a 7-point stencil
	// written out independently (wall neighbours fold into the diagonal)
	{
		constexpr unsigned n = g3::n, M = g3::size;
		dt::vector<M> b, rhs;
		unsigned long long s = 12345;
		for (unsigned i = 0; i < M; ++i) {
			s = s * 6364136223846793005ull + 1442695040888963407ull;
			b[i] = rhs[i] = (double) (s >> 11) / 9007199254740992.0;
		}
		mg::solver<g3> solver(1e-12, 2, 2);
		unsigned cycles = solver.solve(b);
		double mean = 0.0, r2 = 0.0;
		for (unsigned i = 0; i < M; ++i) mean += rhs[i];
		mean /= M;
		for (unsigned i = 0; i < M; ++i) {
			unsigned x = i % n, y = (i / n) % n, z = i / (n * n);
			double lx = 0.0;
			int cnt = 0;
			auto at = [&](unsigned a, unsigned bb, unsigned cc) { return b[a + n * (bb + n * cc)]; };
			lx += at((x + 1) % n, y, z) + at((x + n - 1) % n, y, z) + at(x, y, (z + 1) % n) + at(x, y, (z + n - 1) % n);
			cnt += 4;
			if (y > 0) { lx += at(x, y - 1, z); ++cnt; }
			if (y < n - 1) { lx += at(x, y + 1, z); ++cnt; }
			lx = (double) n * n * (lx - cnt * b[i]);
			const double r = rhs[i] - mean - lx;
			r2 += r * r;
		}
		std::printf("mg3_cycles %u\nmg3_rms %.3e\n", cycles, std::sqrt(r2 / M));
	}
	// ns_test: the reference's 3-D solver with the initial profile its
	// comment suggests, u0 = (H - y) y, no-slip walls, no force: the flow
	// stays a function of y alone (v = w = 0) and decays by diffusion
	{
		constexpr unsigned n = g3::n, nx = g3::nx, ny = g3::ny, nz = g3::nz;
		dt::vector<nx * ny * nz> u0, w0, fx, fz;
		dt::vector<nx * (ny - 1) * nz> v0, fy;
		dt::vector<nx * nz> ut, ub, wt, wb;
		for (unsigned i = 0; i < nx * ny * nz; ++i) {
			double y = (double) ((i % (nx * ny)) / nx) / n;
			u0[i] = (1.0 - y) * y;
		}
		const unsigned mid = nx * (ny / 2);
		const double before = u0[mid];
		ns::solver<g3> s(1.0, 1.0, 5e-5, 1e-12, u0, v0, w0);
		s.step(ut, ub, wt, wb, fx, fy, fz);
		s.step(ut, ub, wt, wb, fx, fy, fz);
		double vw = 0.0, spread = 0.0;
		for (unsigned i = 0; i < nx * (ny - 1) * nz; ++i) vw = std::fmax(vw, std::fabs(v0[i]));
		for (unsigned i = 0; i < nx * ny * nz; ++i) {
			vw = std::fmax(vw, std::fabs(w0[i]));
			const unsigned y = (i % (nx * ny)) / nx;
			spread = std::fmax(spread, std::fabs(u0[i] - u0[nx * y]));
		}
		std::printf("ns3_vw_max %.3e\nns3_xz_spread %.3e\nns3_u_decay %.6e\n", vw, spread, before - u0[mid]);
	}
	// ib_test.cc in 3-D: spread 1 from a marker at the centre to the v-faces,
	// binarise, interpolate back: the 4 x 4 x 4 window's partition of unity.
	// The row at r = 2h in y has a zero delta weight: 4 * 3 * 4 taps.
	{
		ib::mac_grid<g3, ib::mac::y> egrid;
		point3 lgrid(0.5, 0.5, 0.5);
		lgrid[0] = 1.0;
		egrid.interpolate(lgrid);
		unsigned nz = 0;
		for (auto& e : egrid) if (e != 0.0) { e = 1; ++nz; }
		lgrid[0] = 0.0;
		lgrid.interpolate(egrid);
		std::printf("ib3_nonzero %u\nib3_value %.17g\n", nz, lgrid[0]);
	}
	// spreading to the nodes (eulerian_grid) and to the w-faces conserves
	// force: sum_grid f h^3 = sum_k F_k ds_k; a unit field interpolates to 1
	{
		ring3 r(0.45, 0.5, 0.3, 0.2);
		for (unsigned k = 0; k < 96; ++k) r[k] = std::cos(5.0 * k) + 0.25;
		ib::eulerian_grid<g3> f;
		ib::mac_grid<g3, ib::mac::z> fw;
		f.interpolate(r);
		fw.interpolate(r);
		constexpr unsigned M = g3::nx * g3::ny * g3::nz;
		double s = 0.0, sw = 0.0, t = 0.0;
		for (unsigned i = 0; i < M; ++i) { s += f[i]; sw += fw[i]; }
		for (unsigned k = 0; k < 96; ++k) t += r[k] * r.volume(k);
		ib::eulerian_grid<g3> one(1.0);
		r.interpolate(one);
		double unity = 0.0;
		for (unsigned k = 0; k < 96; ++k) unity = std::fmax(unity, std::fabs(r[k] - 1.0));
		std::printf("spread3_node_error %.3e\nspread3_w_error %.3e\ninterp3_unity_error %.3e\n",
		            std::fabs(s / M - t), std::fabs(sw / M - t), unity);
	}
	// mg_test2.cc: L * b on the width-8 cube at level 1 (h = 1/2, 16^3),
	// against the stencil written out above; L * 1 = 0
	{
		constexpr unsigned M = g8::nx * g8::ny * g8::nz;
		open_solver<g8> solver(1e-12, 2, 2);
		dt::vector<M> b, ones(1.0);
		unsigned long long st = 99;
		for (unsigned i = 0; i < M; ++i) {
			st = st * 6364136223846793005ull + 1442695040888963407ull;
			b[i] = (double) (st >> 11) / 9007199254740992.0 - 0.5;
		}
		dt::vector<M> lb = solver.L * b, l1 = solver.L * ones;
		double err = 0.0, zero = 0.0;
		for (unsigned i = 0; i < M; ++i) {
			err = std::fmax(err, std::fabs(lb[i] - stencil7<g8::nx, g8::ny, g8::nz>(b, i, g8::n)));
			zero = std::fmax(zero, std::fabs(l1[i]));
		}
		std::printf("L8_error %.3e\nL8_ones %.3e\n", err, zero);
	}
	// restr_test.cc: one coarse cell of 2 x 2 x 2 fine cells; every fine
	// cell's 64-point weights sum to 64/512, so coarse = mean(fine), added to
	// what coarse holds.  Prolongation of one coarse value is that value.
	{
		restriction_solver<g1> rs(1e-16, 2, 2);
		dt::vector<8> fine = {0.125, -0.5, 1.0, 0.25, -0.75, 2.0, 0.5, -1.0};
		dt::vector<1> coarse;
		rs.restriction(coarse, fine);
		const double first = coarse[0];
		rs.restriction(coarse, fine);
		open_solver<g1> os(1e-16, 2, 2);
		dt::vector<1> c1 = {0.75};
		dt::vector<8> f1(-3.0);
		os.interpolation(f1, c1);
		double pmax = 0.0;
		for (unsigned i = 0; i < 8; ++i) pmax = std::fmax(pmax, std::fabs(f1[i] - 0.75));
		std::printf("restr_first %.17g\nrestr_second %.17g\nprolong_error %.3e\n", first, coarse[0], pmax);
	}
	// alg::conjugate_gradient and alg::smooth on a matrix the program fills:
	// b = A x_true, solve from zero; one sweep against the reference's
	// lexicographic update written out here
	{
		tridiagonal A;
		dt::vector<200> xt, x, b;
		for (unsigned i = 0; i < 200; ++i) xt[i] = std::sin(0.1 * i) + 0.01 * i;
		b = A * xt;
		double ax = 0.0;
		for (unsigned i = 0; i < 200; ++i) {
			const double want = 2 * xt[i] - (i > 0 ? xt[i - 1] : 0.0) - (i + 1 < 200 ? xt[i + 1] : 0.0);
			ax = std::fmax(ax, std::fabs(b[i] - want));
		}
		const unsigned it = alg::conjugate_gradient(x, A, b, 1e-12);
		double e = 0.0;
		for (unsigned i = 0; i < 200; ++i) e = std::fmax(e, std::fabs(x[i] - xt[i]));
		dt::vector<200> y, yh;
		alg::smooth(y, A, b);
		for (unsigned i = 0; i < 200; ++i)
			yh[i] = (b[i] + (i > 0 ? yh[i - 1] : 0.0)) / 2.0;  // GS row i: (b_i + x_{i-1}) / 2, x_{i+1} still 0
		double se = 0.0;
		for (unsigned i = 0; i < 200; ++i) se = std::fmax(se, std::fabs(y[i] - yh[i]));
		std::printf("csr_apply_error %.3e\ncsr_cg_iters %u\ncsr_cg_error %.3e\ncsr_smooth_error %.3e\n", ax, it, e, se);
	}
	std::printf("done 1\n");
	return 0;
}
