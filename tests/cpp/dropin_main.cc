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
		typedef dt::geometry<4u, 1u, 1u, 1u> g3;
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
		typedef dt::geometry<4u, 1u, 1u, 1u> g3;
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
	std::printf("done 1\n");
	return 0;
}
