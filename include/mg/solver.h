// mg/solver.h -- drop-in for the reference's mg/solver.h (mg/solver.h:35-271):
// mg::solver<G>(tol, nu1, nu2).solve(rhs) replaces rhs by the solution of
// L x = rhs - mean(rhs) and returns the V-cycles, stopping when the residual's
// RMS is at most tol (the reference's absolute test).  The hierarchy is the
// reference's (nested start, V(nu1, nu2) cycles, the coarsest level by CG);
// the smoother is red-black Gauss-Seidel on the GPU.  L, restriction and
// interpolation are protected static members, reachable by a subclass's
// `using`, as the reference's test programs reach them.
#pragma once
#include <vector>

#include "../afm/dropin_core.hpp"
#include "../dt/vector.h"
#include "laplacian.h"

namespace mg {

template<class G>
struct coarse_size {
	static constexpr unsigned value = afm_detail::is_plane<G>::value ? G::nx * G::ny / 4 : G::nx * G::ny * G::nz / 8;
};

template<class Geometry, unsigned Level = Geometry::level>
class solver {
public:
	static constexpr unsigned n = Geometry::n;
	static constexpr unsigned nx = Geometry::nx;
	static constexpr unsigned ny = Geometry::ny;
	static constexpr unsigned nz = Geometry::nz;
	typedef Geometry geometry_type;
	typedef laplacian<geometry_type> laplacian_type;
private:
	double tol;
	unsigned int nu1, nu2;
protected:
	static laplacian_type L;
	// mg/solver.h:71-151: fine (every value written) from coarse
	static void interpolation(dt::vector<nx * ny * nz>& fine, dt::vector<coarse_size<Geometry>::value>& coarse)
	{
		constexpr unsigned N = nx * ny * nz, Nc = coarse_size<Geometry>::value;
		const bool plane = afm_detail::is_plane<Geometry>::value;
		afm_ctx* c = plane ? afm_detail::context2<Geometry>() : afm_detail::context3();
		afm_detail::Buffer bf(c, N * 8), bc(c, Nc * 8);
		bc.put(coarse.data(), Nc);
		if (plane) {
			afm_detail::check(afm_mg_prolong(c, Geometry::level, bc.d(), bf.d()), "afm_mg_prolong");
		} else {
			afm_detail::require_cube<Geometry>();
			afm_detail::check(afm_mg3_prolong(c, Geometry::level, Geometry::width, bc.d(), bf.d()), "afm_mg3_prolong");
		}
		bf.get(fine.data(), N);
	}
	// mg/solver.h:153-181: coarse += full weighting of fine (the reference
	// accumulates into coarse; its callers pass zeros, and then the result
	// is bit-identical; otherwise it is rounded once more than there)
	static void restriction(dt::vector<coarse_size<Geometry>::value>& coarse, dt::vector<nx * ny * nz>& fine)
	{
		constexpr unsigned N = nx * ny * nz, Nc = coarse_size<Geometry>::value;
		const bool plane = afm_detail::is_plane<Geometry>::value;
		afm_ctx* c = plane ? afm_detail::context2<Geometry>() : afm_detail::context3();
		afm_detail::Buffer bf(c, N * 8), bc(c, Nc * 8);
		bf.put(fine.data(), N);
		if (plane) {
			afm_detail::check(afm_mg_restrict(c, Geometry::level, bf.d(), bc.d()), "afm_mg_restrict");
		} else {
			afm_detail::require_cube<Geometry>();
			afm_detail::check(afm_mg3_restrict(c, Geometry::level, Geometry::width, bf.d(), bc.d()),
			                  "afm_mg3_restrict");
		}
		std::vector<double> r(Nc);
		bc.get(r.data(), Nc);
		for (unsigned i = 0; i < Nc; ++i) coarse[i] += r[i];
	}
public:
	// on the plane, the form of afm_mg_solve (0: the reference's iteration in
	// FP64; 1: FP32 V-cycles preconditioning FP64 flexible CG)
	int precision = 0;

	solver(double tol_, unsigned int nu1_, unsigned int nu2_) : tol(tol_), nu1(nu1_), nu2(nu2_) {}
	unsigned int solve(dt::vector<nx * ny * nz>& rhs)
	{
		constexpr unsigned N = nx * ny * nz;
		int cyc = 0;
		if (afm_detail::is_plane<Geometry>::value) {
			afm_ctx* c = afm_detail::context2<Geometry>();
			afm_detail::Buffer b(c, N * 8);
			b.put(rhs.data(), N);
			afm_detail::check(afm_mg_solve(c, b.d(), tol, 0, (int) nu1, (int) nu2, precision, 100000, &cyc, nullptr),
			                  "afm_mg_solve");
			b.get(rhs.data(), N);
		} else {
			afm_detail::require_cube<Geometry>();
			afm_ctx* c = afm_detail::context3();
			afm_detail::Buffer b(c, N * 8);
			b.put(rhs.data(), N);
			afm_detail::check(afm_mg3_solve(c, Geometry::level, Geometry::width, b.d(), tol, 0, (int) nu1, (int) nu2,
			                                100000, &cyc, nullptr), "afm_mg3_solve");
			b.get(rhs.data(), N);
		}
		return (unsigned) cyc;
	}
};

template<class Geometry, unsigned Level> laplacian<Geometry> solver<Geometry, Level>::L;

}  // namespace mg
