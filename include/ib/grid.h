// ib/grid.h -- drop-in for the reference's ib/grid.h (ib/grid.h:29-84): a grid
// of values with positions; interpolate(source) sets every value to the
// delta-weighted sum of the source's values near its position.  Here the
// transfer runs on the GPU between markers (ib::lagrangian_grid) and the
// Eulerian grids (ib::mac_grid components, ib::eulerian_grid's nodes):
// markers <- field is interpolation (afm_interp3 / afm_interp), field <-
// markers is spreading (afm_spread3 / afm_spread; per-point sums in marker
// order, no float atomics).  Other pairs throw std::invalid_argument.
#pragma once
#include <cmath>
#include <initializer_list>
#include <stdexcept>
#include <vector>

#include "../afm/dropin_core.hpp"
#include "../dt/point.h"
#include "../dt/vector.h"

// ib/grid.h:31: a long-double pi, which programs written for the reference use
#define pi 3.141592653589793238462643383279502884L

namespace ib {

namespace mac { enum component { x, y, z }; }

template<class Geometry, unsigned int Size>
class grid : public dt::vector<Size> {
public:
	static constexpr unsigned int n = Geometry::n;
	static constexpr unsigned int nx = Geometry::nx;
	static constexpr unsigned int ny = Geometry::ny;
	static constexpr unsigned int nz = Geometry::nz;

	typedef Geometry geometry_type;
	typedef dt::point point_type;
	typedef typename point_type::value_type coordinate_type;
	using typename dt::vector<Size>::value_type;
protected:
	virtual point_type coordinates(unsigned int) = 0;
	virtual std::vector<point_type> search_nearby(point_type) = 0;
	virtual double volume(unsigned int) const = 0;
	// ib/grid.h:51-55 with a correctly rounded cosine, as the GPU evaluates it
	virtual double delta_function(coordinate_type r)
	{
		double out = 0.0;
		afm_detail::check(afm_delta_host(&r, 1, n, &out), "afm_delta_host");
		return out;
	}
	// the GPU transfer the grid takes part in: 0, 1, 2 the MAC component u,
	// v, w; 3 the nodes of eulerian_grid; 4 markers; -1 none
	virtual int transfer_kind() const { return -1; }
public:
	template<unsigned int R> void interpolate(grid<Geometry, R>& source);
	grid(const grid& g) : dt::vector<Size>(g) {}
	grid(dt::vector<Size>& v) : dt::vector<Size>(v) {}
	grid(value_type v = 0.0) : dt::vector<Size>(v) {}
	grid(std::initializer_list<value_type> v) : dt::vector<Size>(v) {}
	using dt::vector<Size>::operator[];

	template<class, unsigned int> friend class grid;
};

template<class Geometry, unsigned int Size>
template<unsigned int R>
void grid<Geometry, Size>::interpolate(grid<Geometry, R>& source)
{
	const bool plane = afm_detail::is_plane<Geometry>::value;
	const int mine = transfer_kind(), theirs = source.transfer_kind();
	const bool spread = mine >= 0 && mine <= 3 && theirs == 4;
	const bool interp = mine == 4 && theirs >= 0 && theirs <= 3;
	if (!spread && !interp)
		throw std::invalid_argument("ib::grid::interpolate runs between markers and an Eulerian grid");
	const int comp = spread ? mine : theirs;
	if (plane && comp > 1) throw std::invalid_argument("the plane has the MAC components u and v only");
	if (!plane) afm_detail::require_unit_cube<Geometry>();
	grid<Geometry, Size>* markers_this = this;
	const unsigned m = spread ? R : Size;          // markers
	const unsigned nf = spread ? Size : R;         // field values
	const unsigned dim = plane ? 2 : 3;
	std::vector<double> X((size_t) dim * m), ds(m);
	for (unsigned k = 0; k < m; ++k) {
		const point_type p = spread ? source.coordinates(k) : markers_this->coordinates(k);
		for (unsigned a = 0; a < dim; ++a) X[(size_t) dim * k + a] = p[a];
		if (spread) ds[k] = source.volume(k);
	}
	afm_ctx* c = plane ? afm_detail::context2<Geometry>() : afm_detail::context3();
	afm_detail::Buffer bX(c, X.size() * 8), bm(c, (size_t) m * 8), bf(c, (size_t) nf * 8);
	bX.put(X.data(), X.size());
	if (spread) {
		afm_detail::Buffer bds(c, (size_t) m * 8);
		bds.put(ds.data(), m);
		bm.put(source.data(), m);
		if (plane)
			afm_detail::check(afm_spread(c, comp, bm.d(), bds.d(), bX.d(), m, bf.d()), "afm_spread");
		else
			afm_detail::check(afm_spread3(c, Geometry::level, comp, bm.d(), bds.d(), bX.d(), m, bf.d()), "afm_spread3");
		bf.get(this->data(), nf);
	} else {
		bf.put(source.data(), nf);
		if (plane)
			afm_detail::check(afm_interp(c, comp, bf.d(), bX.d(), m, bm.d()), "afm_interp");
		else
			afm_detail::check(afm_interp3(c, Geometry::level, comp, bf.d(), bX.d(), m, bm.d()), "afm_interp3");
		bm.get(this->data(), m);
	}
}

}  // namespace ib
