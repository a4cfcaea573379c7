// dt/kdtree.h -- drop-in for the reference's dt/kdtree.h (dt/kdtree.h:39-286):
// the box query the reference's lagrangian_grid spreads through.  The GPU
// spreading bins markers by cell instead (afm_spread, afm_spread3); this
// class keeps the query available to programs that call it: the points
// within 2/n of coords in every axis (inclusive), x and z periodic with the
// geometry's width and depth, periodic images shifted as the reference
// returns them.  A plain host scan, not a hot path.
#pragma once
#include <cmath>
#include <vector>

#include "point.h"

namespace dt {

template<class Geometry, unsigned int = 0>
class kdtree {
public:
	typedef point point_type;
private:
	std::vector<point_type> pts;
public:
	explicit kdtree(std::vector<point_type> list) : pts(std::move(list)) {}
	std::vector<point_type> search_nearby(point_type coords) const
	{
		const double h = 2. / Geometry::n, W = Geometry::width, D = Geometry::depth;
		std::vector<point_type> out;
		for (const point_type& p : pts)
			for (int a = -1; a <= 1; ++a)
				for (int b = -1; b <= 1; ++b) {
					point_type q = p;
					q[0] += a * W;
					q[2] += b * D;
					if (std::fabs(q[0] - coords[0]) <= h && std::fabs(q[1] - coords[1]) <= h &&
					    std::fabs(q[2] - coords[2]) <= h)
						out.push_back(q);
				}
		return out;
	}
};

}  // namespace dt
