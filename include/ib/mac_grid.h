// ib/mac_grid.h -- drop-in for the reference's ib/mac_grid.h
// (ib/mac_grid.h:30-136): one MAC velocity component.  On the walled
// geometry: u, w at the x- and z-faces (shifted by h/2 in the other two
// axes), v at the nx (ny-1) nz interior y-faces ((i+1/2, j+1, k+1/2)/n).  On
// the plane: u at (i, j+1/2)/n, v at (i+1/2, j)/n.  Transfers with markers
// run on the GPU.
#pragma once
#include <vector>

#include "eulerian_grid.h"

namespace ib {

template<class G, mac::component C>
struct mac_size {
	static constexpr unsigned value =
		(C == mac::y && !afm_detail::is_plane<G>::value) ? G::nx * (G::ny - 1) * G::nz : G::nx * G::ny * G::nz;
};

template<class Geometry, mac::component Component>
class mac_grid : public eulerian_grid<Geometry, mac_size<Geometry, Component>::value> {
	static constexpr unsigned int Size = mac_size<Geometry, Component>::value;
	typedef eulerian_grid<Geometry, Size> base;
	static constexpr bool plane = afm_detail::is_plane<Geometry>::value;
public:
	using base::n;
	using base::nx;
	using base::ny;
	using base::nz;
	using typename base::point_type;
	using typename base::value_type;
protected:
	// ib/mac_grid.h:60-73 and 107-136
	virtual std::vector<point_type> search_nearby(point_type point) override
	{
		if (plane || Component != mac::y) {
			const int a = plane ? (Component == mac::x ? 1 : 0) : (Component + 1) % 3;
			const int b = plane ? -1 : (Component + 2) % 3;
			point[a] -= 0.5 / n;
			if (b >= 0) point[b] -= 0.5 / n;
			auto results = base::search_nearby(point);
			for (auto& p : results) {
				p[a] += 0.5 / n;
				if (b >= 0) p[b] += 0.5 / n;
			}
			return results;
		}
		std::vector<point_type> results;
		point[0] = std::fmod(std::fmod(point[0], base::width) + base::width, base::width);
		point[2] = std::fmod(std::fmod(point[2], base::depth) + base::depth, base::depth);
		const double fx = std::floor(point[0] * n - 0.5), fy = std::floor(point[1] * n - 1.0),
		             fz = std::floor(point[2] * n - 0.5);
		for (int i = 0; i < 64; ++i) {
			const int dix = i % 4, diy = (i / 4) % 4, diz = (i / 16) % 4;
			const std::size_t ix = ((std::size_t) fx + dix + nx - 1) % nx, iy = (std::size_t) fy + diy,
			                  iz = ((std::size_t) fz + diz + nz - 1) % nz;
			const double x = (fx + dix - 1.0) / n, y = (fy + diy) / n, z = (fz + diz - 1.0) / n;
			if (y < 0 || y > base::height) continue;
			results.push_back(point_type(x + 0.5 / n, y, z + 0.5 / n, ix + nx * (iy - 1 + (ny - 1) * iz)));
		}
		return results;
	}
	virtual int transfer_kind() const override { return (int) Component; }
public:
	// ib/mac_grid.h:49-58 and 97-105
	virtual point_type coordinates(unsigned int i) override
	{
		if (plane) {
			const double x = (double) (i % nx) / n, y = (double) (i / nx) / n;
			return Component == mac::x ? point_type(x, y + 0.5 / n, i) : point_type(x + 0.5 / n, y, i);
		}
		if (Component == mac::y)
			return point_type(((double) (i % nx) + 0.5) / n, ((double) (i / nx % (ny - 1)) + 1.0) / n,
			                  ((double) (i / (nx * (ny - 1))) + 0.5) / n, i);
		point_type p = base::coordinates(i);
		p[(Component + 1) % 3] += 0.5 / n;
		p[(Component + 2) % 3] += 0.5 / n;
		return p;
	}
	mac_grid(value_type v = value_type()) : base(v) {}
	mac_grid(std::initializer_list<value_type> v) : base(v) {}
};

}  // namespace ib
