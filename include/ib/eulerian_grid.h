// ib/eulerian_grid.h -- drop-in for the reference's ib/eulerian_grid.h
// (ib/eulerian_grid.h:30-93): values at the nodes (i, j, k)/n (x fastest),
// volume h^3 (h^2 on the plane), the 4 x 4 x 4 search window of
// ib/eulerian_grid.h:64-93.  Its transfers with markers run on the GPU (the
// node grid, component 3 of afm_interp3 / afm_spread3).
#pragma once
#include <cmath>
#include <initializer_list>
#include <vector>

#include "grid.h"

namespace ib {

template<class Geometry, unsigned int Size = Geometry::nx * Geometry::ny * Geometry::nz>
class eulerian_grid : public grid<Geometry, Size> {
public:
	using grid<Geometry, Size>::n;
	using grid<Geometry, Size>::nx;
	using grid<Geometry, Size>::ny;
	using grid<Geometry, Size>::nz;
	using typename grid<Geometry, Size>::point_type;
	using typename grid<Geometry, Size>::value_type;
protected:
	static constexpr unsigned int width = Geometry::width;
	static constexpr unsigned int height = Geometry::height;
	static constexpr unsigned int depth = Geometry::depth;

	// ib/eulerian_grid.h:64-93: nodes whose box of half-width 2/n holds coords
	virtual std::vector<point_type> search_nearby(point_type coords) override
	{
		std::vector<point_type> results;
		coords[0] = std::fmod(std::fmod(coords[0], width) + width, width);
		coords[2] = std::fmod(std::fmod(coords[2], depth) + depth, depth);
		const double fx = std::floor(coords[0] * n), fy = std::floor(coords[1] * n), fz = std::floor(coords[2] * n);
		for (int i = 0; i < 64; ++i) {
			const int dix = i % 4, diy = (i / 4) % 4, diz = (i / 16) % 4;
			const std::size_t ix = ((std::size_t) fx + dix + nx - 1) % nx, iy = (std::size_t) fy + diy,
			                  iz = ((std::size_t) fz + diz + nz - 1) % nz;
			const double x = (fx + dix - 1.0) / n, y = (fy + diy - 1.0) / n, z = (fz + diz - 1.0) / n;
			if (y < 0 || y > height) continue;
			results.push_back(point_type(x, y, z, ix + nx * (iy - 1 + ny * iz)));
		}
		return results;
	}
	virtual double volume(unsigned int) const override
	{
		return afm_detail::is_plane<Geometry>::value ? 1. / (n * n) : 1. / (n * n * n);
	}
	virtual point_type coordinates(unsigned int index) override
	{
		return point_type(((double) (index % nx)) / n, ((double) (index % (nx * ny) / nx)) / n,
		                  ((double) (index / (nx * ny))) / n, index);
	}
	virtual int transfer_kind() const override { return Size == nx * ny * nz ? 3 : -1; }
public:
	eulerian_grid(value_type v = value_type()) : grid<Geometry, Size>(v) {}
	eulerian_grid(grid<Geometry, Size>& v) : grid<Geometry, Size>(v) {}
	eulerian_grid(const eulerian_grid& v) : grid<Geometry, Size>(v) {}
	eulerian_grid(std::initializer_list<value_type> v) : grid<Geometry, Size>(v) {}
};

}  // namespace ib
