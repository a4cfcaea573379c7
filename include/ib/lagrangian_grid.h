// ib/lagrangian_grid.h -- drop-in for the reference's ib/lagrangian_grid.h
// (ib/lagrangian_grid.h:33-67): markers at given positions with a value
// each; `volume` (the marker's arclength or area weight) stays abstract, as
// in the reference.  Transfers with the Eulerian grids run on the GPU;
// search_nearby keeps the reference's box query (dt/kdtree.h).
#pragma once
#include <initializer_list>
#include <vector>

#include "../dt/kdtree.h"
#include "grid.h"

namespace ib {

template<class Geometry, unsigned int Size>
class lagrangian_grid : public grid<Geometry, Size> {
public:
	using grid<Geometry, Size>::n;
	static constexpr unsigned int width = Geometry::width;
	static constexpr unsigned int height = Geometry::height;
	static constexpr unsigned int depth = Geometry::depth;
	using typename grid<Geometry, Size>::point_type;
	using typename grid<Geometry, Size>::value_type;
private:
	std::vector<point_type> positions;
	dt::kdtree<Geometry> tree;
protected:
	virtual int transfer_kind() const override { return 4; }
public:
	virtual point_type coordinates(unsigned int index) override { return positions[index]; }
	virtual std::vector<point_type> search_nearby(point_type coords) override
	{
		coords[0] = std::fmod(std::fmod(coords[0], width) + width, width);
		coords[2] = std::fmod(std::fmod(coords[2], depth) + depth, depth);
		return tree.search_nearby(coords);
	}
	lagrangian_grid(std::vector<point_type> p, value_type v = value_type())
		: grid<Geometry, Size>(v), positions(p), tree(p)
	{
		if (positions.size() != Size) throw std::length_error("lagrangian_grid: positions");
	}
	lagrangian_grid(const lagrangian_grid& g) : grid<Geometry, Size>(g), positions(g.positions), tree(g.positions) {}
	lagrangian_grid(std::vector<point_type> p, std::initializer_list<value_type> v)
		: grid<Geometry, Size>(v), positions(p), tree(p)
	{
		if (positions.size() != Size) throw std::length_error("lagrangian_grid: positions");
	}
};

}  // namespace ib
