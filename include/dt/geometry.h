// dt/geometry.h -- drop-in for the reference's dt/geometry.h (dt/geometry.h:29-60).
#pragma once
#include "../afm/dropin_core.hpp"
#include "vector.h"

namespace dt {

// The reference's geometry: n = 2^Level cells per unit length, a box of
// Width x Height x Depth units (walls in y), with its coarse and fine
// neighbours in the level hierarchy.
template<unsigned Level, unsigned Width, unsigned Height, unsigned Depth>
class geometry {
public:
	typedef geometry<Level - 1, Width, Height, Depth> coarse;
	typedef geometry<Level + 1, Width, Height, Depth> fine;
	static constexpr unsigned factor = 2;
	static constexpr unsigned n = 1u << Level;
	static constexpr unsigned nx = Width * n;
	static constexpr unsigned ny = Height * n;
	static constexpr unsigned nz = Depth * n;
	static constexpr unsigned size = nx * ny * nz;
	static constexpr unsigned width = Width;
	static constexpr unsigned height = Height;
	static constexpr unsigned depth = Depth;
	static constexpr unsigned level = Level;
};

template<unsigned Width, unsigned Height, unsigned Depth>
class geometry<0u, Width, Height, Depth> {
public:
	typedef geometry<1u, Width, Height, Depth> fine;
	static constexpr unsigned factor = 2;
	static constexpr unsigned n = 1;
	static constexpr unsigned nx = Width;
	static constexpr unsigned ny = Height;
	static constexpr unsigned nz = Depth;
	static constexpr unsigned size = nx * ny * nz;
	static constexpr unsigned width = Width;
	static constexpr unsigned height = Height;
	static constexpr unsigned depth = Depth;
	static constexpr unsigned level = 0u;
};

// The doubly periodic plane of the benchmark (no reference counterpart: the
// reference's geometry always has walls in y): n = 2^Level, nx x ny cells,
// (x, y) coordinates; nz = 1.
template<unsigned Level, unsigned Width = 1, unsigned Depth = 1>
struct plane {
	static constexpr unsigned level = Level, width = Width, depth = Depth, height = 1;
	static constexpr unsigned n = 1u << Level;
	static constexpr unsigned nx = Width * n, ny = Depth * n, nz = 1, size = nx * ny;
};

}  // namespace dt
