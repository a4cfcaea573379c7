// dt/point.h -- drop-in for the reference's dt/point.h (dt/point.h:29-54):
// coordinates and an index.  The plane geometry uses the 3-argument form
// (x, y, index).
#pragma once
#include <cstddef>
#include <ostream>

namespace dt {

class point {
public:
	typedef double value_type;
	typedef value_type& reference;
	typedef const value_type& const_reference;
	typedef std::size_t size_type;
private:
	value_type values[3];
	size_type index;
public:
	point(value_type x, value_type y, value_type z, const size_type i) : values{x, y, z}, index(i) {}
	point(value_type x, value_type y, const size_type i) : values{x, y, 0.0}, index(i) {}
	point() : values{0.0, 0.0, 0.0}, index(0) {}
	reference operator[](size_type i) { return values[i]; }
	const_reference operator[](size_type i) const { return values[i]; }
	size_type get_index() const { return index; }
};

// "index (x, y, z)" (dt/point.h:50-54)
inline std::ostream& operator<<(std::ostream& out, const point& p)
{
	return out << p.get_index() << " (" << p[0] << ", " << p[1] << ", " << p[2] << ")";
}

}  // namespace dt
