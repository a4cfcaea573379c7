// dt/vector.h -- drop-in for the reference's dt/vector.h (dt/vector.h:33-165):
// a value-semantic vector of doubles with the reference's operators, RMS
// `abs` and printed form.
#pragma once
#include <cmath>
#include <initializer_list>
#include <ostream>
#include <stdexcept>
#include <vector>

namespace dt {

template<unsigned int Size>
class vector {
protected:
	std::vector<double> values;
public:
	typedef double value_type;
	typedef std::vector<double>::iterator iterator;
	typedef std::vector<double>::const_iterator const_iterator;
	typedef double& reference;
	typedef const double& const_reference;
	typedef double* pointer;
	typedef const double* const_pointer;
	typedef std::size_t size_type;

	vector(double value = 0.0) : values(Size, value) {}
	vector(std::initializer_list<double> v) : values(v)
	{
		if (v.size() != Size) throw std::length_error("dt::vector: initializer size");
	}
	vector(const vector& v) = default;
	vector& operator=(const vector& v) = default;
	virtual ~vector() {}
	vector& operator=(std::initializer_list<double> v)
	{
		if (v.size() != Size) throw std::length_error("dt::vector: initializer size");
		values = v;
		return *this;
	}
	virtual reference operator[](unsigned i) { return values[i]; }
	virtual const_reference operator[](unsigned i) const { return values[i]; }
	size_type size() const { return Size; }
	iterator begin() { return values.begin(); }
	iterator end() { return values.end(); }
	const_iterator begin() const { return values.begin(); }
	const_iterator end() const { return values.end(); }
	const_iterator cbegin() const { return values.cbegin(); }
	const_iterator cend() const { return values.cend(); }
	double* data() { return values.data(); }
	const double* data() const { return values.data(); }
	// dot product, summed in index order (dt/vector.h:83-92)
	value_type operator*(const vector& o) const
	{
		double r = 0.0;
		for (unsigned i = 0; i < Size; ++i) r += values[i] * o.values[i];
		return r;
	}
	vector& operator+=(const vector& o) { for (unsigned i = 0; i < Size; ++i) values[i] += o.values[i]; return *this; }
	vector& operator-=(const vector& o) { for (unsigned i = 0; i < Size; ++i) values[i] -= o.values[i]; return *this; }
	vector& operator*=(double c) { for (double& v : values) v *= c; return *this; }
	vector& operator/=(double c) { for (double& v : values) v /= c; return *this; }
};

// RMS (dt/vector.h:75-81)
template<unsigned S> double abs(const vector<S>& v)
{
	double m = 0.0;
	for (auto it = v.cbegin(); it != v.cend(); ++it) m += *it * *it;
	return std::sqrt(m) / std::sqrt((double) S);
}
template<unsigned S> vector<S> operator+(vector<S> a, const vector<S>& b) { return a += b; }
template<unsigned S> vector<S> operator-(vector<S> a, const vector<S>& b) { return a -= b; }
template<unsigned S> vector<S> operator*(double c, vector<S> v) { return v *= c; }
template<unsigned S> vector<S> operator/(vector<S> v, double c) { return v /= c; }
// "[a, b, ...]'" (dt/vector.h:153-165)
template<unsigned S> std::ostream& operator<<(std::ostream& out, const vector<S>& v)
{
	out << "[";
	for (unsigned i = 0; i < S; ++i) out << (i ? ", " : "") << v[i];
	return out << "]'";
}

}  // namespace dt
