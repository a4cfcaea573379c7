// dt/matrix.h -- drop-in for the reference's dt/matrix.h (dt/matrix.h:35-94):
// a square operator on dt::vector<Size>.  The reference's operators
// (mg::laplacian, ns::helmholtz_*) fill CSR arrays; here they are stencil
// kernels on the GPU and override apply / cg / smooth.  A matrix a program
// fills itself through the protected CSR arrays (row_start, column, values,
// ascending columns per row) runs on the GPU as device CSR: operator* and
// alg::conjugate_gradient in the reference's summation order per row,
// alg::smooth as the reference's lexicographic sweep.
#pragma once
#include <ostream>
#include <vector>

#include "../afm/dropin_core.hpp"
#include "vector.h"

namespace dt {

template<unsigned int Size>
class matrix {
public:
	typedef double value_type;
	typedef std::size_t size_type;
protected:
	std::vector<std::size_t> row_start;
	std::vector<std::size_t> column;
	std::vector<value_type> values;

	struct Device {   // the CSR arrays in device memory, for one call
		afm_detail::Buffer rs, col, val;
		Device(afm_ctx* c, const matrix& m)
			: rs(c, (Size + 1) * 8), col(c, m.column.size() * 8), val(c, m.values.size() * 8)
		{
			if (m.row_start.size() != Size + 1 || m.column.size() != m.values.size() ||
			    m.row_start[Size] != m.values.size())
				throw std::invalid_argument("dt::matrix: inconsistent CSR arrays");
			rs.put((const void*) m.row_start.data(), (Size + 1) * 8);
			col.put((const void*) m.column.data(), m.column.size() * 8);
			val.put(m.values.data(), m.values.size());
		}
	};
public:
	matrix() : row_start(Size + 1, 0) {}
	virtual ~matrix() {}

	// y = A x (dt/matrix.h:81-94: each row summed from 0 in column order)
	virtual void apply(const vector<Size>& x, vector<Size>& y) const
	{
		afm_ctx* c = afm_detail::context3();
		Device d(c, *this);
		afm_detail::Buffer bx(c, Size * 8), by(c, Size * 8);
		bx.put(x.data(), Size);
		afm_detail::check(afm_csr_spmv(c, Size, d.rs.template as<unsigned long>(), d.col.template as<unsigned long>(), d.val.d(),
		                               bx.d(), by.d()), "afm_csr_spmv");
		by.get(y.data(), Size);
	}
	// alg::conjugate_gradient (alg/cg.h:32-56): x is the start and the result
	virtual unsigned cg(vector<Size>& x, const vector<Size>& b, double tol) const
	{
		afm_ctx* c = afm_detail::context3();
		Device d(c, *this);
		afm_detail::Buffer bx(c, Size * 8), bb(c, Size * 8);
		bx.put(x.data(), Size);
		bb.put(b.data(), Size);
		int it = 0;
		afm_detail::check(afm_csr_cg(c, Size, d.rs.template as<unsigned long>(), d.col.template as<unsigned long>(), d.val.d(),
		                             bb.d(), bx.d(), tol, 1000000, &it, nullptr), "afm_csr_cg");
		bx.get(x.data(), Size);
		return (unsigned) it;
	}
	// alg::smooth (alg/smooth.h:35-51): one lexicographic sweep
	virtual void smooth(vector<Size>& x, const vector<Size>& b) const
	{
		afm_ctx* c = afm_detail::context3();
		Device d(c, *this);
		afm_detail::Buffer bx(c, Size * 8), bb(c, Size * 8);
		bx.put(x.data(), Size);
		bb.put(b.data(), Size);
		afm_detail::check(afm_csr_smooth(c, Size, d.rs.template as<unsigned long>(), d.col.template as<unsigned long>(), d.val.d(),
		                                 bx.d(), bb.d(), 1), "afm_csr_smooth");
		bx.get(x.data(), Size);
	}

	vector<Size> operator*(const vector<Size>& right) const
	{
		vector<Size> result;
		apply(right, result);
		return result;
	}
	size_type size() const { return Size; }

	// "(row, col) = value" per stored entry, 1-based (dt/matrix.h:54-75)
	friend std::ostream& operator<<(std::ostream& out, const matrix<Size>& mat)
	{
		for (std::size_t row = 0; row < Size && row + 1 < mat.row_start.size(); ++row)
			for (std::size_t k = mat.row_start[row]; k < mat.row_start[row + 1]; ++k)
				out << '(' << (row + 1) << ", " << (mat.column[k] + 1) << ") = " << mat.values[k] << std::endl;
		return out;
	}
};

}  // namespace dt
