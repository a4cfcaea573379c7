// mg/laplacian.h -- drop-in for the reference's mg/laplacian.h
// (mg/laplacian.h:31-153): the cell-centred laplacian of geometry G (7-point,
// x and z periodic, Neumann walls in y; on the plane the 5-point periodic
// operator), as GPU stencil kernels in the reference's CSR summation order.
#pragma once
#include <stdexcept>

#include "../afm/dropin_core.hpp"
#include "../dt/matrix.h"

namespace mg {

template<class Geometry>
class laplacian : public dt::matrix<Geometry::nx * Geometry::ny * Geometry::nz> {
public:
	static constexpr unsigned int nx = Geometry::nx;
	static constexpr unsigned int ny = Geometry::ny;
	static constexpr unsigned int nz = Geometry::nz;
	static constexpr unsigned int n = Geometry::n;
	static constexpr unsigned int S = nx * ny * nz;
	laplacian() {}

	void apply(const dt::vector<S>& x, dt::vector<S>& y) const override
	{
		if (afm_detail::is_plane<Geometry>::value) {
			afm_ctx* c = afm_detail::context2<Geometry>();
			afm_detail::Buffer bx(c, S * 8), by(c, S * 8);
			bx.put(x.data(), S);
			afm_detail::check(afm_laplacian_apply(c, Geometry::level, bx.d(), by.d()), "afm_laplacian_apply");
			by.get(y.data(), S);
			return;
		}
		afm_detail::require_cube<Geometry>();
		afm_ctx* c = afm_detail::context3();
		afm_detail::Buffer bx(c, S * 8), by(c, S * 8);
		bx.put(x.data(), S);
		afm_detail::check(afm_laplacian3_apply(c, Geometry::level, Geometry::width, bx.d(), by.d()),
		                  "afm_laplacian3_apply");
		by.get(y.data(), S);
	}
	unsigned cg(dt::vector<S>& x, const dt::vector<S>& b, double tol) const override
	{
		if (afm_detail::is_plane<Geometry>::value)
			throw std::invalid_argument("alg::conjugate_gradient on the plane's laplacian is not on the GPU path "
			                            "(mg::solver solves it)");
		afm_detail::require_cube<Geometry>();
		afm_ctx* c = afm_detail::context3();
		afm_detail::Buffer bx(c, S * 8), bb(c, S * 8);
		bx.put(x.data(), S);
		bb.put(b.data(), S);
		int it = 0;
		afm_detail::check(afm_laplacian3_cg(c, Geometry::level, Geometry::width, bb.d(), bx.d(), tol, 1000000, &it,
		                                    nullptr), "afm_laplacian3_cg");
		bx.get(x.data(), S);
		return (unsigned) it;
	}
	void smooth(dt::vector<S>& x, const dt::vector<S>& b) const override
	{
		if (afm_detail::is_plane<Geometry>::value) {
			afm_ctx* c = afm_detail::context2<Geometry>();
			afm_detail::Buffer bx(c, S * 8), bb(c, S * 8), bo(c, S * 8);
			bx.put(x.data(), S);
			bb.put(b.data(), S);
			afm_detail::check(afm_mg_smooth(c, Geometry::level, bx.d(), bb.d(), bo.d(), 1, 0), "afm_mg_smooth");
			bo.get(x.data(), S);
			return;
		}
		afm_detail::require_cube<Geometry>();
		afm_ctx* c = afm_detail::context3();
		afm_detail::Buffer bx(c, S * 8), bb(c, S * 8);
		bx.put(x.data(), S);
		bb.put(b.data(), S);
		afm_detail::check(afm_mg3_smooth(c, Geometry::level, Geometry::width, bx.d(), bb.d(), 1), "afm_mg3_smooth");
		bx.get(x.data(), S);
	}
};

}  // namespace mg
