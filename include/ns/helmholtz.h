// ns/helmholtz.h -- drop-in for the reference's ns/helmholtz.h
// (ns/helmholtz.h:29-337): I + scale L with the wall closures, as GPU
// stencil kernels in the reference's CSR summation order.
//   helmholtz_periodic<G>(scale): cell rows (u, w); the second-order wall
//     closure in the first and last row (kind 1); on the plane, I + scale L5;
//   helmholtz_dirichlet<G>(scale): the nx (ny-1) nz interior y-face rows (v),
//     Dirichlet at the walls (kind 3; walled geometries only).
// boundary(top, bottom) gives the wall terms of a right-hand side; cg is
// alg::conjugate_gradient on the operator.
#pragma once
#include <stdexcept>

#include "../afm/dropin_core.hpp"
#include "../dt/matrix.h"

namespace ns {

template<class Geometry, unsigned int Size>
class helmholtz : public dt::matrix<Size> {
protected:
	double scale;
	explicit helmholtz(double s) : scale(s) {}
	// y = (I + scale L) x, or the CG solve, on the walled cube (kind 1 or 3)
	static void apply3(int kind, double scale, const dt::vector<Size>& x, dt::vector<Size>& y)
	{
		afm_detail::require_unit_cube<Geometry>();
		afm_ctx* c = afm_detail::context3();
		afm_detail::Buffer bx(c, Size * 8), by(c, Size * 8);
		bx.put(x.data(), Size);
		afm_detail::check(afm_helmholtz3_apply(c, Geometry::level, kind, scale, bx.d(), by.d()), "afm_helmholtz3_apply");
		by.get(y.data(), Size);
	}
	static unsigned cg3(int kind, double scale, dt::vector<Size>& x, const dt::vector<Size>& b, double tol)
	{
		afm_detail::require_unit_cube<Geometry>();
		afm_ctx* c = afm_detail::context3();
		afm_detail::Buffer bx(c, Size * 8), bb(c, Size * 8);
		bx.put(x.data(), Size);
		bb.put(b.data(), Size);
		int it = 0;
		afm_detail::check(afm_helmholtz3_cg(c, Geometry::level, kind, scale, bb.d(), bx.d(), tol, 1000000, &it,
		                                    nullptr), "afm_helmholtz3_cg");
		bx.get(x.data(), Size);
		return (unsigned) it;
	}
	template<unsigned Out>
	static dt::vector<Out> boundary3(int kind, double scale, const dt::vector<Geometry::nx * Geometry::nz>& top,
	                                 const dt::vector<Geometry::nx * Geometry::nz>& bottom)
	{
		afm_detail::require_unit_cube<Geometry>();
		constexpr unsigned W = Geometry::nx * Geometry::nz;
		afm_ctx* c = afm_detail::context3();
		afm_detail::Buffer bt(c, W * 8), bb(c, W * 8), bo(c, Out * 8);
		bt.put(top.data(), W);
		bb.put(bottom.data(), W);
		afm_detail::check(afm_helmholtz3_boundary(c, Geometry::level, kind, scale, bt.d(), bb.d(), bo.d()),
		                  "afm_helmholtz3_boundary");
		dt::vector<Out> out;
		bo.get(out.data(), Out);
		return out;
	}
public:
	void smooth(dt::vector<Size>&, const dt::vector<Size>&) const override
	{
		throw std::invalid_argument("alg::smooth on a Helmholtz operator is not on the GPU path");
	}
};

template<class Geometry>
class helmholtz_periodic : public helmholtz<Geometry, Geometry::nx * Geometry::ny * Geometry::nz> {
public:
	static constexpr unsigned int nx = Geometry::nx;
	static constexpr unsigned int ny = Geometry::ny;
	static constexpr unsigned int nz = Geometry::nz;
	static constexpr unsigned int n = Geometry::n;
	static constexpr unsigned int S = nx * ny * nz;
private:
	typedef helmholtz<Geometry, S> base;
public:
	explicit helmholtz_periodic(double s) : base(s) {}
	void apply(const dt::vector<S>& x, dt::vector<S>& y) const override
	{
		if (afm_detail::is_plane<Geometry>::value) {
			afm_ctx* c = afm_detail::context2<Geometry>();
			afm_detail::Buffer bx(c, S * 8), by(c, S * 8);
			bx.put(x.data(), S);
			afm_detail::check(afm_helmholtz_apply(c, this->scale, bx.d(), by.d()), "afm_helmholtz_apply");
			by.get(y.data(), S);
			return;
		}
		base::apply3(1, this->scale, x, y);
	}
	unsigned cg(dt::vector<S>& x, const dt::vector<S>& b, double tol) const override
	{
		if (afm_detail::is_plane<Geometry>::value) {
			afm_ctx* c = afm_detail::context2<Geometry>();
			afm_detail::Buffer bx(c, S * 8), bb(c, S * 8);
			bx.put(x.data(), S);
			bb.put(b.data(), S);
			int it = 0;
			afm_detail::check(afm_helmholtz_cg(c, this->scale, bb.d(), bx.d(), tol, 0, 1000000, &it, nullptr),
			                  "afm_helmholtz_cg");
			bx.get(x.data(), S);
			return (unsigned) it;
		}
		return base::cg3(1, this->scale, x, b, tol);
	}
	// ns/helmholtz.h:177-196 (the plane has no walls: zero)
	dt::vector<S> boundary(dt::vector<nx * nz>& top, dt::vector<nx * nz>& bottom)
	{
		if (afm_detail::is_plane<Geometry>::value) return dt::vector<S>();
		return base::template boundary3<S>(1, this->scale, top, bottom);
	}
};

template<class Geometry>
class helmholtz_dirichlet : public helmholtz<Geometry, Geometry::nx * (Geometry::ny - 1) * Geometry::nz> {
public:
	static constexpr unsigned int nx = Geometry::nx;
	static constexpr unsigned int ny = Geometry::ny;
	static constexpr unsigned int nz = Geometry::nz;
	static constexpr unsigned int n = Geometry::n;
	static constexpr unsigned int S = nx * (ny - 1) * nz;
private:
	typedef helmholtz<Geometry, S> base;
public:
	explicit helmholtz_dirichlet(double s) : base(s) {}
	void apply(const dt::vector<S>& x, dt::vector<S>& y) const override { base::apply3(3, this->scale, x, y); }
	unsigned cg(dt::vector<S>& x, const dt::vector<S>& b, double tol) const override
	{
		return base::cg3(3, this->scale, x, b, tol);
	}
	// ns/helmholtz.h:321-337
	dt::vector<S> boundary(dt::vector<nx * nz>& top, dt::vector<nx * nz>& bottom)
	{
		return base::template boundary3<S>(3, this->scale, top, bottom);
	}
};

}  // namespace ns
