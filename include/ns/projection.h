// ns/projection.h -- drop-in for the reference's ns/projection.h
// (ns/projection.h:37-122): project(u, v, w) removes the gradient of the
// potential that solves L phi = div u (multigrid V(2,2) to the absolute RMS
// tol, ns/projection.h:44), all on the GPU; boundary() applies the same
// potential's tangential correction to the wall vectors; get_phidt() is that
// potential.  On the plane: project(u, v).
#pragma once
#include "../afm/dropin_core.hpp"
#include "../dt/vector.h"

namespace ns {

template<class Geometry>
class projection {
	static constexpr unsigned n = Geometry::n;
	static constexpr unsigned nx = Geometry::nx;
	static constexpr unsigned ny = Geometry::ny;
	static constexpr unsigned nz = Geometry::nz;
	static constexpr unsigned N = nx * ny * nz, Nv = nx * (ny - 1) * nz, Nxz = nx * nz;
	dt::vector<N> phidtx, phidtz, phidt;
	dt::vector<Nv> phidty;
	double tol;
public:
	explicit projection(double tol_) : tol(tol_) {}

	// ns/projection.h:60-73 on the walled cube
	void project(dt::vector<N>& u, dt::vector<Nv>& v, dt::vector<N>& w)
	{
		afm_detail::require_unit_cube<Geometry>();
		const unsigned L = Geometry::level;
		afm_ctx* c = afm_detail::context3();
		afm_detail::Buffer bu(c, N * 8), bv(c, Nv * 8), bw(c, N * 8), zero(c, Nxz * 8), phi(c, N * 8), px(c, N * 8),
			py(c, Nv * 8), pz(c, N * 8);
		bu.put(u.data(), N); bv.put(v.data(), Nv); bw.put(w.data(), N);
		const dt::vector<Nxz> dummy;        // the zero wall values (ns/projection.h:66)
		zero.put(dummy.data(), Nxz);
		afm_detail::check(afm_div3(c, L, bu.d(), bv.d(), bw.d(), zero.d(), zero.d(), phi.d()), "afm_div3");
		int cyc = 0;
		afm_detail::check(afm_mg3_solve(c, L, 1, phi.d(), tol, 0, 2, 2, 100000, &cyc, nullptr), "afm_mg3_solve");
		afm_detail::check(afm_grad3(c, L, phi.d(), px.d(), py.d(), pz.d()), "afm_grad3");
		afm_detail::check(afm_axpy(c, N, -1.0, px.d(), bu.d()), "afm_axpy");
		afm_detail::check(afm_axpy(c, Nv, -1.0, py.d(), bv.d()), "afm_axpy");
		afm_detail::check(afm_axpy(c, N, -1.0, pz.d(), bw.d()), "afm_axpy");
		bu.get(u.data(), N); bv.get(v.data(), Nv); bw.get(w.data(), N);
		phi.get(phidt.data(), N); px.get(phidtx.data(), N); py.get(phidty.data(), Nv); pz.get(phidtz.data(), N);
	}
	// ns/projection.h:60-73 on the doubly periodic plane
	void project(dt::vector<N>& u, dt::vector<N>& v)
	{
		afm_ctx* c = afm_detail::context2<Geometry>();
		afm_detail::Buffer du(c, N * 8), dv(c, N * 8), p(c, N * 8);
		du.put(u.data(), N);
		dv.put(v.data(), N);
		afm_detail::check(afm_div(c, du.d(), dv.d(), p.d()), "afm_div");
		int cyc = 0;
		afm_detail::check(afm_mg_solve(c, p.d(), tol, 0, 2, 2, 0, 100000, &cyc, nullptr), "afm_mg_solve");
		afm_detail::check(afm_grad_sub(c, p.d(), du.d(), dv.d()), "afm_grad_sub");
		du.get(u.data(), N);
		dv.get(v.data(), N);
		p.get(phidt.data(), N);
	}
	// ns/projection.h:75-106: the wall vectors corrected by the last
	// projection's potential (the time step argument is unused, as there)
	void boundary(dt::vector<Nxz>& ut, dt::vector<Nxz>& ub, dt::vector<Nxz>& wt, dt::vector<Nxz>& wb, double)
	{
		afm_detail::require_unit_cube<Geometry>();
		afm_ctx* c = afm_detail::context3();
		afm_detail::Buffer px(c, N * 8), pz(c, N * 8), a(c, Nxz * 8), b(c, Nxz * 8), d(c, Nxz * 8), e(c, Nxz * 8);
		px.put(phidtx.data(), N); pz.put(phidtz.data(), N);
		a.put(ut.data(), Nxz); b.put(ub.data(), Nxz); d.put(wt.data(), Nxz); e.put(wb.data(), Nxz);
		afm_detail::check(afm_projection3_boundary(c, Geometry::level, px.d(), pz.d(), a.d(), b.d(), d.d(), e.d()),
		                  "afm_projection3_boundary");
		a.get(ut.data(), Nxz); b.get(ub.data(), Nxz); d.get(wt.data(), Nxz); e.get(wb.data(), Nxz);
	}
	dt::vector<N>& get_phidt() { return phidt; }
};

}  // namespace ns
