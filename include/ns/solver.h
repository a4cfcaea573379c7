// ns/solver.h -- drop-in for the reference's ns/solver.h (ns/solver.h:37-140).
// ns::solver<G>(mu, rho, dt, tol, u0, v0, w0) aliases the caller's velocity
// vectors (ns/solver.h:54-56, 68), projects them, and each step() runs the
// backward-Euler half step, then the Crank-Nicolson full step, alternately,
// on the GPU (afm_ns3_*: geometry<L, 1, 1, 1>, walls in y; afm_ns_* on the
// doubly periodic plane, where step takes the two force components).
#pragma once
#include "../afm/dropin_core.hpp"
#include "../dt/geometry.h"
#include "../dt/vector.h"

namespace ns {

// ns::solver<G>(mu, rho, dt, tol, u0, v0) (ns/solver.h:37-140) on the plane:
// u0, v0 are aliased; step(fx, fy) alternates the backward-Euler half step
// and the Crank-Nicolson full step and leaves the new velocity in u0, v0.
// tol is the reference's absolute bound for both the CG and multigrid.
template<class G>
class solver {
	afm_ctx* c;
	dt::vector<G::size>& u_curr;
	dt::vector<G::size>& v_curr;
	afm_detail::Buffer du, dv, fx_d, fy_d;
	afm_ns* s = nullptr;
public:
	solver(double mu, double rho, double k, double tol, dt::vector<G::size>& u0, dt::vector<G::size>& v0,
	       int precision = 0)
		: c(afm_detail::context2<G>()), u_curr(u0), v_curr(v0), du(c, G::size * 8), dv(c, G::size * 8),
		  fx_d(c, G::size * 8), fy_d(c, G::size * 8)
	{
		du.put(u0.data(), G::size);
		dv.put(v0.data(), G::size);
		afm_detail::check(afm_ns_create(c, mu, rho, k, tol, 0, precision, du.d(), dv.d(), &s), "afm_ns_create");
		du.get(u0.data(), G::size);
		dv.get(v0.data(), G::size);
	}
	~solver() { afm_ns_destroy(s); }
	solver(const solver&) = delete;
	void step(const dt::vector<G::size>& fx, const dt::vector<G::size>& fy)
	{
		fx_d.put(fx.data(), G::size);
		fy_d.put(fy.data(), G::size);
		afm_detail::check(afm_ns_step(s, fx_d.d(), fy_d.d()), "afm_ns_step");
		du.get(u_curr.data(), G::size);
		dv.get(v_curr.data(), G::size);
	}
	const dt::vector<G::size>& get_u() { return u_curr; }
	const dt::vector<G::size>& get_v() { return v_curr; }
	dt::vector<G::size> get_p()
	{
		afm_detail::Buffer p(c, G::size * 8);
		afm_detail::check(afm_ns_get_p(s, p.d()), "afm_ns_get_p");
		dt::vector<G::size> out;
		p.get(out.data(), G::size);
		return out;
	}
};

// ns::solver<geometry<L,1,1,1>>(mu, rho, k, tol, u0, v0, w0) (ns/solver.h:
// 37-140): the reference's 3-D solver with walls in y.  u0, v0, w0 are
// aliased (v0 holds the nx*(ny-1)*nz interior y-faces) and projected at
// construction; step(ut, ub, wt, wb, fx, fy, fz) alternates the
// backward-Euler half step and the Crank-Nicolson full step, leaves the new
// velocity in u0, v0, w0, and adds projection::boundary's correction to the
// wall vectors in place, as the reference does.  tol is the reference's
// absolute bound for the multigrid and CG solves.
template<unsigned L, unsigned W, unsigned H, unsigned D>
class solver<dt::geometry<L, W, H, D>> {
	typedef dt::geometry<L, W, H, D> G;
	static constexpr unsigned nx = G::nx, ny = G::ny, nz = G::nz;
	static constexpr unsigned N = nx * ny * nz, Nv = nx * (ny - 1) * nz, Nxz = nx * nz;
	afm_ctx* c;
	dt::vector<N>& u_curr;
	dt::vector<Nv>& v_curr;
	dt::vector<N>& w_curr;
	afm_detail::Buffer du, dv, dw, fx_d, fy_d, fz_d, ut_d, ub_d, wt_d, wb_d;
	afm_ns3* s = nullptr;
public:
	solver(double mu, double rho, double k, double tol, dt::vector<N>& u0, dt::vector<Nv>& v0, dt::vector<N>& w0)
		: c((afm_detail::require_unit_cube<G>(), afm_detail::context3())), u_curr(u0), v_curr(v0), w_curr(w0),
		  du(c, N * 8), dv(c, Nv * 8),
		  dw(c, N * 8), fx_d(c, N * 8), fy_d(c, Nv * 8), fz_d(c, N * 8), ut_d(c, Nxz * 8), ub_d(c, Nxz * 8),
		  wt_d(c, Nxz * 8), wb_d(c, Nxz * 8)
	{
		du.put(u0.data(), N);
		dv.put(v0.data(), Nv);
		dw.put(w0.data(), N);
		afm_detail::check(afm_ns3_create(c, L, mu, rho, k, tol, du.d(), dv.d(), dw.d(), &s), "afm_ns3_create");
		du.get(u0.data(), N);
		dv.get(v0.data(), Nv);
		dw.get(w0.data(), N);
	}
	~solver() { afm_ns3_destroy(s); }
	solver(const solver&) = delete;
	solver& operator=(const solver&) = delete;
	void step(dt::vector<Nxz>& ut, dt::vector<Nxz>& ub, dt::vector<Nxz>& wt, dt::vector<Nxz>& wb,
	          const dt::vector<N>& fx, const dt::vector<Nv>& fy, const dt::vector<N>& fz)
	{
		ut_d.put(ut.data(), Nxz);
		ub_d.put(ub.data(), Nxz);
		wt_d.put(wt.data(), Nxz);
		wb_d.put(wb.data(), Nxz);
		fx_d.put(fx.data(), N);
		fy_d.put(fy.data(), Nv);
		fz_d.put(fz.data(), N);
		afm_detail::check(afm_ns3_step(s, ut_d.d(), ub_d.d(), wt_d.d(), wb_d.d(), fx_d.d(), fy_d.d(), fz_d.d()),
		                  "afm_ns3_step");
		du.get(u_curr.data(), N);
		dv.get(v_curr.data(), Nv);
		dw.get(w_curr.data(), N);
		ut_d.get(ut.data(), Nxz);
		ub_d.get(ub.data(), Nxz);
		wt_d.get(wt.data(), Nxz);
		wb_d.get(wb.data(), Nxz);
	}
	const dt::vector<N>& get_u() { return u_curr; }
	const dt::vector<Nv>& get_v() { return v_curr; }
	const dt::vector<N>& get_w() { return w_curr; }
	dt::vector<N> get_p()
	{
		afm_detail::Buffer p(c, N * 8);
		afm_detail::check(afm_ns3_get_p(s, p.d()), "afm_ns3_get_p");
		dt::vector<N> out;
		p.get(out.data(), N);
		return out;
	}
};

}  // namespace ns
