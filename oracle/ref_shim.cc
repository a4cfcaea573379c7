// ref_shim.cc -- instantiates the reference's header-only templates for a few
// fixed geometries and exposes them through a C ABI, so that the oracle
// (afm_oracle.cpp) can be pinned against the reference's own code.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile from the reference
// headers where they lie (-I <reference>/src) into oracle/_ref/; this file
// contains no reference code, only calls into it.  The reference sizes are
// compile-time template parameters, so every entry point takes a geometry id:
//   3-D (dt::geometry<L,W,H,D>):  0 -> <4,1,1,1> (16^3),  1 -> <5,1,1,1> (32^3),
//                                  2 -> <1,8,8,8> (16^3, coarsest 8^3 by CG),
//                                  3 -> <1,1,1,1> (2^3; restriction only)
//   2-D planes (nx x 1 x nz, the reference's operators with ny == 1):
//                                 10 -> 16x1x16 (n 16), 11 -> 32x1x32 (n 32),
//                                 12 -> 64x1x64 (n 64)
#include <vector>
#include <cstring>
#include "dt/geometry.h"
#include "dt/vector.h"
#include "dt/matrix.h"
#include "alg/cg.h"
#include "alg/smooth.h"
#include "alg/div.h"
#include "alg/grad.h"
#include "mg/laplacian.h"
#include "mg/solver.h"
#include "ns/helmholtz.h"
#include "ns/solver.h"
#include "ib/grid.h"
#include "ib/eulerian_grid.h"
#include "ib/mac_grid.h"
#include "ib/lagrangian_grid.h"

namespace {

typedef dt::geometry<4u, 1u, 1u, 1u> G16;
typedef dt::geometry<5u, 1u, 1u, 1u> G32;
typedef dt::geometry<1u, 8u, 8u, 8u> G8;
typedef dt::geometry<1u, 1u, 1u, 1u> G1;

template<unsigned NX, unsigned NZ, unsigned N>
struct Plane {
	static constexpr unsigned n = N, nx = NX, ny = 1, nz = NZ;
	static constexpr unsigned width = NX / N, height = 1, depth = NZ / N;
};
typedef Plane<16, 16, 16> P16;
typedef Plane<32, 32, 32> P32;
typedef Plane<64, 64, 64> P64;

constexpr unsigned M = 64;   // marker count of the 3-D transfer entry points

// ---- delta: a minimal concrete grid exposing the protected delta_function.
template<class G>
struct DeltaProbe : ib::grid<G, 1u> {
	using typename ib::grid<G, 1u>::point_type;
	point_type coordinates(unsigned) { return point_type(); }
	std::vector<point_type> search_nearby(point_type) { return {}; }
	double volume(unsigned) const { return 1.0; }
	double d(double r) { return this->delta_function(r); }
};

template<unsigned L>
double delta_l(double r) { DeltaProbe<dt::geometry<L, 1u, 1u, 1u>> p; return p.d(r); }

// ---- markers with a caller-supplied volume (ib/lagrangian_grid.h is abstract).
template<class G>
struct Markers : ib::lagrangian_grid<G, M> {
	const double* ds;
	Markers(std::vector<dt::point> p, const double* ds_) : ib::lagrangian_grid<G, M>(p), ds(ds_) {}
	double volume(unsigned i) const { return ds[i]; }
};

// one marker of a given volume (test/ib_test.cc:33-43 defines such a grid)
template<class G>
struct OneMarker : ib::lagrangian_grid<G, 1u> {
	double vol;
	OneMarker(double x, double y, double z, double v)
		: ib::lagrangian_grid<G, 1u>({dt::point(x, y, z, 0u)}), vol(v) {}
	double volume(unsigned) const { return vol; }
};

template<class G, ib::mac::component C>
void interp(const double* field, const double* X, double* U)
{
	ib::mac_grid<G, C> e;
	for (unsigned i = 0; i < e.size(); ++i) e[i] = field[i];
	std::vector<dt::point> pts;
	for (unsigned k = 0; k < M; ++k) pts.push_back(dt::point(X[3 * k], X[3 * k + 1], X[3 * k + 2], k));
	Markers<G> l(pts, nullptr);
	l.interpolate(e);
	for (unsigned k = 0; k < M; ++k) U[k] = l[k];
}

// the node grid ib::eulerian_grid<G> (comp 3) in place of a MAC component
template<class G>
void interp_nodes(const double* field, const double* X, double* U)
{
	ib::eulerian_grid<G> e;
	for (unsigned i = 0; i < e.size(); ++i) e[i] = field[i];
	std::vector<dt::point> pts;
	for (unsigned k = 0; k < M; ++k) pts.push_back(dt::point(X[3 * k], X[3 * k + 1], X[3 * k + 2], k));
	Markers<G> l(pts, nullptr);
	l.interpolate(e);
	for (unsigned k = 0; k < M; ++k) U[k] = l[k];
}

template<class G>
void spread_nodes(const double* F, const double* ds, const double* X, double* f)
{
	ib::eulerian_grid<G> e;
	std::vector<dt::point> pts;
	for (unsigned k = 0; k < M; ++k) pts.push_back(dt::point(X[3 * k], X[3 * k + 1], X[3 * k + 2], k));
	Markers<G> l(pts, ds);
	for (unsigned k = 0; k < M; ++k) l[k] = F[k];
	e.interpolate(l);
	for (unsigned i = 0; i < e.size(); ++i) f[i] = e[i];
}

template<class G, ib::mac::component C>
void spread(const double* F, const double* ds, const double* X, double* f)
{
	ib::mac_grid<G, C> e;
	std::vector<dt::point> pts;
	for (unsigned k = 0; k < M; ++k) pts.push_back(dt::point(X[3 * k], X[3 * k + 1], X[3 * k + 2], k));
	Markers<G> l(pts, ds);
	for (unsigned k = 0; k < M; ++k) l[k] = F[k];
	e.interpolate(l);
	for (unsigned i = 0; i < e.size(); ++i) f[i] = e[i];
}

// ---- multigrid internals exposed through a subclass (as test/mg_test2.cc and
// test/restr_test.cc do).
template<class G>
struct MGX : mg::solver<G> {
	using mg::solver<G>::L;
	using mg::solver<G>::restriction;
	using mg::solver<G>::interpolation;
	MGX(double tol, unsigned a, unsigned b) : mg::solver<G>(tol, a, b) {}
};

template<class Mat>
struct CsrView : Mat {
	using Mat::values; using Mat::column; using Mat::row_start;
	template<class... A> CsrView(A... a) : Mat(a...) {}
};

template<class G> constexpr unsigned SZ = G::nx * G::ny * G::nz;

template<class G>
void lap_apply(const double* x, double* y)
{
	mg::laplacian<G> L;
	dt::vector<SZ<G>> v;
	for (unsigned i = 0; i < SZ<G>; ++i) v[i] = x[i];
	auto r = L * v;
	for (unsigned i = 0; i < SZ<G>; ++i) y[i] = r[i];
}

template<class G>
void lap_smooth(double* x, const double* b, int sweeps)
{
	mg::laplacian<G> L;
	dt::vector<SZ<G>> xv, bv;
	for (unsigned i = 0; i < SZ<G>; ++i) { xv[i] = x[i]; bv[i] = b[i]; }
	for (int s = 0; s < sweeps; ++s) alg::smooth(xv, L, bv);
	for (unsigned i = 0; i < SZ<G>; ++i) x[i] = xv[i];
}

template<class G, class Mat>
unsigned run_cg(Mat& A, double* x, const double* b, double tol)
{
	dt::vector<SZ<G>> xv, bv;
	for (unsigned i = 0; i < SZ<G>; ++i) { xv[i] = x[i]; bv[i] = b[i]; }
	unsigned it = alg::conjugate_gradient(xv, A, bv, tol);
	for (unsigned i = 0; i < SZ<G>; ++i) x[i] = xv[i];
	return it;
}

template<class G>
long csr_export(int kind, double scale, long* rs, long* col, double* val)
{
	if (kind == 0) return -1;   // mg::laplacian keeps its arrays private: pinned through SpMV
	if (kind == 3) {
		if (G::ny < 2) return -1;
		CsrView<ns::helmholtz_dirichlet<G>> A(scale);
		constexpr unsigned rows = G::nx * (G::ny - 1) * G::nz;
		if (rs) {
			for (unsigned i = 0; i <= rows; ++i) rs[i] = A.row_start[i];
			for (size_t k = 0; k < A.values.size(); ++k) { col[k] = A.column[k]; val[k] = A.values[k]; }
		}
		return (long) A.values.size();
	}
	CsrView<ns::helmholtz_periodic<G>> A(scale);
	if (rs) {
		for (unsigned i = 0; i <= SZ<G>; ++i) rs[i] = A.row_start[i];
		for (size_t k = 0; k < A.values.size(); ++k) { col[k] = A.column[k]; val[k] = A.values[k]; }
	}
	return (long) A.values.size();
}

template<class G>
void div_grad(const double* u, const double* v, const double* w, double* d,
              const double* p, double* px, double* py, double* pz)
{
	constexpr unsigned nx = G::nx, ny = G::ny, nz = G::nz, n = G::n;
	dt::vector<nx * ny * nz> uu, ww, dd, pp, gx, gz;
	dt::vector<nx * (ny - 1) * nz> vv, gy;
	dt::vector<nx * nz> top, bottom;
	for (unsigned i = 0; i < nx * ny * nz; ++i) { uu[i] = u[i]; ww[i] = w[i]; pp[i] = p[i]; }
	for (unsigned i = 0; i < nx * (ny - 1) * nz; ++i) vv[i] = v[i];
	alg::div<nx, ny, nz, n>(dd, uu, vv, ww, top, bottom);
	alg::grad<nx, ny, nz, n>(gx, gy, gz, pp);
	for (unsigned i = 0; i < nx * ny * nz; ++i) { d[i] = dd[i]; px[i] = gx[i]; pz[i] = gz[i]; }
	for (unsigned i = 0; i < nx * (ny - 1) * nz; ++i) py[i] = gy[i];
}

// ns::solver<G> (ns/solver.h:37-140): construct over u, v, w (projects
// them), run `steps` step() calls with fixed forces; wall vectors in/out.
template<class G>
void ns_run(double mu, double rho, double dt, double tol, double* u, double* v, double* w, double* ut, double* ub,
            double* wt, double* wb, const double* fx, const double* fy, const double* fz, int steps, double* p)
{
	constexpr unsigned nx = G::nx, ny = G::ny, nz = G::nz, N = nx * ny * nz, Nv = nx * (ny - 1) * nz, Nxz = nx * nz;
	dt::vector<N> uu, ww, ffx, ffz;
	dt::vector<Nv> vv, ffy;
	dt::vector<Nxz> t1, b1, t2, b2;
	for (unsigned i = 0; i < N; ++i) { uu[i] = u[i]; ww[i] = w[i]; ffx[i] = fx[i]; ffz[i] = fz[i]; }
	for (unsigned i = 0; i < Nv; ++i) { vv[i] = v[i]; ffy[i] = fy[i]; }
	for (unsigned i = 0; i < Nxz; ++i) { t1[i] = ut[i]; b1[i] = ub[i]; t2[i] = wt[i]; b2[i] = wb[i]; }
	ns::solver<G> s(mu, rho, dt, tol, uu, vv, ww);
	for (int k = 0; k < steps; ++k) s.step(t1, b1, t2, b2, ffx, ffy, ffz);
	for (unsigned i = 0; i < N; ++i) { u[i] = uu[i]; w[i] = ww[i]; }
	for (unsigned i = 0; i < Nv; ++i) v[i] = vv[i];
	for (unsigned i = 0; i < Nxz; ++i) { ut[i] = t1[i]; ub[i] = b1[i]; wt[i] = t2[i]; wb[i] = b2[i]; }
	if (p) { auto pp = s.get_p(); for (unsigned i = 0; i < N; ++i) p[i] = pp[i]; }
}

// ---- ns::solver<G>::advection (ns/solver.h:142-236) and its outputs u_adv,
// v_adv, w_adv are private.  An explicit instantiation may name private
// members (access is not checked there), and a friend it defines hands out
// the member pointers; the reference's code is then called as it stands.
template<class Tag, typename Tag::type Ptr>
struct Expose { friend typename Tag::type member_of(Tag) { return Ptr; } };

constexpr unsigned N16 = SZ<G16>, Nv16 = G16::nx * (G16::ny - 1) * G16::nz;
struct AdvFn { typedef void (ns::solver<G16>::*type)(dt::vector<N16>&, dt::vector<Nv16>&, dt::vector<N16>&);
               friend type member_of(AdvFn); };
struct AdvU { typedef dt::vector<N16> ns::solver<G16>::*type; friend type member_of(AdvU); };
struct AdvV { typedef dt::vector<Nv16> ns::solver<G16>::*type; friend type member_of(AdvV); };
struct AdvW { typedef dt::vector<N16> ns::solver<G16>::*type; friend type member_of(AdvW); };
template struct Expose<AdvFn, &ns::solver<G16>::advection>;
template struct Expose<AdvU, &ns::solver<G16>::u_adv>;
template struct Expose<AdvV, &ns::solver<G16>::v_adv>;
template struct Expose<AdvW, &ns::solver<G16>::w_adv>;

}  // namespace

extern "C" {

// ns::solver<geometry<4,1,1,1>>::advection on caller data (u, w: 16^3;
// v: 16*15*16 interior y-faces); the solver itself is built over zero fields.
int ref_advect(const double* u, const double* v, const double* w, double* ua, double* va, double* wa)
{
	dt::vector<N16> u0, w0, uu, ww;
	dt::vector<Nv16> v0, vv;
	ns::solver<G16> s(1e-2, 1.0, 1e-3, 1e-12, u0, v0, w0);
	for (unsigned i = 0; i < N16; ++i) { uu[i] = u[i]; ww[i] = w[i]; }
	for (unsigned i = 0; i < Nv16; ++i) vv[i] = v[i];
	(s.*member_of(AdvFn()))(uu, vv, ww);
	const dt::vector<N16>& a = s.*member_of(AdvU());
	const dt::vector<Nv16>& b = s.*member_of(AdvV());
	const dt::vector<N16>& c = s.*member_of(AdvW());
	for (unsigned i = 0; i < N16; ++i) { ua[i] = a[i]; wa[i] = c[i]; }
	for (unsigned i = 0; i < Nv16; ++i) va[i] = b[i];
	return 0;
}

// ib/grid.h:51-55 for n = 2^level, level 1..12.
double ref_delta(double r, unsigned level)
{
	switch (level) {
	case 1: return delta_l<1>(r);   case 2: return delta_l<2>(r);   case 3: return delta_l<3>(r);
	case 4: return delta_l<4>(r);   case 5: return delta_l<5>(r);   case 6: return delta_l<6>(r);
	case 7: return delta_l<7>(r);   case 8: return delta_l<8>(r);   case 9: return delta_l<9>(r);
	case 10: return delta_l<10>(r); case 11: return delta_l<11>(r); case 12: return delta_l<12>(r);
	}
	return -1.0;
}

unsigned ref_markers(void) { return M; }

// ib/grid.h:68-84, marker targets; comp 0/1/2 = mac::x/y/z.  X is M x 3.
int ref_interp(int geo, int comp, const double* field, const double* X, double* U)
{
	using namespace ib::mac;
	if (geo == 0) {
		if (comp == 0) interp<G16, x>(field, X, U); else if (comp == 1) interp<G16, y>(field, X, U);
		else if (comp == 2) interp<G16, z>(field, X, U); else interp_nodes<G16>(field, X, U);
	} else if (geo == 1) {
		if (comp == 0) interp<G32, x>(field, X, U); else if (comp == 1) interp<G32, y>(field, X, U);
		else if (comp == 2) interp<G32, z>(field, X, U); else interp_nodes<G32>(field, X, U);
	} else return -1;
	return 0;
}

// ib/grid.h:68-84, grid targets (spreading through the kd-tree search).
int ref_spread(int geo, int comp, const double* F, const double* ds, const double* X, double* f)
{
	using namespace ib::mac;
	if (geo == 0) {
		if (comp == 0) spread<G16, x>(F, ds, X, f); else if (comp == 1) spread<G16, y>(F, ds, X, f);
		else if (comp == 2) spread<G16, z>(F, ds, X, f); else spread_nodes<G16>(F, ds, X, f);
	} else if (geo == 1) {
		if (comp == 0) spread<G32, x>(F, ds, X, f); else if (comp == 1) spread<G32, y>(F, ds, X, f);
		else if (comp == 2) spread<G32, z>(F, ds, X, f); else spread_nodes<G32>(F, ds, X, f);
	} else return -1;
	return 0;
}

// mg/laplacian.h * x (dt/matrix.h:81-94).
int ref_lap_apply(int geo, const double* x, double* y)
{
	switch (geo) {
	case 0: lap_apply<G16>(x, y); return 0;  case 1: lap_apply<G32>(x, y); return 0;
	case 2: lap_apply<G8>(x, y); return 0;   case 10: lap_apply<P16>(x, y); return 0;
	case 11: lap_apply<P32>(x, y); return 0; case 12: lap_apply<P64>(x, y); return 0;
	}
	return -1;
}

// alg/smooth.h sweeps on the laplacian.
int ref_smooth(int geo, double* x, const double* b, int sweeps)
{
	switch (geo) {
	case 0: lap_smooth<G16>(x, b, sweeps); return 0;  case 1: lap_smooth<G32>(x, b, sweeps); return 0;
	case 10: lap_smooth<P16>(x, b, sweeps); return 0; case 11: lap_smooth<P32>(x, b, sweeps); return 0;
	case 12: lap_smooth<P64>(x, b, sweeps); return 0;
	}
	return -1;
}

// alg/cg.h on the laplacian (kind 0) or helmholtz_periodic(scale) (kind 1).
long ref_cg(int geo, int kind, double scale, double* x, const double* b, double tol)
{
	if (kind == 0) {
		switch (geo) {
		case 0: { mg::laplacian<G16> A; return run_cg<G16>(A, x, b, tol); }
		case 10: { mg::laplacian<P16> A; return run_cg<P16>(A, x, b, tol); }
		case 11: { mg::laplacian<P32> A; return run_cg<P32>(A, x, b, tol); }
		}
	} else {
		switch (geo) {
		case 0: { ns::helmholtz_periodic<G16> A(scale); return run_cg<G16>(A, x, b, tol); }
		case 1: { ns::helmholtz_periodic<G32> A(scale); return run_cg<G32>(A, x, b, tol); }
		}
	}
	return -1;
}

// CSR arrays of ns::helmholtz_periodic (kind 1) and ns::helmholtz_dirichlet
// (kind 3, nx*(ny-1)*nz rows); kind 0 is not exported.
// With rs == NULL only the entry count is returned.
long ref_csr(int geo, int kind, double scale, long* rs, long* col, double* val)
{
	switch (geo) {
	case 0: return csr_export<G16>(kind, scale, rs, col, val);
	case 1: return csr_export<G32>(kind, scale, rs, col, val);
	case 2: return csr_export<G8>(kind, scale, rs, col, val);
	case 10: return csr_export<P16>(kind, scale, rs, col, val);
	case 11: return csr_export<P32>(kind, scale, rs, col, val);
	}
	return -1;
}

// mg/solver.h:153-181 (restriction) and :71-151 (interpolation, overwriting
// only the rows it visits; `fine` is read-modify-write).
int ref_restrict(int geo, const double* fine, double* coarse)
{
	if (geo == 0) {
		dt::vector<SZ<G16>> f; dt::vector<SZ<G16> / 8> c;
		for (unsigned i = 0; i < SZ<G16>; ++i) f[i] = fine[i];
		MGX<G16>::restriction(c, f);
		for (unsigned i = 0; i < SZ<G16> / 8; ++i) coarse[i] = c[i];
		return 0;
	}
	if (geo == 1) {
		dt::vector<SZ<G32>> f; dt::vector<SZ<G32> / 8> c;
		for (unsigned i = 0; i < SZ<G32>; ++i) f[i] = fine[i];
		MGX<G32>::restriction(c, f);
		for (unsigned i = 0; i < SZ<G32> / 8; ++i) coarse[i] = c[i];
		return 0;
	}
	return -1;
}

int ref_prolong(int geo, const double* coarse, double* fine)
{
	if (geo == 0) {
		dt::vector<SZ<G16>> f; dt::vector<SZ<G16> / 8> c;
		for (unsigned i = 0; i < SZ<G16>; ++i) f[i] = fine[i];
		for (unsigned i = 0; i < SZ<G16> / 8; ++i) c[i] = coarse[i];
		MGX<G16>::interpolation(f, c);
		for (unsigned i = 0; i < SZ<G16>; ++i) fine[i] = f[i];
		return 0;
	}
	if (geo == 1) {
		dt::vector<SZ<G32>> f; dt::vector<SZ<G32> / 8> c;
		for (unsigned i = 0; i < SZ<G32>; ++i) f[i] = fine[i];
		for (unsigned i = 0; i < SZ<G32> / 8; ++i) c[i] = coarse[i];
		MGX<G32>::interpolation(f, c);
		for (unsigned i = 0; i < SZ<G32>; ++i) fine[i] = f[i];
		return 0;
	}
	return -1;
}

// mg::solver<G>(tol, nu1, nu2).solve(rhs) -> cycles; rhs overwritten.
long ref_mg_solve(int geo, double tol, unsigned nu1, unsigned nu2, double* rhs)
{
	if (geo == 0) {
		mg::solver<G16> s(tol, nu1, nu2);
		dt::vector<SZ<G16>> b;
		for (unsigned i = 0; i < SZ<G16>; ++i) b[i] = rhs[i];
		unsigned c = s.solve(b);
		for (unsigned i = 0; i < SZ<G16>; ++i) rhs[i] = b[i];
		return c;
	}
	if (geo == 2) {
		mg::solver<G8> s(tol, nu1, nu2);
		dt::vector<SZ<G8>> b;
		for (unsigned i = 0; i < SZ<G8>; ++i) b[i] = rhs[i];
		unsigned c = s.solve(b);
		for (unsigned i = 0; i < SZ<G8>; ++i) rhs[i] = b[i];
		return c;
	}
	return -1;
}

// alg/div.h and alg/grad.h.  v, py have nx*(ny-1)*nz entries (none for planes).
int ref_div_grad(int geo, const double* u, const double* v, const double* w, double* d,
                 const double* p, double* px, double* py, double* pz)
{
	switch (geo) {
	case 0: div_grad<G16>(u, v, w, d, p, px, py, pz); return 0;
	case 10: div_grad<P16>(u, v, w, d, p, px, py, pz); return 0;
	case 11: div_grad<P32>(u, v, w, d, p, px, py, pz); return 0;
	}
	return -1;
}

// ns::solver steps on the 3-D geometries (0: 16^3, 1: 32^3).
int ref_ns3_run(int geo, double mu, double rho, double dt, double tol, double* u, double* v, double* w, double* ut,
                double* ub, double* wt, double* wb, const double* fx, const double* fy, const double* fz, int steps,
                double* p)
{
	switch (geo) {
	case 0: ns_run<G16>(mu, rho, dt, tol, u, v, w, ut, ub, wt, wb, fx, fy, fz, steps, p); return 0;
	case 1: ns_run<G32>(mu, rho, dt, tol, u, v, w, ut, ub, wt, wb, fx, fy, fz, steps, p); return 0;
	}
	return -1;
}

// The computation of test/ib_test.cc:73-93 on geometry<5,1,1,1>: one marker
// of volume 1/n^3 at (x, y, z) holding 1 is spread to mac_grid<G, y>; the
// field is returned (nx (ny-1) nz entries), its non-zero entries are set to
// 1 and interpolated back to the marker, whose value is returned in *out.
int ref_ib_test(double x, double y, double z, double* field, double* out)
{
	constexpr unsigned n = G32::n;
	ib::mac_grid<G32, ib::mac::y> egrid;
	OneMarker<G32> lgrid(x, y, z, 1. / (n * n * n));
	for (auto& e : lgrid) e = 1.0;
	egrid.interpolate(lgrid);
	for (unsigned i = 0; i < egrid.size(); ++i) field[i] = egrid[i];
	for (auto& e : egrid) if (e != 0.0) e = 1;
	for (auto& e : lgrid) e = 0.0;
	lgrid.interpolate(egrid);
	*out = lgrid[0];
	return (int) egrid.size();
}

// mg::solver<geometry<1,1,1,1>>::restriction of 8 values into one (the
// computation of test/restr_test.cc:38-50; the output starts at zero).
int ref_restrict_g1(const double* fine, double* coarse)
{
	dt::vector<8> f; dt::vector<1> c;
	for (unsigned i = 0; i < 8; ++i) f[i] = fine[i];
	MGX<G1>::restriction(c, f);
	coarse[0] = c[0];
	return 0;
}

}  // extern "C"
