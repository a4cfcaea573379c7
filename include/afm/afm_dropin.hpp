// afm_dropin.hpp -- C++ drop-in for the reference's dt / mg / ns / ib class
// interfaces, on the doubly periodic plane, forwarding to libafm's C ABI.
//
// Same class names, constructor arguments and member functions as the
// reference headers (cited per class); the data stays value-semantic on the
// host (dt::vector), and every call moves it to a device mirror, runs the
// gfx950 kernels and copies results back, so programs written against the
// reference keep their semantics (including ns::solver aliasing the caller's
// velocity vectors, ns/solver.h:54-56,68).
//
// Geometry: dt::plane<Level> is the 2-D analogue of
// dt::geometry<Level, W, H, D> (dt/geometry.h:29-44) for the unit square:
// n = nx = ny = 2^Level.  Points are (x, y, index).
//
// One context per geometry is created on first use (device 0, or the device
// named by AFM_DEVICE) and lives until the process ends, as the reference's
// static per-level laplacians do (mg/solver.h:68-69).
#pragma once
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <memory>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../afm.h"

namespace afm_detail {

inline void check(int rc, const char* what)
{
	if (rc != 0) throw std::runtime_error(std::string(what) + ": " + afm_last_error());
}

struct Buffer {
	afm_ctx* c = nullptr;
	void* p = nullptr;
	Buffer(afm_ctx* ctx, size_t bytes) : c(ctx) { check(afm_alloc(c, bytes, &p), "afm_alloc"); }
	~Buffer() { if (p) afm_free(c, p); }
	Buffer(const Buffer&) = delete;
	Buffer& operator=(const Buffer&) = delete;
	double* d() const { return static_cast<double*>(p); }
	void put(const double* h, size_t count) { check(afm_copy_to_device(c, p, h, count * 8), "copy_to_device"); }
	void get(double* h, size_t count) const { check(afm_copy_to_host(c, h, p, count * 8), "copy_to_host"); }
};

// the context the 3-D geometry's calls run on (their arrays are their own;
// only the device and stream of the context are used)
inline afm_ctx* context3()
{
	static std::unique_ptr<afm_ctx, int (*)(afm_ctx*)> holder(nullptr, afm_ctx_destroy);
	if (!holder) {
		afm_geom g{2, 8, 1, 1};
		const char* dev = std::getenv("AFM_DEVICE");
		afm_ctx* c = nullptr;
		check(afm_ctx_create(&g, dev ? std::atoi(dev) : 0, &c), "afm_ctx_create");
		holder.reset(c);
	}
	return holder.get();
}

template<class G>
afm_ctx* context()
{
	static std::unique_ptr<afm_ctx, int (*)(afm_ctx*)> holder(nullptr, afm_ctx_destroy);
	if (!holder) {
		afm_geom g{2, G::n, G::width, G::depth};
		const char* dev = std::getenv("AFM_DEVICE");
		afm_ctx* c = nullptr;
		check(afm_ctx_create(&g, dev ? std::atoi(dev) : 0, &c), "afm_ctx_create");
		holder.reset(c);
	}
	return holder.get();
}

}  // namespace afm_detail

namespace dt {

template<unsigned Level, unsigned Width = 1, unsigned Depth = 1>
struct plane {
	static constexpr unsigned level = Level, width = Width, depth = Depth;
	static constexpr unsigned n = 1u << Level;
	static constexpr unsigned nx = Width * n, ny = Depth * n, size = nx * ny;
};

// dt/geometry.h:29-60: the reference's 3-D geometry, n = 2^Level cells per
// unit length, walls in y.  mg::solver and ns::solver accept the cube
// geometry<Level, 1, 1, 1> (the reference's own tests' geometry).
template<unsigned Level, unsigned Width, unsigned Height, unsigned Depth>
struct geometry {
	static constexpr unsigned factor = 2;
	static constexpr unsigned n = 1u << Level;
	static constexpr unsigned nx = Width * n, ny = Height * n, nz = Depth * n, size = nx * ny * nz;
	static constexpr unsigned width = Width, height = Height, depth = Depth, level = Level;
};

// dt/point.h:29-48, in the plane.
class point {
	double v[2];
	std::size_t idx;
public:
	typedef double value_type;
	point(double x, double y, std::size_t index) : v{x, y}, idx(index) {}
	point() : v{0.0, 0.0}, idx(0) {}
	double& operator[](std::size_t i) { return v[i]; }
	const double& operator[](std::size_t i) const { return v[i]; }
	std::size_t get_index() const { return idx; }
};

// dt/vector.h:33-165: a value-semantic vector of doubles.
template<unsigned Size>
class vector {
protected:
	std::vector<double> values;
public:
	typedef double value_type;
	typedef std::vector<double>::iterator iterator;
	typedef std::vector<double>::const_iterator const_iterator;
	vector(double value = 0.0) : values(Size, value) {}
	vector(std::initializer_list<double> v) : values(v)
	{
		if (v.size() != Size) throw std::length_error("dt::vector: initializer size");
	}
	virtual ~vector() {}
	virtual double& operator[](unsigned i) { return values[i]; }
	virtual const double& operator[](unsigned i) const { return values[i]; }
	unsigned size() const { return Size; }
	iterator begin() { return values.begin(); }
	iterator end() { return values.end(); }
	const_iterator cbegin() const { return values.cbegin(); }
	const_iterator cend() const { return values.cend(); }
	double* data() { return values.data(); }
	const double* data() const { return values.data(); }
	double operator*(const vector& o) const
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
template<unsigned S> std::ostream& operator<<(std::ostream& out, const vector<S>& v)
{
	out << "[";
	for (unsigned i = 0; i < S; ++i) out << (i ? ", " : "") << v[i];
	return out << "]'";
}

}  // namespace dt

namespace mg {

// mg::solver<G>(tol, nu1, nu2).solve(rhs) -> V-cycles (mg/solver.h:60-62,
// 218-237): rhs is replaced by the solution; tol is the absolute RMS bound.
template<class G>
class solver {
	double tol;
	unsigned nu1, nu2;
public:
	int precision = 0;          // as afm_mg_solve's (1: FP32 V-cycle preconditioning FP64 CG)
	solver(double tol_, unsigned nu1_, unsigned nu2_) : tol(tol_), nu1(nu1_), nu2(nu2_) {}
	unsigned solve(dt::vector<G::size>& rhs)
	{
		afm_ctx* c = afm_detail::context<G>();
		afm_detail::Buffer b(c, G::size * 8);
		b.put(rhs.data(), G::size);
		int cyc = 0;
		afm_detail::check(afm_mg_solve(c, b.d(), tol, 0, (int) nu1, (int) nu2, precision, 100000, &cyc, nullptr),
		                  "afm_mg_solve");
		b.get(rhs.data(), G::size);
		return (unsigned) cyc;
	}
};

// mg::solver<geometry<L,1,1,1>>: the 3-D solve with walls in y (red-black
// smoothing, the reference's absolute RMS tolerance).
template<unsigned L, unsigned W, unsigned H, unsigned D>
class solver<dt::geometry<L, W, H, D>> {
	static_assert(W == 1 && H == 1 && D == 1 && L >= 1, "the 3-D solve takes geometry<L, 1, 1, 1>, L >= 1");
	typedef dt::geometry<L, W, H, D> G;
	double tol;
	unsigned nu1, nu2;
public:
	solver(double tol_, unsigned nu1_, unsigned nu2_) : tol(tol_), nu1(nu1_), nu2(nu2_) {}
	unsigned solve(dt::vector<G::size>& rhs)
	{
		afm_ctx* c = afm_detail::context3();
		afm_detail::Buffer b(c, G::size * 8);
		b.put(rhs.data(), G::size);
		int cyc = 0;
		afm_detail::check(afm_mg3_solve(c, L, b.d(), tol, 0, (int) nu1, (int) nu2, 100000, &cyc, nullptr),
		                  "afm_mg3_solve");
		b.get(rhs.data(), G::size);
		return (unsigned) cyc;
	}
};

}  // namespace mg

namespace ns {

// ns::projection<G>(tol).project(u, v) (ns/projection.h:44-73), V(2,2).
template<class G>
class projection {
	double tol;
	dt::vector<G::size> phidt;
public:
	projection(double tol_) : tol(tol_) {}
	void project(dt::vector<G::size>& u, dt::vector<G::size>& v)
	{
		afm_ctx* c = afm_detail::context<G>();
		afm_detail::Buffer du(c, G::size * 8), dv(c, G::size * 8), p(c, G::size * 8);
		du.put(u.data(), G::size);
		dv.put(v.data(), G::size);
		afm_detail::check(afm_div(c, du.d(), dv.d(), p.d()), "afm_div");
		int cyc = 0;
		afm_detail::check(afm_mg_solve(c, p.d(), tol, 0, 2, 2, 0, 100000, &cyc, nullptr), "afm_mg_solve");
		afm_detail::check(afm_grad_sub(c, p.d(), du.d(), dv.d()), "afm_grad_sub");
		du.get(u.data(), G::size);
		dv.get(v.data(), G::size);
		p.get(phidt.data(), G::size);
	}
	dt::vector<G::size>& get_phidt() { return phidt; }
};

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
		: c(afm_detail::context<G>()), u_curr(u0), v_curr(v0), du(c, G::size * 8), dv(c, G::size * 8),
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
	static_assert(W == 1 && H == 1 && D == 1 && L >= 1, "the 3-D solver takes geometry<L, 1, 1, 1>, L >= 1");
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
		: c(afm_detail::context3()), u_curr(u0), v_curr(v0), w_curr(w0), du(c, N * 8), dv(c, Nv * 8),
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

namespace ib {

namespace mac { enum component { x, y }; }

template<class G, unsigned Size> class lagrangian_grid;
template<class G, mac::component C> class mac_grid;

// ib::lagrangian_grid<G, Size> (ib/lagrangian_grid.h:33-67): markers with a
// user-supplied volume (arclength weight); abstract, as in the reference.
template<class G, unsigned Size>
class lagrangian_grid : public dt::vector<Size> {
	std::vector<dt::point> positions;
public:
	lagrangian_grid(std::vector<dt::point> p, double v = 0.0) : dt::vector<Size>(v), positions(std::move(p))
	{
		if (positions.size() != Size) throw std::length_error("lagrangian_grid: positions");
	}
	virtual double volume(unsigned) const = 0;
	virtual dt::point coordinates(unsigned i) { return positions[i]; }
	// interpolation (ib/grid.h:68-84 with a MAC source): this[k] = J u
	template<mac::component C> void interpolate(mac_grid<G, C>& source);
	std::vector<double> packed_positions() const
	{
		std::vector<double> X(2 * Size);
		for (unsigned k = 0; k < Size; ++k) { X[2 * k] = positions[k][0]; X[2 * k + 1] = positions[k][1]; }
		return X;
	}
};

// ib::mac_grid<G, C> (ib/mac_grid.h:30-136): a MAC component on the plane,
// u on x-faces (C = x), v on y-faces (C = y).
template<class G, mac::component C>
class mac_grid : public dt::vector<G::size> {
public:
	mac_grid(double v = 0.0) : dt::vector<G::size>(v) {}
	// spreading (ib/grid.h:68-84 with a marker source): this = S (F ds)
	template<unsigned R> void interpolate(lagrangian_grid<G, R>& source)
	{
		afm_ctx* c = afm_detail::context<G>();
		std::vector<double> X = source.packed_positions(), ds(R);
		for (unsigned k = 0; k < R; ++k) ds[k] = source.volume(k);
		afm_detail::Buffer dX(c, 2 * R * 8), dF(c, R * 8), dds(c, R * 8), f(c, G::size * 8);
		dX.put(X.data(), 2 * R);
		dF.put(source.data(), R);
		dds.put(ds.data(), R);
		afm_detail::check(afm_spread(c, C == mac::x ? 0 : 1, dF.d(), dds.d(), dX.d(), R, f.d()), "afm_spread");
		f.get(this->data(), G::size);
	}
};

template<class G, unsigned Size>
template<mac::component C>
void lagrangian_grid<G, Size>::interpolate(mac_grid<G, C>& source)
{
	afm_ctx* c = afm_detail::context<G>();
	std::vector<double> X = packed_positions();
	afm_detail::Buffer dX(c, 2 * Size * 8), f(c, G::size * 8), U(c, Size * 8);
	dX.put(X.data(), 2 * Size);
	f.put(source.data(), G::size);
	afm_detail::check(afm_interp(c, C == mac::x ? 0 : 1, f.d(), dX.d(), Size, U.d()), "afm_interp");
	U.get(this->data(), Size);
}

}  // namespace ib
