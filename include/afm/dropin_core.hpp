// dropin_core.hpp -- what the drop-in headers at the reference's include
// paths (dt/, alg/, mg/, ns/, ib/) share: error checks, device buffers, the
// contexts their calls run on, and the geometry traits.
//
// Every class of those headers keeps its data value-semantic on the host, as
// the reference does (dt/vector.h:33-47); each operation copies its operands
// to device memory, runs the gfx950 kernels of libafm through the C ABI
// (include/afm.h) and copies the results back.  Nothing is computed on the
// host beyond index bookkeeping.
//
// Geometries: the reference's dt::geometry<L, W, H, D> (3-D, walls in y);
// the GPU path takes cubes W = H = D = c (c a power of two, c 2^L <= 512
// cells per axis) for the multigrid operators, and c = 1 for the
// Navier-Stokes solver and the IB transfer.  dt::plane<L, W, D> is the
// doubly periodic unit square of the benchmark (n = 2^L, W = D = 1).  Other
// shapes compile and throw std::invalid_argument when used.
#pragma once
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../afm.h"

namespace dt {
template<unsigned Level, unsigned Width, unsigned Height, unsigned Depth> class geometry;
template<unsigned Level, unsigned Width, unsigned Depth> struct plane;
}

namespace afm_detail {

inline void check(int rc, const char* what)
{
	if (rc != 0) throw std::runtime_error(std::string(what) + ": " + afm_last_error());
}

inline int device_from_env()
{
	const char* dev = std::getenv("AFM_DEVICE");
	return dev ? std::atoi(dev) : 0;
}

struct Buffer {
	afm_ctx* c = nullptr;
	void* p = nullptr;
	Buffer(afm_ctx* ctx, size_t bytes) : c(ctx) { check(afm_alloc(c, bytes ? bytes : 8, &p), "afm_alloc"); }
	~Buffer() { if (p) afm_free(c, p); }
	Buffer(const Buffer&) = delete;
	Buffer& operator=(const Buffer&) = delete;
	double* d() const { return static_cast<double*>(p); }
	template<class T> T* as() const { return static_cast<T*>(p); }
	void put(const void* h, size_t bytes) { if (bytes) check(afm_copy_to_device(c, p, h, bytes), "copy_to_device"); }
	void get(void* h, size_t bytes) const { if (bytes) check(afm_copy_to_host(c, h, p, bytes), "copy_to_host"); }
	void put(const double* h, size_t count) { put((const void*) h, count * 8); }
	void get(double* h, size_t count) const { get((void*) h, count * 8); }
};

// the context the 3-D calls run on: their arrays are their own, only the
// device and the stream of this context are used
inline afm_ctx* context3()
{
	static std::unique_ptr<afm_ctx, int (*)(afm_ctx*)> holder(nullptr, afm_ctx_destroy);
	if (!holder) {
		afm_geom g{2, 8, 1, 1};
		afm_ctx* c = nullptr;
		check(afm_ctx_create(&g, device_from_env(), &c), "afm_ctx_create");
		holder.reset(c);
	}
	return holder.get();
}

// one context per plane geometry (its multigrid hierarchy), created on first
// use and kept until the process ends, as the reference keeps its static
// per-level laplacians (mg/solver.h:68-69)
template<class G>
afm_ctx* context2()
{
	static std::unique_ptr<afm_ctx, int (*)(afm_ctx*)> holder(nullptr, afm_ctx_destroy);
	if (!holder) {
		afm_geom g{2, G::n, G::width, G::depth};
		afm_ctx* c = nullptr;
		check(afm_ctx_create(&g, device_from_env(), &c), "afm_ctx_create");
		holder.reset(c);
	}
	return holder.get();
}

template<class G> struct is_plane { static constexpr bool value = false; };
template<unsigned L, unsigned W, unsigned D> struct is_plane<dt::plane<L, W, D>> { static constexpr bool value = true; };

// cubes of c 2^L cells per axis, c a power of two (the 3-D multigrid path)
template<class G>
inline void require_cube()
{
	constexpr unsigned c = G::width;
	if (!(G::width == G::height && G::height == G::depth && c >= 1 && (c & (c - 1)) == 0 && G::nx <= 512 &&
	      G::nx >= 2))
		throw std::invalid_argument("the GPU path takes cubes geometry<L, c, c, c> (c a power of two, "
		                            "2..512 cells per axis)");
}

// geometry<L, 1, 1, 1> (the 3-D Navier-Stokes solver and IB transfer)
template<class G>
inline void require_unit_cube()
{
	if (!(G::width == 1 && G::height == 1 && G::depth == 1 && G::level >= 1 && G::level <= 8))
		throw std::invalid_argument("the GPU path takes geometry<L, 1, 1, 1>, 1 <= L <= 8, here");
}

inline unsigned log2u(unsigned v)
{
	unsigned l = 0;
	while ((1u << l) < v) ++l;
	return l;
}

}  // namespace afm_detail
