// alg/grad.h -- drop-in for the reference's alg/grad.h (alg/grad.h:29-44): the
// MAC gradient n (p_i - p_i-1) on the x-, interior y- and z-faces, on the GPU
// (afm_grad3).  The GPU path takes cubes nx = ny = nz = n = 2^L.
#pragma once
#include <stdexcept>

#include "../afm/dropin_core.hpp"
#include "../dt/vector.h"

namespace alg {

template<unsigned nx, unsigned ny, unsigned nz, unsigned n>
void grad(dt::vector<nx * ny * nz>& px, dt::vector<nx * (ny - 1) * nz>& py, dt::vector<nx * ny * nz>& pz,
          dt::vector<nx * ny * nz>& p)
{
	if (!(nx == ny && ny == nz && nz == n && n >= 2 && (n & (n - 1)) == 0 && n <= 256))
		throw std::invalid_argument("alg::grad: the GPU path takes nx = ny = nz = n = 2^L <= 256");
	constexpr unsigned N = nx * ny * nz, Nv = nx * (ny - 1) * nz;
	afm_ctx* c = afm_detail::context3();
	afm_detail::Buffer bp(c, N * 8), bx(c, N * 8), by(c, Nv * 8), bz(c, N * 8);
	bp.put(p.data(), N);
	afm_detail::check(afm_grad3(c, afm_detail::log2u(n), bp.d(), bx.d(), by.d(), bz.d()), "afm_grad3");
	bx.get(px.data(), N); by.get(py.data(), Nv); bz.get(pz.data(), N);
}

}  // namespace alg
