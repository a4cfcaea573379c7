// alg/div.h -- drop-in for the reference's alg/div.h (alg/div.h:29-60): the
// MAC divergence n ((u_x+1 - u_x) + dv + (w_z+1 - w_z)), dv from the interior
// y-faces or the wall values top / bottom, on the GPU (afm_div3).  Sizes as
// the reference's template arguments; the GPU path takes cubes nx = ny = nz
// = n = 2^L.
#pragma once
#include <stdexcept>

#include "../afm/dropin_core.hpp"
#include "../dt/vector.h"

namespace alg {

template<unsigned int nx, unsigned int ny, unsigned int nz, unsigned int n>
void div(dt::vector<nx * ny * nz>& d, dt::vector<nx * ny * nz>& u, dt::vector<nx * (ny - 1) * nz>& v,
         dt::vector<nx * ny * nz>& w, dt::vector<nx * nz>& top, dt::vector<nx * nz>& bottom)
{
	if (!(nx == ny && ny == nz && nz == n && n >= 2 && (n & (n - 1)) == 0 && n <= 256))
		throw std::invalid_argument("alg::div: the GPU path takes nx = ny = nz = n = 2^L <= 256");
	constexpr unsigned N = nx * ny * nz, Nv = nx * (ny - 1) * nz, Nxz = nx * nz;
	afm_ctx* c = afm_detail::context3();
	afm_detail::Buffer bu(c, N * 8), bv(c, Nv * 8), bw(c, N * 8), bt(c, Nxz * 8), bb(c, Nxz * 8), bd(c, N * 8);
	bu.put(u.data(), N); bv.put(v.data(), Nv); bw.put(w.data(), N); bt.put(top.data(), Nxz); bb.put(bottom.data(), Nxz);
	afm_detail::check(afm_div3(c, afm_detail::log2u(n), bu.d(), bv.d(), bw.d(), bt.d(), bb.d(), bd.d()), "afm_div3");
	bd.get(d.data(), N);
}

}  // namespace alg
