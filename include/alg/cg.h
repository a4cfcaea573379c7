// alg/cg.h -- drop-in for the reference's alg/cg.h (alg/cg.h:32-56): the
// operator's own GPU conjugate gradient (stencil kernels for mg::laplacian
// and ns::helmholtz_*, device CSR for a filled dt::matrix); x is the start
// and the result, the stop test is the reference's |r|_2 <= tol, and the
// return value the number of iterations.
#pragma once
#include "../dt/matrix.h"
#include "../dt/vector.h"

namespace alg {

template<unsigned int Size>
unsigned int conjugate_gradient(dt::vector<Size>& x, const dt::matrix<Size>& A, const dt::vector<Size>& b, double tol)
{
	return A.cg(x, b, tol);
}

}  // namespace alg
