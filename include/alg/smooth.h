// alg/smooth.h -- drop-in for the reference's alg/smooth.h (alg/smooth.h:35-51):
// one Gauss-Seidel sweep of A x = b on the GPU.  A filled dt::matrix is swept
// in the reference's lexicographic row order; mg::laplacian is swept red-black
// (the per-row arithmetic is the reference's, the order is red then black).
#pragma once
#include "../dt/matrix.h"
#include "../dt/vector.h"

namespace alg {

template<unsigned int Size>
void smooth(dt::vector<Size>& x, const dt::matrix<Size>& A, const dt::vector<Size>& b)
{
	A.smooth(x, b);
}

}  // namespace alg
