// Host build of afm_delta.h: the fast cos path (series + rounding test)
// against the full double-double series and against cosq rounded to double.
// Usage: cos_paths N seed -> prints "mismatch_full mismatch_quad slow_count".
#define AFM_HD
#include "afm_delta.h"
#include <quadmath.h>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <random>
using namespace afm;

int main(int argc, char** argv)
{
	const long N = argc > 1 ? atol(argv[1]) : 1000000;
	std::mt19937_64 g(argc > 2 ? atoi(argv[2]) : 1);
	std::uniform_real_distribution<double> U(-3 * M_PI / 2, 3 * M_PI / 2);
	long bad_full = 0, bad_quad = 0;
	for (long i = 0; i < N; ++i) {
		double x = U(g);
		if (i % 4 == 0) {
			// near multiples of pi/2: small reduced arguments
			const int kq = (int) (g() % 4);
			const double d = std::ldexp((double) (g() >> 11) * 0x1p-53, -(int) (g() % 50));
			x = kq * (M_PI / 2) + ((g() & 1) ? d : -d);
		}
		const double a = cos_cr(x);
		if (a != cos_cr_full(x)) ++bad_full;
		if (i % 8 == 0 && a != (double) cosq((__float128) x)) ++bad_quad;
	}
	printf("%ld %ld\n", bad_full, bad_quad);
	return 0;
}
