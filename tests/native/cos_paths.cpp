// Host build of afm_delta.h: the fast cos path (table + short series +
// rounding test) against the full double-double series and against cosq
// rounded to double; the fast delta argument (fma products, midpoint test)
// against the exact integer one; the delta against its exact definition.
// Usage: cos_paths N seed -> prints "mismatch_full mismatch_quad mismatch_arg".
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
	// delta arguments: r uniform over the delta's support |n r| < 3 for
	// n = 2 .. 2^13, a quarter of them tiny or near the support's ends
	long bad_arg = 0;
	for (long i = 0; i < N; ++i) {
		const unsigned n = 1u << (1 + (int) (g() % 13));
		double r = (3.0 / n) * ((double) (g() >> 11) * 0x1p-53) * ((g() & 1) ? 1 : -1);
		if (i % 4 == 1) r = std::ldexp(r, -(int) (g() % 60));
		if (i % 4 == 2) r = std::nextafter(3.0 / n, 0.0) * ((g() & 1) ? 1 : -1) - (double) (g() % 1000) * 0x1p-60;
		if (delta_arg(r, n) != delta_arg_exact(r, n)) ++bad_arg;
		// the whole delta (fast paths, flags, support test) against the exact
		// definition, also outside the support (|n r| up to 4)
		const double r2 = (4.0 / n) * ((double) (g() >> 11) * 0x1p-53) * ((g() & 1) ? 1 : -1);
		if (delta(r, n) != delta_exact(r, n) || delta(r2, n) != delta_exact(r2, n)) ++bad_arg;
	}
	printf("%ld %ld %ld\n", bad_full, bad_quad, bad_arg);
	return 0;
}
