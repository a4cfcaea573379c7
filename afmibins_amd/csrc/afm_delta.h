// afm_delta.h -- the 4-point cosine delta of ib/grid.h:51-55, evaluated
// identically on the host and on gfx950.
//
//   delta(r) = 0 if |trunc(n r)| > 2, else (0.25 n) * (1 + cos(theta)),
//   theta    = RN53( RN64(pi_64 * r) * n / 2 ).
//
// The reference forms theta in x87 long double (a 64-bit-mantissa pi literal
// times a double, ib/grid.h:31,54) and rounds it to double on the way into
// cos().  GPUs have no 80-bit type, so theta is reproduced exactly with
// integer arithmetic: the 117-bit product of r's significand and pi's 64-bit
// significand is rounded to 64 bits, then to 53 bits, both to nearest-even.
// n is a power of two, so the scaling by n/2 is exact in either order.
//
// cos is correctly rounded: Cody-Waite reduction by pi/2 (four 50-bit
// pieces, so k*P_i is exact for |k| < 8), then double-double Taylor series
// for sin/cos on |y| <= pi/4 (truncation below 2^-105 relative), rounded
// once.  glibc's cos is not correctly rounded (it differs in about 1e-3 of
// arguments), so this is the build's definition of delta; every operation is
// an explicit IEEE op or fma, so host and device produce the same bits.
#pragma once
#include <cstdint>
#include <cmath>
#include <cstring>

#ifndef AFM_HD
#define AFM_HD __host__ __device__
#endif

namespace afm {

struct dd { double hi, lo; };

AFM_HD inline dd two_sum(double a, double b)
{
	double s = a + b;
	double bb = s - a;
	double e = (a - (s - bb)) + (b - bb);
	return {s, e};
}

AFM_HD inline dd fast_two_sum(double a, double b)
{
	double s = a + b;
	double e = b - (s - a);
	return {s, e};
}

AFM_HD inline dd two_prod(double a, double b)
{
	double p = a * b;
	double e = fma(a, b, -p);
	return {p, e};
}

AFM_HD inline dd dd_mul(dd a, dd b)
{
	dd p = two_prod(a.hi, b.hi);
	double t1 = a.hi * b.lo;
	double t2 = a.lo * b.hi;
	double e = p.lo + (t1 + t2);
	return fast_two_sum(p.hi, e);
}

AFM_HD inline dd dd_add(dd a, dd b)
{
	dd s = two_sum(a.hi, b.hi);
	dd t = two_sum(a.lo, b.lo);
	double e = s.lo + t.hi;
	dd u = fast_two_sum(s.hi, e);
	double e2 = u.lo + t.lo;
	return fast_two_sum(u.hi, e2);
}

// (-1)^j / (2j)!  and  (-1)^j / (2j+1)!  as double-double.
AFM_HD constexpr dd cos_coef(int j)
{
	switch (j) {
	case 0: return {0x1.0000000000000p+0, 0x0.0p+0};
	case 1: return {-0x1.0000000000000p-1, 0x0.0p+0};
	case 2: return {0x1.5555555555555p-5, 0x1.5555555555555p-59};
	case 3: return {-0x1.6c16c16c16c17p-10, 0x1.f49f49f49f49fp-65};
	case 4: return {0x1.a01a01a01a01ap-16, 0x1.a01a01a01a01ap-76};
	case 5: return {-0x1.27e4fb7789f5cp-22, -0x1.cbbc05b4fa99ap-76};
	case 6: return {0x1.1eed8eff8d898p-29, -0x1.2aec959e14c06p-83};
	case 7: return {-0x1.93974a8c07c9dp-37, -0x1.05d6f8a2efd1fp-92};
	case 8: return {0x1.ae7f3e733b81fp-45, 0x1.1d8656b0ee8cbp-101};
	case 9: return {-0x1.6827863b97d97p-53, -0x1.eec01221a8b0bp-107};
	case 10: return {0x1.e542ba4020225p-62, 0x1.ea72b4afe3c2fp-120};
	case 11: return {-0x1.0ce396db7f853p-70, 0x1.aebcdbd20331cp-124};
	case 12: return {0x1.f2cf01972f578p-80, -0x1.9ada5fcc1ab14p-135};
	case 13: return {-0x1.88e85fc6a4e5ap-89, 0x1.71c37ebd16540p-143};
	default: return {0x1.0a18a2635085dp-98, 0x1.b9e2e28e1aa54p-153};
	}
}

AFM_HD constexpr dd sin_coef(int j)
{
	switch (j) {
	case 0: return {0x1.0000000000000p+0, 0x0.0p+0};
	case 1: return {-0x1.5555555555555p-3, -0x1.5555555555555p-57};
	case 2: return {0x1.1111111111111p-7, 0x1.1111111111111p-63};
	case 3: return {-0x1.a01a01a01a01ap-13, -0x1.a01a01a01a01ap-73};
	case 4: return {0x1.71de3a556c734p-19, -0x1.c154f8ddc6c00p-73};
	case 5: return {-0x1.ae64567f544e4p-26, 0x1.c062e06d1f209p-80};
	case 6: return {0x1.6124613a86d09p-33, 0x1.f28e0cc748ebep-87};
	case 7: return {-0x1.ae7f3e733b81fp-41, -0x1.1d8656b0ee8cbp-97};
	case 8: return {0x1.952c77030ad4ap-49, 0x1.ac981465ddc6cp-103};
	case 9: return {-0x1.2f49b46814157p-57, -0x1.2650f61dbdcb4p-112};
	case 10: return {0x1.71b8ef6dcf572p-66, -0x1.d043ae40c4647p-120};
	case 11: return {-0x1.761b41316381ap-75, 0x1.3423c7d91404fp-130};
	case 12: return {0x1.3f3ccdd165fa9p-84, -0x1.58ddadf344487p-139};
	default: return {-0x1.d1ab1c2dccea3p-94, -0x1.054d0c78aea14p-149};
	}
}

#if defined(__HIP_DEVICE_COMPILE__)
// The fast path's coefficients of one parity (0 cos, 1 sin), in the order the
// series takes them: c10 .. c5 (double), then c4 .. c0 as (hi, lo).  On the
// GPU a lane reads its parity's row: the lanes of a wave alternate between
// the two series, and loads replace the per-lane selection of every constant
// between two register pairs.  The values are cos_coef / sin_coef's.
#define AFM_FC(c) {c(10).hi, c(9).hi, c(8).hi, c(7).hi, c(6).hi, c(5).hi, c(4).hi, c(4).lo, c(3).hi, c(3).lo, \
                   c(2).hi, c(2).lo, c(1).hi, c(1).lo, c(0).hi, c(0).lo}
static __constant__ double AFM_FAST_COEF[2][16] = {AFM_FC(cos_coef), AFM_FC(sin_coef)};
#undef AFM_FC
#endif

// a + b without the careful low-order sum of dd_add; fine when the sum does
// not cancel (every step of the series below adds a smaller term to a larger)
AFM_HD inline dd dd_add_nc(dd a, dd b)
{
	dd s = two_sum(a.hi, b.hi);
	double e = s.lo + (a.lo + b.lo);
	return fast_two_sum(s.hi, e);
}

// Correctly rounded cos(x) for |x| < 8 (the delta needs |x| <= 3*pi/2).
//
// Ziv's strategy.  The fast evaluation takes the series to degree 10 in z =
// y^2 with the first five coefficients in double-double and the other six in
// double: its relative error is below 2^-75 (tail rounding < 2^-75.5,
// truncation < 2^-77 for cos and < 2^-82 for sin on |y| <= pi/4).  If the
// rounding of every value within 2^-70 of it is the same, that is the
// correctly rounded result; otherwise (about one argument in 2^16) the full
// degree-14 double-double series (2^-105) decides.  The cos and sin series
// share one loop with the coefficients chosen per lane: the four taps of a
// window are pi/2 apart in x, so neighbouring lanes alternate between them.
AFM_HD inline double cos_cr(double x)
{
	const double P1 = 0x1.921fb54442d18p+0, P2 = 0x1.1a62633145c08p-54,
	             P3 = -0x1.1f1976b7ed8f8p-106, P4 = -0x1.dd660ce2ff7d1p-157;
	double ax = fabs(x);
	double k = rint(ax * 0x1.45f306dc9c883p-1);
	// y = ax - k*(P1+P2+P3+P4); k*P_i exact, ax - k*P1 exact (Sterbenz).
	double t = ax - k * P1;
	dd s = two_sum(t, -(k * P2));
	double ylo = (s.lo - k * P3) - k * P4;
	dd y = fast_two_sum(s.hi, ylo);
	dd z = dd_mul(y, y);
	int q = ((int) k) & 3;
	const bool odd = (q & 1) != 0;
	// fast: tail c5..c10 in double (Horner on z.hi), then c4..c0 in dd
#if defined(__HIP_DEVICE_COMPILE__)
	const double* cf = AFM_FAST_COEF[odd ? 1 : 0];
	double pt = cf[0];
#pragma unroll
	for (int j = 1; j < 6; ++j) pt = fma(pt, z.hi, cf[j]);
	dd p = {pt, 0.0};
#pragma unroll
	for (int j = 0; j < 5; ++j) p = dd_add_nc(dd_mul(p, z), dd{cf[6 + 2 * j], cf[7 + 2 * j]});
#else
	double pt = odd ? sin_coef(10).hi : cos_coef(10).hi;
#pragma unroll
	for (int j = 9; j >= 5; --j) pt = fma(pt, z.hi, odd ? sin_coef(j).hi : cos_coef(j).hi);
	dd p = {pt, 0.0};
#pragma unroll
	for (int j = 4; j >= 0; --j) {
		const dd cj = odd ? sin_coef(j) : cos_coef(j);
		p = dd_add_nc(dd_mul(p, z), cj);
	}
#endif
	const dd one = {1.0, 0.0};
	dd r = dd_mul(p, odd ? y : one);   // times 1 is exact
	const double e = 0x1p-70 * fabs(r.hi);
	if (!(r.hi + (r.lo + e) == r.hi && r.hi + (r.lo - e) == r.hi)) {
		if (!odd) {
			dd pc = cos_coef(14);
			for (int j = 13; j >= 0; --j) pc = dd_add(dd_mul(pc, z), cos_coef(j));
			r = pc;
		} else {
			dd ps = sin_coef(13);
			for (int j = 12; j >= 0; --j) ps = dd_add(dd_mul(ps, z), sin_coef(j));
			r = dd_mul(ps, y);
		}
	}
	// cos(y + k pi/2): q=0 cos y, q=1 -sin y, q=2 -cos y, q=3 sin y
	double v = r.hi;
	return (q == 1 || q == 2) ? -v : v;
}

// The degree-14 series alone (the definition the fast path reproduces);
// kept for the tests that compare the two.
AFM_HD inline double cos_cr_full(double x)
{
	const double P1 = 0x1.921fb54442d18p+0, P2 = 0x1.1a62633145c08p-54,
	             P3 = -0x1.1f1976b7ed8f8p-106, P4 = -0x1.dd660ce2ff7d1p-157;
	double ax = fabs(x);
	double k = rint(ax * 0x1.45f306dc9c883p-1);
	double t = ax - k * P1;
	dd s = two_sum(t, -(k * P2));
	double ylo = (s.lo - k * P3) - k * P4;
	dd y = fast_two_sum(s.hi, ylo);
	dd z = dd_mul(y, y);
	int q = ((int) k) & 3;
	dd r;
	if ((q & 1) == 0) {
		dd pc = cos_coef(14);
		for (int j = 13; j >= 0; --j) pc = dd_add(dd_mul(pc, z), cos_coef(j));
		r = pc;
	} else {
		dd ps = sin_coef(13);
		for (int j = 12; j >= 0; --j) ps = dd_add(dd_mul(ps, z), sin_coef(j));
		r = dd_mul(ps, y);
	}
	double v = r.hi;
	return (q == 1 || q == 2) ? -v : v;
}

AFM_HD inline uint64_t dbits(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }

// theta = RN53(RN64(pi_64 * r)) * n / 2, exactly as the x87 sequence of
// ib/grid.h:54 rounds it (r normal or zero; n a power of two).
AFM_HD inline double delta_arg(double r, unsigned n)
{
	if (r == 0.0) return r;
	const uint64_t PM = 0xC90FDAA22168C235ull;   // pi = PM * 2^-62
	uint64_t b = dbits(r);
	int neg = (int) (b >> 63);
	int be = (int) ((b >> 52) & 0x7ff);
	uint64_t m = b & 0xFFFFFFFFFFFFFull;
	int e2;
	if (be == 0) e2 = -1074; else { m |= 1ull << 52; e2 = be - 1075; }
	unsigned __int128 P = (unsigned __int128) m * PM;
	uint64_t hi = (uint64_t) (P >> 64), lo = (uint64_t) P;
	int L = hi ? 128 - __builtin_clzll(hi) : 64 - __builtin_clzll(lo);
	int sh = L - 64;
	uint64_t q;
	if (sh > 0) {
		q = (uint64_t) (P >> sh);
		uint64_t rbit = (lo >> (sh - 1)) & 1ull;
		uint64_t sticky = sh > 1 ? (lo & ((1ull << (sh - 1)) - 1ull)) : 0ull;
		if (rbit && (sticky || (q & 1ull))) {
			q += 1ull;
			if (q == 0ull) { q = 1ull << 63; sh += 1; }
		}
	} else {
		q = lo << (-sh);
	}
	// q * 2^(e2 - 62 + sh) is RN64(pi_64 * |r|); now round to 53 bits.
	uint64_t t = q >> 11;
	uint64_t rb = (q >> 10) & 1ull, st = q & 0x3FFull;
	int ex = e2 - 62 + sh + 11;
	if (rb && (st || (t & 1ull))) {
		t += 1ull;
		if (t == (1ull << 53)) { t >>= 1; ex += 1; }
	}
	double v = ldexp((double) t, ex);
	v = v * (double) n;
	v = v / 2;
	return neg ? -v : v;
}

AFM_HD inline double delta(double r, unsigned n)
{
	double nr = n * r;
	if (fabs(nr) >= 3.0)   // |(int)(n r)| > 2 without the out-of-range cast
		return 0;
	double c = cos_cr(delta_arg(r, n));
	return 0.25 * n * (1 + c);
}

}  // namespace afm
