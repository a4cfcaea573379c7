// afm_delta.h -- the 4-point cosine delta of ib/grid.h:51-55, evaluated
// identically on the host and on gfx950.
//
//   delta(r) = 0 if |trunc(n r)| > 2, else (0.25 n) * (1 + cos(theta)),
//   theta    = RN53( RN64(pi_64 * r) * n / 2 ).
//
// The reference forms theta in x87 long double (a 64-bit-mantissa pi literal
// times a double, ib/grid.h:31,54) and rounds it to double on the way into
// cos().  GPUs have no 80-bit type.  pi_64 is split as A + B (A the double
// nearest pi, B = pi_64 - A exact in 10 bits), so pi_64 * r is a sum of four
// doubles formed exactly with fmas; its double rounding RN53(RN64(.)) differs
// from a single rounding only when RN64 lands on a 53-bit midpoint, which the
// sum's low part shows directly.  The few arguments whose low part sits
// within 2^-32 ulp of that decision, or whose rounding is a power of two, take
// the exact integer path (the 117-bit product rounded to 64, then 53 bits).
// n is a power of two, so the scaling by n/2 is exact in either order.
//
// cos is correctly rounded (Ziv): Cody-Waite reduction by pi/2 (four 50-bit
// pieces, so k*P_i is exact for |k| < 8) to y, |y| <= pi/4; then y = j/64 + d,
// |d| <= 1/128, and cos or sin of y is U cos d + V sin d with U, V from a
// double-double table (afm_cos_table.h) and short series in d; if every value
// within 2^-65 of the result rounds alike, that is the result, otherwise (a
// few arguments in 10^4) the full double-double Taylor series on y decides.
// glibc's cos is not correctly rounded (it differs in about 1e-3 of
// arguments), so this is the build's definition of delta; every operation is
// an explicit IEEE op or fma, so host and device produce the same bits.
#pragma once
#include <cstdint>
#include <cmath>
#include <cstring>
#include "afm_cos_table.h"

#ifndef AFM_HD
#define AFM_HD __host__ __device__
#endif
// the fast paths are inlined into the kernels; the rarely taken exact paths
// stay out of line, so a wave that needs one branches to it
#define AFM_INLINE __attribute__((always_inline))
#define AFM_OUTLINE __attribute__((noinline))

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

// cos(j/64), sin(j/64) as double-double, j = -51..51, per parity (see
// afm_cos_table.h).  The device reads its copy from constant memory.
#if defined(__HIP_DEVICE_COMPILE__)
static __constant__ double AFM_COS_TAB[2][2 * AFM_COS_TAB_JMAX + 1][4] = AFM_COS_TAB_INIT;
#else
static const double AFM_COS_TAB[2][2 * AFM_COS_TAB_JMAX + 1][4] = AFM_COS_TAB_INIT;
#endif

// Correctly rounded cos(x) for |x| < 8 (the delta needs |x| <= 3*pi/2).
//
// Fast path, for y = y.hi + y.lo reduced: j = rint(64 y.hi), d = y.hi - j/64
// (exact, |d| <= 2^-7), and cos y (q even) or sin y (q odd) is
//   R = U cos(d + y.lo) + V sin(d + y.lo)
//     = U - U (z/2 + a) - U y.lo d + V d + V b + V y.lo (1 - z/2)
// with d^2 = z + ze exactly, a = ze/2 - z^2/24 + z^3/720 - z^4/40320
// (1 - cos d - z/2; truncation < 2^-91) and b = d z (-1/6 + z/120 -
// z^2/5040) (sin d - d; truncation < 2^-81).  U + V d - U z/2 are summed
// exactly (two_prod, two_sum), the rest in double.  Error, relative to R
// (R >= 0.7 for cos, R >= max(|U|, |d|)/2 for sin): b's own rounding < 2^-67.4
// (|b| < 2^-16.6 R), the fma that adds V b < 2^-69.6, all other terms below
// 2^-70; so R is within 2^-66.5 R of the true value and the test below, with
// a band of 2^-65 R, accepts only results that the exact value rounds to.
// The table-independent part of the fast path: reduction, table row, d and
// the short series.  Split from cos_finish so that a lane evaluating several
// factors issues all its table loads before it waits for any.
struct CosPrep {
	double ylo, d, z, ze, a, b;
	int q, row;           // quadrant; table row (parity * rows + j + JMAX)
};

AFM_HD AFM_INLINE inline void cos_prep(double x, CosPrep& c)
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
	c.q = ((int) k) & 3;
	const double jf = rint(y.hi * 64.0);
	c.d = fma(jf, -0x1p-6, y.hi);
	// |jf| <= 51 for |x| < 8; the clamp keeps the read in the table for
	// any input (NaN included)
	const int ji = (int) fmin(fmax(jf, -(double) AFM_COS_TAB_JMAX), (double) AFM_COS_TAB_JMAX);
	c.row = (c.q & 1) * (2 * AFM_COS_TAB_JMAX + 1) + ji + AFM_COS_TAB_JMAX;
	c.ylo = y.lo;
	const double d = c.d;
	c.z = d * d;
	c.ze = fma(d, d, -c.z);
	const double z = c.z;
	c.a = fma(z * z, fma(z, fma(z, -0x1.a01a01a01a01ap-16, 0x1.6c16c16c16c17p-10), -0x1.5555555555555p-5),
	          0.5 * c.ze);
	c.b = (d * z) * fma(z, fma(z, -0x1.a01a01a01a01ap-13, 0x1.1111111111111p-7), -0x1.5555555555555p-3);
}

AFM_HD AFM_INLINE inline const double* cos_row(const CosPrep& c) { return AFM_COS_TAB[0][c.row]; }

// The rest of the fast path from the table row T = (U hi, lo, V hi, lo).
AFM_HD AFM_INLINE inline double cos_finish(const CosPrep& c, double Uh, double Ul, double Vh, double Vl, bool& ok)
{
	const double d = c.d, z = c.z, ylo = c.ylo;
	const dd p = two_prod(Vh, d);
	const dd w = two_prod(0.5 * Uh, z);
	// terms below 2^-52 R first, then the two larger ones by fma
	double small = ((Ul + p.lo) - w.lo) + fma(Vl, d, fma(Vh, ylo, -(Uh * ylo) * d))
	             - (Ul * z) * 0.5 - ((Vh * ylo) * z) * 0.5;
	small = fma(Vh, c.b, fma(-Uh, c.a, small));
	const dd s1 = two_sum(Uh, p.hi);
	const dd s2 = two_sum(s1.hi, -w.hi);
	const dd r = fast_two_sum(s2.hi, (s1.lo + s2.lo) + small);
	const double e = 0x1p-65 * fabs(r.hi);
	// bitwise & (no short-circuit): the flag costs no branch
	ok = ok & (r.hi + (r.lo + e) == r.hi) & (r.hi + (r.lo - e) == r.hi);
	// cos(y + k pi/2): q=0 cos y, q=1 -sin y, q=2 -cos y, q=3 sin y
	return (c.q == 1 || c.q == 2) ? -r.hi : r.hi;
}

AFM_HD AFM_INLINE inline double cos_fast(double x, bool& ok)
{
	CosPrep c;
	cos_prep(x, c);
	const double* T = cos_row(c);
	return cos_finish(c, T[0], T[1], T[2], T[3], ok);
}

// The degree-14 double-double series on the reduced argument (relative error
// below 2^-105), rounded once: the definition the fast path reproduces, and
// its fallback.
AFM_HD AFM_INLINE inline double cos_cr_full(double x)
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

AFM_HD AFM_OUTLINE inline double cos_cr_slow(double x) { return cos_cr_full(x); }

// Correctly rounded cos(x) for |x| < 8: the fast path, or the full series
// where its rounding test fails.
AFM_HD AFM_INLINE inline double cos_cr(double x)
{
	bool ok = true;
	const double v = cos_fast(x, ok);
	return ok ? v : cos_cr_slow(x);
}

AFM_HD inline uint64_t dbits(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }

AFM_HD inline double dfrom(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }

// theta = RN53(RN64(pi_64 * r)) * n / 2, exactly as the x87 sequence of
// ib/grid.h:54 rounds it (r normal or zero; n a power of two), in integer
// arithmetic: the exact path of delta_arg.
AFM_HD AFM_INLINE inline double delta_arg_exact(double r, unsigned n)
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

// The same value from the exact product pi_64 |r| = h + l + m + m2 (A r = h + l,
// B r = m + m2, by fma).  With H = RN53 of that sum and T its remainder (to
// 2^-100 ulp(H)), RN64 rounds onto the midpoint H + sign(T) ulp/2 exactly
// when ||T| - ulp/2| <= ulp/2^12 (the midpoint is even on the 64-bit grid, so
// both ends of that interval round to it); then RN53 ties to even.  Outside
// that interval RN53(RN64(.)) = H.  Within 2^-32 ulp of its ends, for H a
// power of two (the grid changes below it) and for tiny |r|, it clears ok
// and the exact path decides.
AFM_HD AFM_INLINE inline double delta_arg_fast(double r, unsigned n, bool& ok)
{
	const double A = 0x1.921fb54442d18p+1, B = 0x1.1a8p-53;
	const double ar = fabs(r);
	const double h = A * ar, l = fma(A, ar, -h);
	const double m = B * ar, m2 = fma(B, ar, -m);
	const dd s = two_sum(l, m);
	const dd x = fast_two_sum(h, s.hi);
	const double T = x.lo + (s.lo + m2);
	const uint64_t hb = dbits(x.hi);
	const uint64_t eb = hb & 0x7FF0000000000000ull;
	const double half = dfrom(eb - (53ull << 52));   // ulp(H) / 2
	const double q64 = dfrom(eb - (64ull << 52));    // half the 64-bit ulp at H
	const double D = fabs(fabs(T) - half);
	double v = x.hi;
	if (D < q64 * (1 - 0x1p-20) && (hb & 1ull))
		v = x.hi + copysign(2 * half, T);
	ok = ok & ((D < q64 * (1 - 0x1p-20)) | (D > q64 * (1 + 0x1p-20))) & ((hb & 0xFFFFFFFFFFFFFull) != 0)
	     & (ar >= 0x1p-900);
	v = v * (double) n;
	v = v / 2;
	return r < 0 ? -v : v;
}

AFM_HD AFM_OUTLINE inline double delta_arg_slow(double r, unsigned n) { return delta_arg_exact(r, n); }

AFM_HD AFM_INLINE inline double delta_arg(double r, unsigned n)
{
	bool ok = true;
	const double v = delta_arg_fast(r, n, ok);
	return ok ? v : delta_arg_slow(r, n);
}

// delta by the fast paths; ok is cleared where delta_exact must decide.
AFM_HD AFM_INLINE inline double delta_fast(double r, unsigned n, bool& ok)
{
	// |(int)(n r)| > 2 without the out-of-range cast: outside the support
	const bool in = fabs(n * r) < 3.0;
	bool fok = true;
	const double arg = delta_arg_fast(r, n, fok);
	const double c = cos_fast(in ? arg : 0.0, fok);
	ok = ok & (fok | !in);
	return in ? 0.25 * n * (1 + c) : 0.0;
}

// delta_fast of N factors at once: all arguments and reductions, then all
// table loads, then the rest; bit j of the result set where factor j needs
// delta_exact
template<int N>
AFM_HD AFM_INLINE inline unsigned delta_fast_n(const double (&r)[N], unsigned n, double (&f)[N])
{
	CosPrep c[N];
	bool ok[N], in[N];
#pragma unroll
	for (int j = 0; j < N; ++j) {
		in[j] = fabs(n * r[j]) < 3.0;
		ok[j] = true;
		const double arg = delta_arg_fast(r[j], n, ok[j]);
		cos_prep(in[j] ? arg : 0.0, c[j]);
	}
	double T[N][4];
#pragma unroll
	for (int j = 0; j < N; ++j) {
		const double* t = cos_row(c[j]);
#pragma unroll
		for (int e = 0; e < 4; ++e) T[j][e] = t[e];
	}
#if defined(__HIP_DEVICE_COMPILE__)
	// keep the scheduler from sinking each load to its first use: the loads
	// are then in flight together and the lane waits once
	__builtin_amdgcn_sched_barrier(0);
#endif
	unsigned bad = 0;
#pragma unroll
	for (int j = 0; j < N; ++j) {
		const double cv = cos_finish(c[j], T[j][0], T[j][1], T[j][2], T[j][3], ok[j]);
		f[j] = in[j] ? 0.25 * n * (1 + cv) : 0.0;
		bad |= (ok[j] | !in[j]) ? 0u : 1u << j;
	}
	return bad;
}

// delta by the exact argument and the full series (the definition the fast
// paths reproduce); delta_exact is the same out of line
AFM_HD AFM_INLINE inline double delta_exact_inl(double r, unsigned n)
{
	if (fabs(n * r) >= 3.0) return 0;
	return 0.25 * n * (1 + cos_cr_full(delta_arg_exact(r, n)));
}

AFM_HD AFM_OUTLINE inline double delta_exact(double r, unsigned n) { return delta_exact_inl(r, n); }

AFM_HD AFM_INLINE inline double delta(double r, unsigned n)
{
	bool ok = true;
	const double v = delta_fast(r, n, ok);
	return ok ? v : delta_exact(r, n);
}

}  // namespace afm
