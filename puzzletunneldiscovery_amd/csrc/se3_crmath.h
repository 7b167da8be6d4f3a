/*
 * se3_crmath.h -- correctly rounded (round to nearest, ties to even) sin,
 * cos, acos and atan2 in binary64, bit-identical on x86-64 hosts and gfx950.
 *
 * The state arithmetic of lib/osr/osr_state.cc calls the C library through
 * Eigen: Quaternion::slerp uses acos and sin (osr_state.cc:39),
 * angularDistance uses atan2 (osr_state.cc:135), AngleAxis uses sin and cos
 * (osr_state.cc:121-128, 243-259).  On the reference's platform (Ubuntu
 * 18.04, glibc 2.27) these four functions come from the IBM Accurate
 * Mathematical Library, which returns the correctly rounded result.  A
 * correctly rounded result is unique, so computing it reproduces those
 * values bit for bit, on the CPU and on the GPU alike.
 *
 * Method (Ziv's strategy, after the published designs of CRlibm and
 * CORE-MATH; the code here is this library's own):
 *   1. a fast evaluation in double-double arithmetic with a proven error
 *      bound of about 2^-95 relative: argument reduction (Cody-Waite with a
 *      four-part pi/2), a table at multiples of 1/64, short Taylor series;
 *   2. a rounding test: if every value within the error bound of the
 *      double-double result rounds to the same binary64 number, that number
 *      is the correctly rounded result;
 *   3. otherwise (about once in 2^40 calls) the same function evaluated in
 *      256-bit fixed-point arithmetic (error about 2^-240), exact integer
 *      arithmetic that needs no further test for these functions' hardest
 *      known cases, which lie near 2^-(53+70).
 * Exact products use fma on the GPU (and on hosts compiled with FMA) and
 * Dekker's splitting otherwise; both give the exact product error, so the
 * targets agree to the bit.
 *
 * Domains: sin and cos are correctly rounded for |x| < 2^20; beyond that
 * they fall back to a deterministic, not correctly rounded evaluation.
 * acos and atan2 are correctly rounded for all inputs, except that atan2
 * decides ratios |y/x| in [2^-900, 2^-30] from a 2^-150-accurate value.
 */
#ifndef SE3_CRMATH_H
#define SE3_CRMATH_H

#include <math.h>
#include <stdint.h>
#include <string.h>

#include "se3_cr_tables.h"

#ifndef SE3_HD
#if defined(__HIPCC__) || defined(__HIP__)
#define SE3_HD __host__ __device__
#else
#define SE3_HD
#endif
#endif

/* The slow paths run about once in 2^40 calls: their loops stay rolled and
 * their limb arrays in private memory, so that inlining them does not raise
 * the register count of the kernels that use these functions. */
#if defined(__HIP_DEVICE_COMPILE__)
#define SE3_CR_NO_UNROLL _Pragma("clang loop unroll(disable)")
#else
#define SE3_CR_NO_UNROLL
#endif

/* the fast paths inline on the device when SE3_CR_INLINE is set */
#if defined(__HIP_DEVICE_COMPILE__) && defined(SE3_CR_INLINE) && SE3_CR_INLINE
#define SE3_CR_FAST2 __attribute__((always_inline))
#else
#define SE3_CR_FAST2
#endif
#if defined(__GNUC__) || defined(__clang__)
#define SE3_CR_NOINLINE __attribute__((unused))
#else
#define SE3_CR_NOINLINE
#endif

/* ---------------------------------------------------------------------- */
/* Bits and exact arithmetic                                              */
/* ---------------------------------------------------------------------- */
SE3_HD static inline double se3_cr_bits2d(uint64_t b)
{
	double d;
	memcpy(&d, &b, sizeof d);
	return d;
}

SE3_HD static inline uint64_t se3_cr_d2bits(double d)
{
	uint64_t b;
	memcpy(&b, &d, sizeof b);
	return b;
}

/* The exact error a*b - p of p = fl(a*b) (no overflow or underflow). */
SE3_HD static inline double se3_cr_prod_err(double a, double b, double p)
{
#if defined(__HIP_DEVICE_COMPILE__) || defined(__FMA__)
	return fma(a, b, -p);
#else
	const double c = 134217729.0; /* 2^27 + 1: Veltkamp splitting */
	double t = c * a;
	const double ah = t - (t - a), al = a - ah;
	t = c * b;
	const double bh = t - (t - b), bl = b - bh;
	return ((ah * bh - p) + ah * bl + al * bh) + al * bl;
#endif
}

typedef struct {
	double h, l;
} se3_cr_dd;

SE3_HD static inline se3_cr_dd se3_cr_dd_of(double h, double l)
{
	se3_cr_dd r;
	r.h = h;
	r.l = l;
	return r;
}

SE3_HD static inline se3_cr_dd se3_cr_neg(se3_cr_dd a) { return se3_cr_dd_of(-a.h, -a.l); }

/* a + b exactly, |a| >= |b| or a == 0 */
SE3_HD static inline se3_cr_dd se3_cr_fast2sum(double a, double b)
{
	const double s = a + b;
	return se3_cr_dd_of(s, b - (s - a));
}

/* a + b exactly */
SE3_HD static inline se3_cr_dd se3_cr_2sum(double a, double b)
{
	const double s = a + b, bb = s - a;
	return se3_cr_dd_of(s, (a - (s - bb)) + (b - bb));
}

SE3_HD static inline se3_cr_dd se3_cr_2prod(double a, double b)
{
	const double p = a * b;
	return se3_cr_dd_of(p, se3_cr_prod_err(a, b, p));
}

SE3_HD static inline se3_cr_dd se3_cr_add(se3_cr_dd a, se3_cr_dd b)
{
	se3_cr_dd s = se3_cr_2sum(a.h, b.h);
	const se3_cr_dd t = se3_cr_2sum(a.l, b.l);
	double e = s.l + t.h;
	s = se3_cr_fast2sum(s.h, e);
	e = s.l + t.l;
	return se3_cr_fast2sum(s.h, e);
}

SE3_HD static inline se3_cr_dd se3_cr_mul(se3_cr_dd a, se3_cr_dd b)
{
	const se3_cr_dd p = se3_cr_2prod(a.h, b.h);
	const double e = p.l + (a.h * b.l + a.l * b.h);
	return se3_cr_fast2sum(p.h, e);
}

SE3_HD static inline se3_cr_dd se3_cr_mul_d(se3_cr_dd a, double b)
{
	const se3_cr_dd p = se3_cr_2prod(a.h, b);
	const double e = p.l + a.l * b;
	return se3_cr_fast2sum(p.h, e);
}

SE3_HD static inline se3_cr_dd se3_cr_div(se3_cr_dd a, se3_cr_dd b)
{
	const double q1 = a.h / b.h;
	se3_cr_dd r = se3_cr_add(a, se3_cr_neg(se3_cr_mul_d(b, q1)));
	const double q2 = r.h / b.h;
	r = se3_cr_add(r, se3_cr_neg(se3_cr_mul_d(b, q2)));
	const double q3 = r.h / b.h;
	return se3_cr_add(se3_cr_fast2sum(q1, q2), se3_cr_dd_of(q3, 0.0));
}

SE3_HD static inline se3_cr_dd se3_cr_sqrt(se3_cr_dd a)
{
	const double s = sqrt(a.h);
	const se3_cr_dd sq = se3_cr_2prod(s, s);
	const double e = ((a.h - sq.h) - sq.l) + a.l;
	return se3_cr_fast2sum(s, e / (2.0 * s));
}

/*
 * 1 when every real within err of y = h + l (|l| <= ulp(h)/2) rounds to
 * nearest to h, so h is the correctly rounded value of whatever y
 * approximates.  Zero, subnormal and huge h are left to the slow path.
 */
SE3_HD static inline int se3_cr_rounds_to_h(double h, double l, double err)
{
	const uint64_t b = se3_cr_d2bits(h) & 0x7fffffffffffffffull;
	const int e = (int)(b >> 52);
	if (e <= 53 || e >= 0x7fe)
		return 0;
	const double u = se3_cr_bits2d((uint64_t)(e - 52) << 52);          /* ulp(|h|) */
	const double up = 0.5 * u * (1.0 - 0x1p-40);
	const double down = ((b & 0x000fffffffffffffull) == 0 ? 0.25 : 0.5) * u * (1.0 - 0x1p-40);
	const double lm = h < 0.0 ? -l : l;
	return (lm + err < up) && (lm - err > -down);
}

/* ---------------------------------------------------------------------- */
/* 256-bit fixed point: value = W / 2^254, W < 2^256 (limbs LSB first)     */
/* ---------------------------------------------------------------------- */
typedef struct {
	uint64_t w[4];
} se3_fx;

SE3_HD static inline void se3_fx_mul64(uint64_t a, uint64_t b, uint64_t* hi, uint64_t* lo)
{
#if defined(__HIP_DEVICE_COMPILE__)
	*lo = a * b;
	*hi = __umul64hi(a, b);
#else
	const unsigned __int128 p = (unsigned __int128)a * b;
	*lo = (uint64_t)p;
	*hi = (uint64_t)(p >> 64);
#endif
}

SE3_HD static inline se3_fx se3_fx_zero(void)
{
	se3_fx r;
	r.w[0] = r.w[1] = r.w[2] = r.w[3] = 0;
	return r;
}

SE3_HD static inline int se3_fx_cmp(const se3_fx* a, const se3_fx* b)
{
	SE3_CR_NO_UNROLL
	for (int k = 3; k >= 0; k--)
		if (a->w[k] != b->w[k])
			return a->w[k] < b->w[k] ? -1 : 1;
	return 0;
}

SE3_HD static inline int se3_fx_is_zero(const se3_fx* a) { return !(a->w[0] | a->w[1] | a->w[2] | a->w[3]); }

SE3_HD static inline se3_fx se3_fx_add(se3_fx a, se3_fx b)
{
	se3_fx r;
	uint64_t c = 0;
	SE3_CR_NO_UNROLL
	for (int k = 0; k < 4; k++) {
		const uint64_t s = a.w[k] + b.w[k];
		const uint64_t c1 = s < a.w[k];
		r.w[k] = s + c;
		c = c1 | (r.w[k] < s);
	}
	return r;
}

/* a - b, a >= b */
SE3_HD static inline se3_fx se3_fx_sub(se3_fx a, se3_fx b)
{
	se3_fx r;
	uint64_t br = 0;
	SE3_CR_NO_UNROLL
	for (int k = 0; k < 4; k++) {
		const uint64_t d = a.w[k] - b.w[k];
		const uint64_t b1 = a.w[k] < b.w[k];
		r.w[k] = d - br;
		br = b1 | (d < br);
	}
	return r;
}

/* |a - b| and whether a < b */
SE3_HD static inline se3_fx se3_fx_absdiff(se3_fx a, se3_fx b, int* neg)
{
	*neg = se3_fx_cmp(&a, &b) < 0;
	return *neg ? se3_fx_sub(b, a) : se3_fx_sub(a, b);
}

/* a * b (truncated) */
SE3_HD static inline se3_fx se3_fx_mul(se3_fx a, se3_fx b)
{
	uint64_t p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	SE3_CR_NO_UNROLL
	for (int i = 0; i < 4; i++) {
		uint64_t carry = 0;
		SE3_CR_NO_UNROLL
		for (int j = 0; j < 4; j++) {
			uint64_t hi, lo;
			se3_fx_mul64(a.w[i], b.w[j], &hi, &lo);
			uint64_t s = p[i + j] + lo;
			hi += s < lo;
			s += carry;
			hi += s < carry;
			p[i + j] = s;
			carry = hi;
		}
		p[i + 4] = carry;
	}
	se3_fx r;
	SE3_CR_NO_UNROLL
	for (int k = 0; k < 4; k++)
		r.w[k] = (p[k + 3] >> 62) | (p[k + 4] << 2);
	return r;
}

/* a / d for 0 < d < 2^32 (truncated) */
SE3_HD static inline se3_fx se3_fx_div_u32(se3_fx a, uint32_t d)
{
	se3_fx r;
	uint64_t rem = 0;
	SE3_CR_NO_UNROLL
	for (int k = 3; k >= 0; k--) {
		const uint64_t hi = (rem << 32) | (a.w[k] >> 32);
		const uint64_t qh = hi / d;
		rem = hi - qh * d;
		const uint64_t lo = (rem << 32) | (a.w[k] & 0xffffffffull);
		const uint64_t ql = lo / d;
		rem = lo - ql * d;
		r.w[k] = (qh << 32) | ql;
	}
	return r;
}

/* a / b (truncated), a < 4 b, b > 0: restoring division of a * 2^254 by b */
SE3_HD static SE3_CR_NOINLINE se3_fx se3_fx_div(se3_fx a, se3_fx b)
{
	uint64_t r[5] = {0, 0, 0, 0, 0};
	const uint64_t bb[5] = {b.w[0], b.w[1], b.w[2], b.w[3], 0};
	se3_fx q = se3_fx_zero();
	SE3_CR_NO_UNROLL
	for (int i = 509; i >= 0; i--) {
		/* r = 2 r + bit i of a * 2^254 (bit i - 254 of a) */
		SE3_CR_NO_UNROLL
		for (int k = 4; k > 0; k--)
			r[k] = (r[k] << 1) | (r[k - 1] >> 63);
		const int j = i - 254;
		r[0] = (r[0] << 1) | (j >= 0 ? (a.w[j >> 6] >> (j & 63)) & 1u : 0u);
		int ge = 1;
		SE3_CR_NO_UNROLL
		for (int k = 4; k >= 0; k--)
			if (r[k] != bb[k]) {
				ge = r[k] > bb[k];
				break;
			}
		if (ge) {
			uint64_t br = 0;
			SE3_CR_NO_UNROLL
			for (int k = 0; k < 5; k++) {
				const uint64_t d = r[k] - bb[k];
				const uint64_t b1 = r[k] < bb[k];
				r[k] = d - br;
				br = b1 | (d < br);
			}
			if (i < 256)
				q.w[i >> 6] |= 1ull << (i & 63);
		}
	}
	return q;
}

/* x >= 0, x < 4, as fixed point (truncated below 2^-254) */
SE3_HD static inline se3_fx se3_fx_from_double(double x)
{
	se3_fx r = se3_fx_zero();
	if (x == 0.0)
		return r;
	const uint64_t b = se3_cr_d2bits(x);
	int e = (int)((b >> 52) & 0x7ff);
	uint64_t m = b & 0x000fffffffffffffull;
	if (e == 0)
		e = 1;
	else
		m |= 1ull << 52;
	/* x = m 2^(e - 1075); W = m 2^(e - 1075 + 254) */
	const int sh = e - 1075 + 254;
	if (sh >= 0) {
		const int w = sh >> 6, bit = sh & 63;
		r.w[w] = m << bit;
		if (bit && w + 1 < 4)
			r.w[w + 1] = m >> (64 - bit);
	} else if (sh > -53) {
		r.w[0] = m >> (-sh);
	}
	return r;
}

/* the binary64 nearest to (neg ? -1 : 1) * W / 2^254 (ties to even); W is
 * at least 2^60 in the domains that call this */
SE3_HD static inline double se3_fx_to_double(se3_fx a, int neg)
{
	int top = -1;
	SE3_CR_NO_UNROLL
	for (int k = 3; k >= 0 && top < 0; k--)
		if (a.w[k])
			top = 64 * k + 63 - __builtin_clzll(a.w[k]);
	if (top < 0)
		return neg ? -0.0 : 0.0;
	/* mantissa = bits [top - 52, top], then the round bit and the sticky bits */
	uint64_t m = 0;
	SE3_CR_NO_UNROLL
	for (int i = 0; i < 53; i++) {
		const int bi = top - i;
		m = (m << 1) | (bi >= 0 ? (a.w[bi >> 6] >> (bi & 63)) & 1u : 0u);
	}
	const int rb = top - 53;
	int round = 0, sticky = 0;
	if (rb >= 0) {
		round = (int)((a.w[rb >> 6] >> (rb & 63)) & 1u);
		SE3_CR_NO_UNROLL
		for (int k = 0; k < 4 && !sticky; k++) {
			uint64_t mask;
			if (64 * k + 63 < rb)
				mask = ~0ull;
			else if (64 * k > rb - 1)
				mask = 0;
			else
				mask = (rb & 63) == 0 ? 0 : ((1ull << (rb & 63)) - 1);
			if (a.w[k] & mask)
				sticky = 1;
		}
	}
	int exp2 = top - 254; /* value = m 2^(exp2 - 52) before rounding */
	if (round && (sticky || (m & 1u))) {
		m++;
		if (m >> 53) {
			m >>= 1;
			exp2++;
		}
	}
	const uint64_t bits = ((uint64_t)(exp2 + 1023) << 52) | (m & 0x000fffffffffffffull);
	const double v = se3_cr_bits2d(bits);
	return neg ? -v : v;
}

SE3_HD static inline se3_fx se3_fx_const(const uint64_t* c)
{
	se3_fx r;
	SE3_CR_NO_UNROLL
	for (int k = 0; k < 4; k++)
		r.w[k] = c[k];
	return r;
}

/* a * 2^s for the small left shifts the callers need (0 <= s < 64) */
SE3_HD static inline se3_fx se3_fx_shl(se3_fx a, int s)
{
	if (s == 0)
		return a;
	se3_fx r;
	SE3_CR_NO_UNROLL
	for (int k = 3; k > 0; k--)
		r.w[k] = (a.w[k] << s) | (a.w[k - 1] >> (64 - s));
	r.w[0] = a.w[0] << s;
	return r;
}

/* sum of (-1)^k x^(2k+1) / (2k+1)! (sine) or (-1)^k x^(2k) / (2k)! (cosine), 0 <= x < 1 */
SE3_HD static SE3_CR_NOINLINE se3_fx se3_fx_sincos_series(se3_fx x, int want_cos)
{
	const se3_fx x2 = se3_fx_mul(x, x);
	se3_fx one = se3_fx_zero();
	one.w[3] = 1ull << 62;
	se3_fx term = want_cos ? one : x;
	se3_fx pos = term, neg = se3_fx_zero();
	SE3_CR_NO_UNROLL
	for (uint32_t n = want_cos ? 1 : 2, k = 1; k < 60; k++, n += 2) {
		term = se3_fx_div_u32(se3_fx_mul(term, x2), n * (n + 1));
		if (se3_fx_is_zero(&term))
			break;
		if (k & 1)
			neg = se3_fx_add(neg, term);
		else
			pos = se3_fx_add(pos, term);
	}
	return se3_fx_sub(pos, neg);
}

/* atan(t), 0 <= t <= 1/64: t - t^3/3 + t^5/5 - ... */
SE3_HD static SE3_CR_NOINLINE se3_fx se3_fx_atan_series(se3_fx t)
{
	const se3_fx t2 = se3_fx_mul(t, t);
	se3_fx p = t, pos = t, neg = se3_fx_zero();
	SE3_CR_NO_UNROLL
	for (uint32_t k = 1; k < 60; k++) {
		p = se3_fx_mul(p, t2);
		const se3_fx term = se3_fx_div_u32(p, 2 * k + 1);
		if (se3_fx_is_zero(&term))
			break;
		if (k & 1)
			neg = se3_fx_add(neg, term);
		else
			pos = se3_fx_add(pos, term);
	}
	return se3_fx_sub(pos, neg);
}

/* atan(n / d) for 0 <= n <= d, d > 0, both below 4: the angle in [0, pi/4] */
SE3_HD static SE3_CR_NOINLINE se3_fx se3_fx_atan_ratio(se3_fx n, se3_fx d)
{
	const se3_fx z = se3_fx_div(n, d);
	/* the nearest multiple c = i/64 of 1/64, then atan(z) = atan(c) + atan((z - c) / (1 + z c)) */
	se3_fx half = se3_fx_zero();
	half.w[3] = 1ull << 55; /* 2^-7 = 1/128 */
	const se3_fx zr = se3_fx_add(z, half);
	uint32_t i = (uint32_t)(zr.w[3] >> 56);
	if (i > 64)
		i = 64;
	se3_fx c = se3_fx_zero();
	c.w[3] = (uint64_t)i << 56;
	int neg = 0;
	const se3_fx num = se3_fx_absdiff(z, c, &neg);
	se3_fx one = se3_fx_zero();
	one.w[3] = 1ull << 62;
	const se3_fx den = se3_fx_add(one, se3_fx_mul(z, c));
	const se3_fx t = se3_fx_div(num, den);
	const se3_fx at = se3_fx_atan_series(t);
	const se3_fx base = se3_fx_const(se3_cr_fx_atan64[i]);
	return neg ? se3_fx_sub(base, at) : se3_fx_add(base, at);
}

/* ---------------------------------------------------------------------- */
/* Slow paths                                                             */
/* ---------------------------------------------------------------------- */
/* sin (want_cos = 0) or cos of x, 2^-27 <= |x| < 2^20, to 2^-240 */
SE3_HD static SE3_CR_NOINLINE double se3_cr_sincos_slow(double x, int want_cos)
{
	const double ax = fabs(x);
	/* k = the integer nearest ax * 2/pi (off by one near half-integers,
	 * which only moves r outside [-pi/4, pi/4] by a little) */
	const double kd = floor(ax * SE3_CR_TWO_OVER_PI + 0.5);
	const uint64_t k = (uint64_t)kd;
	/* D = ax 2^318 - k floor(pi/2 2^318), in six limbs, two's complement */
	uint64_t X[6] = {0, 0, 0, 0, 0, 0}, KP[6] = {0, 0, 0, 0, 0, 0};
	{
		const uint64_t b = se3_cr_d2bits(ax);
		const int e = (int)((b >> 52) & 0x7ff);
		const uint64_t m = (b & 0x000fffffffffffffull) | (1ull << 52);
		const int sh = e - 1075 + 318; /* >= 318 - 27 - 53 */
		const int w = sh >> 6, bit = sh & 63;
		X[w] = m << bit;
		if (bit && w + 1 < 6)
			X[w + 1] = m >> (64 - bit);
	}
	{
		uint64_t carry = 0;
		SE3_CR_NO_UNROLL
		for (int j = 0; j < 6; j++) {
			uint64_t hi, lo;
			se3_fx_mul64(se3_cr_pio2_318[j], k, &hi, &lo);
			const uint64_t s = lo + carry;
			hi += s < lo;
			KP[j] = s;
			carry = hi;
		}
	}
	uint64_t D[6];
	uint64_t br = 0;
	SE3_CR_NO_UNROLL
	for (int j = 0; j < 6; j++) {
		const uint64_t d = X[j] - KP[j];
		const uint64_t b1 = X[j] < KP[j];
		D[j] = d - br;
		br = b1 | (d < br);
	}
	const int rneg = (int)(D[5] >> 63);
	if (rneg) {
		/* negate */
		uint64_t c = 1;
		SE3_CR_NO_UNROLL
		for (int j = 0; j < 6; j++) {
			D[j] = ~D[j] + c;
			c = (c && D[j] == 0) ? 1 : 0;
		}
	}
	/* r = D 2^-318; as W = r 2^254 = D >> 64 */
	se3_fx r;
	SE3_CR_NO_UNROLL
	for (int j = 0; j < 4; j++)
		r.w[j] = D[j + 1];
	const int q = (int)(k & 3u);
	/* sin(x) = sgn(x) sin(ax); sin(ax) = [sin r, cos r, -sin r, -cos r][q];
	 * cos(x) = cos(ax) = [cos r, -sin r, -cos r, sin r][q] */
	int use_cos, neg;
	if (!want_cos) {
		use_cos = q & 1;
		neg = (q & 2) != 0;
		if (!use_cos && rneg)
			neg = !neg;
		if (x < 0.0)
			neg = !neg;
	} else {
		use_cos = !(q & 1);
		neg = (q == 1 || q == 2);
		if (!use_cos && rneg)
			neg = !neg;
	}
	const se3_fx v = se3_fx_sincos_series(r, use_cos);
	return se3_fx_to_double(v, neg);
}

/* atan2 for finite nonzero y, x with 2^-30 <= min/max ratio, to 2^-240 */
SE3_HD static SE3_CR_NOINLINE double se3_cr_atan2_slow(double y, double x)
{
	const double ax = fabs(x), ay = fabs(y);
	const int swap = ay > ax;
	double n = swap ? ax : ay, d = swap ? ay : ax;
	/* scale both so d lies in [1, 2) (exact) */
	const int e = ilogb(d);
	n = ldexp(n, -e);
	d = ldexp(d, -e);
	se3_fx a = se3_fx_atan_ratio(se3_fx_from_double(n), se3_fx_from_double(d));
	if (swap)
		a = se3_fx_sub(se3_fx_const(se3_cr_fx_pio2), a);
	if (x < 0.0)
		a = se3_fx_sub(se3_fx_const(se3_cr_fx_pi), a);
	return se3_fx_to_double(a, y < 0.0);
}

/* acos for |x| < 1, x != 0, to 2^-240 */
SE3_HD static SE3_CR_NOINLINE double se3_cr_acos_slow(double x)
{
	const se3_fx X = se3_fx_from_double(fabs(x));
	se3_fx one = se3_fx_zero();
	one.w[3] = 1ull << 62;
	const se3_fx s2 = se3_fx_sub(one, se3_fx_mul(X, X));
	/* Y = sqrt(1 - x^2) by Newton from the binary64 square root */
	se3_fx Y = se3_fx_from_double(sqrt(1.0 - fabs(x)) * sqrt(1.0 + fabs(x)));
	SE3_CR_NO_UNROLL
	for (int it = 0; it < 4; it++)
		Y = se3_fx_div_u32(se3_fx_add(Y, se3_fx_div(s2, Y)), 2);
	/* the angle of (X, Y) in [0, pi/2] */
	se3_fx a;
	if (se3_fx_cmp(&Y, &X) <= 0)
		a = se3_fx_atan_ratio(Y, X);
	else
		a = se3_fx_sub(se3_fx_const(se3_cr_fx_pio2), se3_fx_atan_ratio(X, Y));
	if (x < 0.0)
		a = se3_fx_sub(se3_fx_const(se3_cr_fx_pi), a);
	return se3_fx_to_double(a, 0);
}

/* ---------------------------------------------------------------------- */
/* Fast paths                                                             */
/* ---------------------------------------------------------------------- */
/* sin(s) and cos(s) for |s| <= 2^-7 (and a little), double-double */
SE3_HD static inline void se3_cr_sincos_small(se3_cr_dd s, se3_cr_dd* so, se3_cr_dd* co)
{
	const se3_cr_dd z = se3_cr_mul(s, s);
	const double zh = z.h;
	/* S(z) = -1/6 + z (1/120 + z (-1/5040 + z (1/362880 + z (-1/39916800 + z / 6227020800)))) */
	const double t2 = -0x1.a01a01a01a01ap-13 +
	                  zh * (0x1.71de3a556c734p-19 + zh * (-0x1.ae64567f544e4p-26 + zh * 0x1.6124613a86d09p-33));
	se3_cr_dd S = se3_cr_add(se3_cr_dd_of(SE3_CR_INV120_HI, SE3_CR_INV120_LO), se3_cr_mul_d(z, t2));
	S = se3_cr_add(se3_cr_dd_of(-SE3_CR_INV6_HI, -SE3_CR_INV6_LO), se3_cr_mul(z, S));
	*so = se3_cr_add(s, se3_cr_mul(s, se3_cr_mul(z, S)));
	/* C(z) = -1/2 + z (1/24 + z (-1/720 + z (1/40320 + z (-1/3628800 + z / 479001600)))) */
	const double u3 = 0x1.a01a01a01a01ap-16 + zh * (-0x1.27e4fb7789f5cp-22 + zh * 0x1.1eed8eff8d898p-29);
	se3_cr_dd C = se3_cr_add(se3_cr_dd_of(-SE3_CR_INV720_HI, -SE3_CR_INV720_LO), se3_cr_dd_of(zh * u3, 0.0));
	C = se3_cr_add(se3_cr_dd_of(SE3_CR_INV24_HI, SE3_CR_INV24_LO), se3_cr_mul(z, C));
	C = se3_cr_add(se3_cr_dd_of(-0.5, 0.0), se3_cr_mul(z, C));
	*co = se3_cr_add(se3_cr_dd_of(1.0, 0.0), se3_cr_mul(z, C));
}

/* sin (want_cos = 0) or cos of x, correctly rounded for |x| < 2^20; 0 when
 * |x| is out of that range (the caller falls back) */
SE3_HD static SE3_CR_FAST2 inline int se3_cr_sincos(double x, int want_cos, double* out)
{
	const double ax = fabs(x);
	if (!(ax < 0x1p20))
		return 0;
	if (ax < 0x1p-27) {
		/* sin x = x (1 - x^2/6 ...) and cos x = 1 - x^2/2: the corrections
		 * stay below a quarter ulp */
		*out = want_cos ? 1.0 : x;
		return 1;
	}
	double k = 0.0;
	se3_cr_dd r;
	double er = 0.0;
	if (ax <= SE3_CR_PIO4_HI) {
		r = se3_cr_dd_of(ax, 0.0);
	} else {
		k = floor(ax * SE3_CR_TWO_OVER_PI + 0.5);
		const double t1 = ax - k * SE3_CR_PIO2_C1; /* exact: k C1 is exact and close to ax */
		r = se3_cr_2sum(t1, -(k * SE3_CR_PIO2_C2)); /* k C2 exact */
		r = se3_cr_add(r, se3_cr_neg(se3_cr_2prod(k, SE3_CR_PIO2_C3)));
		r = se3_cr_add(r, se3_cr_neg(se3_cr_2prod(k, SE3_CR_PIO2_C4)));
		er = k * 0x1p-150 + 0x1p-150;
	}
	const int rneg = r.h < 0.0;
	if (rneg)
		r = se3_cr_neg(r);
	int i = (int)floor(r.h * 64.0 + 0.5);
	if (i > 56)
		i = 56;
	const se3_cr_dd s = se3_cr_add(se3_cr_dd_of(r.h - (double)i * 0.015625, 0.0), se3_cr_dd_of(r.l, 0.0));
	se3_cr_dd ss, cs;
	se3_cr_sincos_small(s, &ss, &cs);
	const se3_cr_dd sa = se3_cr_dd_of(se3_cr_sin64[i][0], se3_cr_sin64[i][1]);
	const se3_cr_dd ca = se3_cr_dd_of(se3_cr_cos64[i][0], se3_cr_cos64[i][1]);
	const se3_cr_dd sr = se3_cr_add(se3_cr_mul(sa, cs), se3_cr_mul(ca, ss));              /* sin |r| */
	const se3_cr_dd cr = se3_cr_add(se3_cr_mul(ca, cs), se3_cr_neg(se3_cr_mul(sa, ss))); /* cos |r| */
	const int q = (int)(k - 4.0 * floor(k * 0.25));
	int use_cos, neg;
	if (!want_cos) {
		use_cos = q & 1;
		neg = (q & 2) != 0;
		if (!use_cos && rneg)
			neg = !neg;
		if (x < 0.0)
			neg = !neg;
	} else {
		use_cos = !(q & 1);
		neg = (q == 1 || q == 2);
		if (!use_cos && rneg)
			neg = !neg;
	}
	se3_cr_dd v = use_cos ? cr : sr;
	if (neg)
		v = se3_cr_neg(v);
	if (se3_cr_rounds_to_h(v.h, v.l, fabs(v.h) * 0x1p-95 + er)) {
		*out = v.h;
		return 1;
	}
	*out = se3_cr_sincos_slow(x, want_cos);
	return 1;
}

/* atan(n/d) for 0 <= n <= d (double-double), d > 0: the angle, and an
 * absolute error bound; *tiny = 1 when the ratio is below 2^-900 (then
 * the angle is the ratio itself, see se3_cr_atan2) */
SE3_HD static inline se3_cr_dd se3_cr_atan_ratio(se3_cr_dd n, se3_cr_dd d, double* err, int* tiny)
{
	const se3_cr_dd z = se3_cr_div(n, d);
	*tiny = 0;
	if (z.h < 0x1p-900) {
		*tiny = 1;
		*err = 0.0;
		return z;
	}
	if (z.h < 0x1p-30) {
		/* atan z = z - z^3/3 + z^5/5 ..., z^5/5 below 2^-150 z */
		const double c3 = z.h * z.h * z.h * (1.0 / 3.0);
		const se3_cr_dd a = se3_cr_add(z, se3_cr_dd_of(-c3, 0.0));
		*err = a.h * 0x1p-100;
		return a;
	}
	int i = (int)floor(z.h * 64.0 + 0.5);
	if (i > 64)
		i = 64;
	const double c = (double)i * 0.015625;
	const se3_cr_dd num = se3_cr_add(se3_cr_dd_of(z.h - c, 0.0), se3_cr_dd_of(z.l, 0.0));
	const se3_cr_dd den = se3_cr_add(se3_cr_dd_of(1.0, 0.0), se3_cr_mul_d(z, c));
	const se3_cr_dd t = se3_cr_div(num, den);
	const se3_cr_dd w = se3_cr_mul(t, t);
	const double wh = w.h;
	/* atan t = t + t w A(w), A = -1/3 + w (1/5 + w (-1/7 + w D)), D = 1/9 - w/11 + ... */
	const double D = 0x1.c71c71c71c71cp-4 +
	                 wh * (-0x1.745d1745d1746p-4 +
	                       wh * (0x1.3b13b13b13b14p-4 + wh * (-0x1.1111111111111p-4 + wh * 0x1.e1e1e1e1e1e1ep-5)));
	se3_cr_dd A = se3_cr_add(se3_cr_dd_of(-SE3_CR_INV7_HI, -SE3_CR_INV7_LO), se3_cr_dd_of(wh * D, 0.0));
	A = se3_cr_add(se3_cr_dd_of(SE3_CR_INV5_HI, SE3_CR_INV5_LO), se3_cr_mul(w, A));
	A = se3_cr_add(se3_cr_dd_of(-SE3_CR_INV3_HI, -SE3_CR_INV3_LO), se3_cr_mul(w, A));
	const se3_cr_dd at = se3_cr_add(t, se3_cr_mul(t, se3_cr_mul(w, A)));
	const se3_cr_dd a = se3_cr_add(se3_cr_dd_of(se3_cr_atan64[i][0], se3_cr_atan64[i][1]), at);
	*err = a.h * 0x1p-95;
	return a;
}

/* atan(n/d) rounded to nearest for ratios in [2^-900, 2^-30], from a value
 * accurate to about 2^-150: the quotient in three parts (exact remainders)
 * minus z^3/3 (the next term, z^5/5, is below 2^-150 z) */
SE3_HD static inline double se3_cr_atan_tiny_slow(double n, double d, double sign)
{
	const double q1 = n / d;
	const double r1 = (n - q1 * d) - se3_cr_prod_err(q1, d, q1 * d);
	const double q2 = r1 / d;
	const double r2 = (r1 - q2 * d) - se3_cr_prod_err(q2, d, q2 * d);
	const double q3 = r2 / d;
	const double c3 = q1 * q1 * q1 * (1.0 / 3.0);
	const se3_cr_dd lo = se3_cr_add(se3_cr_dd_of(q2, q3), se3_cr_dd_of(-c3, 0.0));
	const se3_cr_dd v = se3_cr_add(se3_cr_dd_of(q1, 0.0), lo);
	return sign * v.h;
}

SE3_HD static SE3_CR_FAST2 inline double se3_cr_atan2(double y, double x)
{
	if (x != x || y != y)
		return x + y;
	const double ax = fabs(x), ay = fabs(y);
	const int yneg = (se3_cr_d2bits(y) >> 63) != 0;
	const int xneg = (se3_cr_d2bits(x) >> 63) != 0;
	const double sg = yneg ? -1.0 : 1.0;
	if (ay == 0.0)
		return xneg ? sg * SE3_CR_PI_HI : y;
	if (ax == 0.0)
		return sg * SE3_CR_PIO2_HI;
	if (ay == INFINITY) {
		if (ax == INFINITY)
			return sg * (xneg ? 0x1.2d97c7f3321d2p+1 : SE3_CR_PIO4_HI); /* 3 pi / 4 */
		return sg * SE3_CR_PIO2_HI;
	}
	if (ax == INFINITY)
		return xneg ? sg * SE3_CR_PI_HI : sg * 0.0;
	const int swap = ay > ax;
	double n = swap ? ax : ay, d = swap ? ay : ax;
	const double q = n / d;
	if (q >= 0x1p-900) {
		/* scale both so d lies in [1, 2) (exact: n stays normal) */
		const int e = ilogb(d);
		n = ldexp(n, -e);
		d = ldexp(d, -e);
	}
	double err = 0.0;
	int tiny = 0;
	se3_cr_dd a = se3_cr_atan_ratio(se3_cr_dd_of(n, 0.0), se3_cr_dd_of(d, 0.0), &err, &tiny);
	if (tiny) {
		/* |y/x| below 2^-900: atan2 is the ratio itself, or pi/2, pi or
		 * pi/2 adjusted by it, which round to the constants */
		if (!swap && !xneg)
			return sg * q;
		if (swap && !xneg)
			return sg * SE3_CR_PIO2_HI;
		if (!swap && xneg)
			return sg * SE3_CR_PI_HI;
		return sg * SE3_CR_PIO2_HI;
	}
	if (swap)
		a = se3_cr_add(se3_cr_dd_of(SE3_CR_PIO2_HI, SE3_CR_PIO2_LO), se3_cr_neg(a));
	if (xneg)
		a = se3_cr_add(se3_cr_dd_of(SE3_CR_PI_HI, SE3_CR_PI_LO), se3_cr_neg(a));
	err += fabs(a.h) * 0x1p-100;
	if (se3_cr_rounds_to_h(a.h, a.l, err))
		return sg * a.h;
	if (q < 0x1p-30 && !swap && !xneg)
		return se3_cr_atan_tiny_slow(n, d, sg);
	return se3_cr_atan2_slow(y, x);
}

SE3_HD static SE3_CR_FAST2 inline double se3_cr_acos(double x)
{
	if (x != x)
		return x + x;
	const double ax = fabs(x);
	if (ax > 1.0)
		return (x - x) / (x - x); /* NaN */
	if (ax == 1.0)
		return x > 0.0 ? 0.0 : SE3_CR_PI_HI;
	if (ax == 0.0)
		return SE3_CR_PIO2_HI;
	/* y = sqrt((1 - |x|)(1 + |x|)) in double-double */
	const se3_cr_dd om = se3_cr_2sum(1.0, -ax), op = se3_cr_2sum(1.0, ax);
	const se3_cr_dd y = se3_cr_sqrt(se3_cr_mul(om, op));
	const se3_cr_dd xx = se3_cr_dd_of(ax, 0.0);
	double err = 0.0;
	int tiny = 0;
	se3_cr_dd a;
	const int swap = y.h > ax;
	if (!swap) {
		a = se3_cr_atan_ratio(y, xx, &err, &tiny);
	} else {
		a = se3_cr_atan_ratio(xx, y, &err, &tiny);
		a = se3_cr_add(se3_cr_dd_of(SE3_CR_PIO2_HI, SE3_CR_PIO2_LO), se3_cr_neg(a));
	}
	if (x < 0.0)
		a = se3_cr_add(se3_cr_dd_of(SE3_CR_PI_HI, SE3_CR_PI_LO), se3_cr_neg(a));
	err += fabs(a.h) * 0x1p-98;
	if (!tiny && se3_cr_rounds_to_h(a.h, a.l, err))
		return a.h;
	if (tiny && swap)
		return x < 0.0 ? SE3_CR_PIO2_HI + 0.0 : SE3_CR_PIO2_HI; /* pi/2 -/+ a ratio below 2^-900 */
	return se3_cr_acos_slow(x);
}

#endif
