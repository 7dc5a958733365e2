/*
 * Correctly rounded cos() and pow() for the FP64 propagation path.
 *
 * The reference calls glibc's cos() in the plane-wave mode sum
 * (src/magneticField/turbulentField/PlaneWaveTurbulence.cpp:237) and pow(r,-0.2)
 * in the adaptive step control (src/module/PropagationBP.cpp:104,111,
 * src/module/PropagationCK.cpp:116,123).  The GPU cannot call glibc, and its own
 * libm is not correctly rounded, so both are evaluated here in double-double
 * arithmetic (about 100 correct bits) and rounded once at the end.
 *
 * Method (textbook; Muller, "Elementary Functions", ch. 11, and the
 * Dekker/Knuth error-free transformations):
 *   cos:  Cody-Waite reduction by pi/2 split in 30-bit chunks, then a table of
 *         sin/cos(j/32) combined with Taylor series of the remainder.
 *   pow:  x^y = exp(y log x); log via a 256-entry table of log(1+(j+.5)/256)
 *         and a log1p series, exp via 2^(j/256) and an expm1 series.
 * Tables come from gen_crmath_tables.py (Python decimal, 90 digits).
 *
 * The same source compiles as C11 (host; the oracle's "cr" mode) and as HIP
 * device code.  Products are split exactly with fma on the device and with
 * Dekker's splitting on the host; both are exact, so the two builds give the
 * same bits.  Build with -ffp-contract=off.
 */
#ifndef CRP_CRMATH_H
#define CRP_CRMATH_H

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define CRP_MATH_FN static inline __device__
#define CRP_MATH_COLD static __device__ __attribute__((noinline)) /* rare paths: kept out of the callers' registers */
#define CRP_TABLE static __device__ __constant__ const
#else
#define CRP_MATH_FN static inline
#define CRP_MATH_COLD static
#define CRP_TABLE static const
#endif

#include "crp_crmath_tables.h"

typedef struct {
	double hi, lo;
} crp_dd;

CRP_MATH_FN crp_dd crp_mkdd(double hi, double lo) {
	crp_dd r;
	r.hi = hi;
	r.lo = lo;
	return r;
}

CRP_MATH_FN uint64_t crp_as_u64(double x) {
	uint64_t u;
	memcpy(&u, &x, 8);
	return u;
}

CRP_MATH_FN double crp_from_u64(uint64_t u) {
	double x;
	memcpy(&x, &u, 8);
	return x;
}

/* round to nearest integer (ties away from zero is irrelevant here: only used
 * to pick reduction indices; any nearby integer keeps the method exact) */
CRP_MATH_FN double crp_rint(double x) {
#if defined(__HIPCC__)
	return __builtin_rint(x);
#else
	const double big = 6755399441055744.0; /* 1.5 * 2^52 */
	if (x > 4503599627370496.0 || x < -4503599627370496.0)
		return x;
	return (x + big) - big;
#endif
}

/* 2^e for integer e in the normal range */
CRP_MATH_FN double crp_pow2i(int e) {
	return crp_from_u64((uint64_t)(e + 1023) << 52);
}

/* ---------------------------------------------------------- error-free ops */
CRP_MATH_FN crp_dd crp_two_sum(double a, double b) {
	double s = a + b;
	double bb = s - a;
	double e = (a - (s - bb)) + (b - bb);
	return crp_mkdd(s, e);
}

CRP_MATH_FN crp_dd crp_fast_two_sum(double a, double b) {
	double s = a + b;
	double e = b - (s - a);
	return crp_mkdd(s, e);
}

CRP_MATH_FN crp_dd crp_two_prod(double a, double b) {
	double p = a * b;
#if defined(__HIP_DEVICE_COMPILE__)
	double e = __builtin_fma(a, b, -p);
#else
	const double split = 134217729.0; /* 2^27 + 1 */
	double ca = split * a, cb = split * b;
	double ah = ca - (ca - a), al = a - ah;
	double bh = cb - (cb - b), bl = b - bh;
	double e = ((ah * bh - p) + ah * bl + al * bh) + al * bl;
#endif
	return crp_mkdd(p, e);
}

CRP_MATH_FN crp_dd crp_dd_add(crp_dd a, crp_dd b) {
	crp_dd s = crp_two_sum(a.hi, b.hi);
	crp_dd t = crp_two_sum(a.lo, b.lo);
	s.lo += t.hi;
	s = crp_fast_two_sum(s.hi, s.lo);
	s.lo += t.lo;
	return crp_fast_two_sum(s.hi, s.lo);
}

CRP_MATH_FN crp_dd crp_dd_add_d(crp_dd a, double b) {
	crp_dd s = crp_two_sum(a.hi, b);
	s.lo += a.lo;
	return crp_fast_two_sum(s.hi, s.lo);
}

CRP_MATH_FN crp_dd crp_dd_mul(crp_dd a, crp_dd b) {
	crp_dd p = crp_two_prod(a.hi, b.hi);
	p.lo += a.hi * b.lo + a.lo * b.hi;
	return crp_fast_two_sum(p.hi, p.lo);
}

CRP_MATH_FN crp_dd crp_dd_mul_d(crp_dd a, double b) {
	crp_dd p = crp_two_prod(a.hi, b);
	p.lo += a.lo * b;
	return crp_fast_two_sum(p.hi, p.lo);
}

CRP_MATH_FN crp_dd crp_dd_neg(crp_dd a) {
	return crp_mkdd(-a.hi, -a.lo);
}

CRP_MATH_FN crp_dd crp_tab_dd(const double *t, int i) {
	return crp_mkdd(t[2 * i], t[2 * i + 1]);
}

/* ---------------------------------------------------------- log (dd) */
/* log(x) for finite x > 0, absolute error about 2^-104. */
CRP_MATH_FN crp_dd crp_log_dd(double x) {
	int e = 0;
	uint64_t u = crp_as_u64(x);
	if ((u >> 52) == 0) { /* subnormal */
		x *= 18014398509481984.0; /* 2^54 */
		u = crp_as_u64(x);
		e = -54;
	}
	e += (int)(u >> 52) - 1023;
	double m = crp_from_u64((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL); /* [1,2) */
	int j = (int)((m - 1.0) * 256.0);
	double c = 1.0 + ((double)j + 0.5) * (1.0 / 256.0);
	double d = m - c; /* exact */
	/* t = d / c as a double-double */
	double th = d / c;
	crp_dd pc = crp_two_prod(th, c);
	double tl = ((d - pc.hi) - pc.lo) / c;
	crp_dd t = crp_fast_two_sum(th, tl);
	/* log1p(t) = t - t^2/2 + t^3/3 - ... , |t| <= 2^-9; terms from t^6 in double */
	double tt = t.hi;
	double tail = -1.0 / 6 + tt * (1.0 / 7 + tt * (-1.0 / 8 + tt * (1.0 / 9 + tt * (-1.0 / 10 + tt * (1.0 / 11)))));
	/* Horner in dd: q = 1/5 + t*tail */
	crp_dd q = crp_dd_add(crp_tab_dd(crp_recip_tab, 5), crp_dd_mul_d(t, tail));
	q = crp_dd_add(crp_mkdd(-0.25, 0.0), crp_dd_mul(t, q));
	q = crp_dd_add(crp_tab_dd(crp_recip_tab, 3), crp_dd_mul(t, q));
	q = crp_dd_add(crp_mkdd(-0.5, 0.0), crp_dd_mul(t, q));
	q = crp_dd_add(crp_mkdd(1.0, 0.0), crp_dd_mul(t, q));
	crp_dd l1p = crp_dd_mul(t, q);
	/* e*ln2 + log(c_j) + log1p(t) */
	double fe = (double)e;
	crp_dd acc = crp_two_sum(fe * CRP_LN2_1, crp_log_tab[2 * j]);
	acc = crp_dd_add_d(acc, fe * CRP_LN2_2);
	acc = crp_dd_add_d(acc, crp_log_tab[2 * j + 1]);
	acc = crp_dd_add_d(acc, fe * CRP_LN2_3);
	acc = crp_dd_add(acc, l1p);
	return acc;
}

/* ---------------------------------------------------------- exp (dd) */
/* exp(z) for |z| < 700 as an unscaled double-double (value = r * 2^*scale). */
CRP_MATH_FN crp_dd crp_exp_dd(crp_dd z, int *scale) {
	double k = crp_rint(z.hi * CRP_INV_LN2_256);
	/* r = z - k*ln2/256; k*M1, k*M2 are exact (|k| < 2^20) */
	double r0 = z.hi - k * CRP_LN2_256_1; /* exact (Sterbenz) */
	crp_dd r = crp_two_sum(r0, -k * CRP_LN2_256_2);
	r = crp_dd_add_d(r, z.lo);
	r = crp_dd_add_d(r, -k * CRP_LN2_256_3);
	/* expm1(r), |r| <= 2^-9.5: r + r^2/2 + ... ; r^5.. in double */
	double rr = r.hi;
	double tail = 1.0 / 120 + rr * (1.0 / 720 + rr * (1.0 / 5040 + rr * (1.0 / 40320 + rr * (1.0 / 362880))));
	crp_dd q = crp_dd_add(crp_tab_dd(crp_invfact_tab, 4), crp_dd_mul_d(r, tail)); /* 1/24 + r*tail */
	q = crp_dd_add(crp_tab_dd(crp_invfact_tab, 3), crp_dd_mul(r, q));              /* 1/6 */
	q = crp_dd_add(crp_mkdd(0.5, 0.0), crp_dd_mul(r, q));                           /* 1/2 */
	q = crp_dd_add(crp_mkdd(1.0, 0.0), crp_dd_mul(r, q));
	crp_dd em1 = crp_dd_mul(r, q);
	long long ki = (long long)k;
	int jj = (int)(ki & 255);
	long long kk = (ki - jj) / 256;
	crp_dd T = crp_tab_dd(crp_exp2_tab, jj);
	crp_dd res = crp_dd_add(T, crp_dd_mul(T, em1));
	*scale = (int)kk;
	return res;
}

/*
 * pow(x, y), correctly rounded for finite x > 0 and |y log2 x| < 1000
 * (the step-control use has y = -0.2).  Special values follow C99 pow for
 * non-integer y: pow(0,y<0)=+inf, pow(0,y>0)=0, pow(inf,y<0)=0, pow(x<0,y)=NaN.
 */
CRP_MATH_FN double crp_pow_cr(double x, double y) {
	if (x != x || y != y)
		return x + y;
	if (x == 1.0 || y == 0.0)
		return 1.0;
	if (x < 0.0)
		return (x - x) / (x - x); /* NaN */
	if (x == 0.0)
		return y < 0 ? 1.0 / 0.0 : 0.0;
	if (x == 1.0 / 0.0)
		return y < 0 ? 0.0 : x;
	crp_dd l = crp_log_dd(x);
	crp_dd z = crp_dd_mul_d(l, y);
	int sc;
	crp_dd r = crp_exp_dd(z, &sc);
	double v = r.hi + r.lo; /* the one rounding */
	/* 2^sc in two halves keeps the intermediate normal */
	int s1 = sc / 2, s2 = sc - s1;
	return v * crp_pow2i(s1) * crp_pow2i(s2);
}

/* ---------------------------------------------------------- large arguments */
/*
 * |x| = M 2^e reduced modulo pi/2 for any finite |x| >= 2^23 pi/2 (Payne and Hanek): only
 * the bits of 2/pi from 2^-(e-1) on matter modulo 4, so a 224-bit window of the table
 * is multiplied by the 53-bit integer M exactly (32-bit limbs).  The two bits above
 * the binary point are the quadrant, the 128 below it the fraction f; with f >= 1/2
 * the quadrant advances and f - 1 is used, so r = f pi/2 has |r| <= pi/4.  The
 * window carries at least 190 fraction bits, so f is exact to 2^-128.
 */
CRP_MATH_COLD crp_dd crp_reduce_large(double ax, int *quadrant) {
	const uint64_t u = crp_as_u64(ax);
	const int e = (int)(u >> 52) - 1075; /* ax = M 2^e */
	const uint64_t M = (u & 0x000fffffffffffffULL) | (1ULL << 52);
	int k0 = e - 1; /* first bit of 2/pi that matters modulo 4 */
	if (k0 < 1)
		k0 = 1;
	const int w0 = (k0 - 1) / 32, sh = (k0 - 1) % 32;
	uint32_t W[7]; /* little endian: W[0] least significant */
	for (int i = 0; i < 7; i++)
		W[i] = crp_two_over_pi_bits[w0 + 6 - i];
	if (sh)
		W[6] &= 0xffffffffu >> sh;
	const uint32_t Mw[2] = {(uint32_t)M, (uint32_t)(M >> 32)};
	uint32_t R[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
	for (int i = 0; i < 7; i++) {
		uint64_t carry = 0;
		for (int j = 0; j < 2; j++) {
			const uint64_t t = (uint64_t)W[i] * Mw[j] + R[i + j] + carry;
			R[i + j] = (uint32_t)t;
			carry = t >> 32;
		}
		for (int k = i + 2; carry && k < 9; k++) {
			const uint64_t t = (uint64_t)R[k] + carry;
			R[k] = (uint32_t)t;
			carry = t >> 32;
		}
	}
	const int F = 32 * (w0 + 7) - e; /* fraction bits of R */
	uint64_t hi = 0, lo = 0;
	{
		/* hi = bits F-64 .. F-1, lo = bits F-128 .. F-65 */
		const int p1 = F - 64, p2 = F - 128;
		uint64_t t1 = 0, t2 = 0;
		for (int b = 0; b < 64; b++) {
			const int q1 = p1 + b, q2 = p2 + b;
			t1 |= (uint64_t)((R[q1 / 32] >> (q1 % 32)) & 1u) << b;
			t2 |= (uint64_t)((R[q2 / 32] >> (q2 % 32)) & 1u) << b;
		}
		hi = t1;
		lo = t2;
	}
	int n = (int)((R[F / 32] >> (F % 32)) & 1u) | (int)(((R[(F + 1) / 32] >> ((F + 1) % 32)) & 1u) << 1);
	double sign = 1.0;
	if (hi >> 63) { /* f >= 1/2: use f - 1 = -(2^128 - (hi, lo)) / 2^128 */
		n = (n + 1) & 3;
		lo = ~lo + 1;
		hi = ~hi + (lo == 0 ? 1 : 0);
		sign = -1.0;
	}
	/* f as a double-double: the top 53 bits, then the next 64 rounded */
	const double f1 = (double)(hi >> 11) * 0x1p-53;
	const double f2 = (double)(((hi & 0x7ffULL) << 53) | (lo >> 11)) * 0x1p-117;
	crp_dd f = crp_fast_two_sum(f1, f2);
	crp_dd r = crp_dd_mul(f, crp_mkdd(CRP_PIO2_DD_HI, CRP_PIO2_DD_LO));
	*quadrant = n;
	return crp_mkdd(sign * r.hi, sign * r.lo);
}

/* ---------------------------------------------------------- cos */
/* sin and cos of a double-double a with 0 <= a <= pi/4 + tiny */
CRP_MATH_FN void crp_sincos_dd_reduced(crp_dd a, crp_dd *s_out, crp_dd *c_out) {
	int j = (int)crp_rint(a.hi * 32.0);
	double dh = a.hi - (double)j * (1.0 / 32.0); /* exact */
	crp_dd d = crp_two_sum(dh, a.lo);
	crp_dd z = crp_dd_mul(d, d);
	double zz = z.hi;
	/* sin(d) = d + d*z*(-1/6 + z*(1/120 + z*(-1/5040 + ...))) */
	double ts = -1.0 / 5040 + zz * (1.0 / 362880 + zz * (-1.0 / 39916800 + zz * (1.0 / 6227020800.0)));
	crp_dd ps = crp_dd_add(crp_tab_dd(crp_invfact_tab, 5), crp_dd_mul_d(z, ts)); /* 1/120 */
	ps = crp_dd_add(crp_dd_neg(crp_tab_dd(crp_invfact_tab, 3)), crp_dd_mul(z, ps)); /* -1/6 */
	crp_dd sd_m = crp_dd_mul(crp_dd_mul(d, z), ps);                                /* sin(d) - d */
	crp_dd sd = crp_dd_add(d, sd_m);
	/* cos(d) - 1 = z*(-1/2 + z*(1/24 + z*(-1/720 + ...))) */
	double tc = 1.0 / 40320 + zz * (-1.0 / 3628800 + zz * (1.0 / 479001600.0));
	crp_dd pcq = crp_dd_add(crp_dd_neg(crp_tab_dd(crp_invfact_tab, 6)), crp_dd_mul_d(z, tc)); /* -1/720 */
	pcq = crp_dd_add(crp_tab_dd(crp_invfact_tab, 4), crp_dd_mul(z, pcq));                    /* 1/24 */
	pcq = crp_dd_add(crp_mkdd(-0.5, 0.0), crp_dd_mul(z, pcq));
	crp_dd cdm1 = crp_dd_mul(z, pcq); /* cos(d) - 1 */
	if (j == 0) {
		*s_out = sd;
		*c_out = crp_dd_add_d(cdm1, 1.0);
		return;
	}
	crp_dd S = crp_mkdd(crp_sincos_tab[4 * j], crp_sincos_tab[4 * j + 1]);
	crp_dd C = crp_mkdd(crp_sincos_tab[4 * j + 2], crp_sincos_tab[4 * j + 3]);
	/* sin(t+d) = S + (S*(cos d - 1) + C*sin d);  cos(t+d) = C + (C*(cos d - 1) - S*sin d) */
	crp_dd s_corr = crp_dd_add(crp_dd_mul(S, cdm1), crp_dd_mul(C, sd));
	crp_dd c_corr = crp_dd_add(crp_dd_mul(C, cdm1), crp_dd_neg(crp_dd_mul(S, sd)));
	*s_out = crp_dd_add(S, s_corr);
	*c_out = crp_dd_add(C, c_corr);
}

/*
 * cos(x), accurate phase: about 2^-100 relative accuracy.  The reduction keeps
 * about 2^-130 absolute accuracy for |x| < 2^23 * pi/2 (1.3e7); beyond it
 * products q*P_k stay exact (two_prod) but cancellation is only tracked to
 * double-double precision.
 */
CRP_MATH_COLD double crp_cos_accurate_shift(double x, int shift);
CRP_MATH_FN double crp_cos_accurate(double x) { return crp_cos_accurate_shift(x, 0); }

/* cos(x) for shift 0, sin(x) for shift 3: sin(q pi/2 + r) = cos((q - 1) pi/2 + r) */
CRP_MATH_COLD double crp_cos_accurate_shift(double x, int shift) {
	int flip = 0;
	double ax = x < 0 ? -x : x;
	if (!(ax < 1.0 / 0.0))
		return x - x; /* NaN for inf and NaN */
	if (ax < 0x1p-27)
		return shift ? x : 1.0; /* sin x = x, cos x = 1 after rounding */
	double q = crp_rint(x * CRP_TWO_OVER_PI);
	crp_dd r;
	if (q == 0.0) {
		r = crp_mkdd(x, 0.0);
	} else if (q < 8388608.0 && q > -8388608.0) {
		double r0 = x - q * CRP_PIO2_1; /* exact */
		r = crp_two_sum(r0, -q * CRP_PIO2_2);
		r = crp_dd_add_d(r, -q * CRP_PIO2_3);
		r = crp_dd_add_d(r, -q * CRP_PIO2_4);
		r = crp_dd_add_d(r, -q * CRP_PIO2_5);
	} else {
		/* cos is even and sin odd: reduce |x|, flip the sine's sign at the end */
		int n0;
		r = crp_reduce_large(ax, &n0);
		q = (double)n0;
		if (x < 0 && shift)
			flip = 1;
	}
	int neg = r.hi < 0;
	crp_dd a = neg ? crp_dd_neg(r) : r;
	crp_dd s, c;
	crp_sincos_dd_reduced(a, &s, &c);
	if (neg)
		s = crp_dd_neg(s);
	long long qi = (long long)q;
	int n = (int)((qi + shift) & 3);
	crp_dd res;
	switch (n) {
	case 0: res = c; break;
	case 1: res = crp_dd_neg(s); break;
	case 2: res = crp_dd_neg(c); break;
	default: res = s; break;
	}
	return flip ? -(res.hi + res.lo) : res.hi + res.lo;
}

/*
 * cos(x), quick phase (Ziv's strategy).  On a GPU every lane of a wave waits for
 * the slowest, so the accurate phase must be rare per wave, not only per call:
 * this phase aims at 2^-67 relative accuracy.  After the same reduction,
 * a = |r| = j/32 + d with d = d_h + d_l, z = d_h^2 = z_h + z_l (exact), and
 *   cos a = C - S d_h - C z_h/2 + (small),   sin a = S + C d_h - S z_h/2 + (small),
 * the three leading terms formed and added exactly (two_prod, two_sum), the rest
 * (at most 2^-19 of the result: the low parts, C or S times the z^2 series, the
 * d^3 series, the d_l cross terms) in double.  Error budget relative to the
 * result: series truncation < 2^-80, the small terms' rounding < 2^-69.5, the
 * reduction < 2^-96 for |r| >= 2^-30 (smaller |r| goes to the accurate phase).
 * The result is returned when every value within 2^-66 |y_h| of y_h + y_l rounds
 * to y_h (half an ulp, a quarter below a power of two): about 1 argument in
 * 4000 is left to the accurate phase.  Returns 0 when it does not decide.
 */
CRP_MATH_FN int crp_cos_quick_shift(double x, int shift, double *out) {
	double ax = x < 0 ? -x : x;
	if (!(ax < 1.0 / 0.0) || ax < 0x1p-27)
		return 0;
	double q = crp_rint(x * CRP_TWO_OVER_PI);
	if (!(q < 8388608.0 && q > -8388608.0))
		return 0;
	crp_dd r;
	if (q == 0.0) {
		r = crp_mkdd(x, 0.0);
	} else {
		double r0 = x - q * CRP_PIO2_1; /* exact */
		r = crp_two_sum(r0, -q * CRP_PIO2_2);
		r = crp_dd_add_d(r, -q * CRP_PIO2_3);
		r = crp_dd_add_d(r, -q * CRP_PIO2_4);
		r = crp_dd_add_d(r, -q * CRP_PIO2_5);
	}
	int neg = r.hi < 0;
	double ah = neg ? -r.hi : r.hi, al = neg ? -r.lo : r.lo;
	if (ah < 0x1p-30)
		return 0;
	int j = (int)crp_rint(ah * 32.0);
	double dh = ah - (double)j * (1.0 / 32.0); /* exact (Sterbenz for j >= 1) */
	double dl = al;                             /* |dl| <= 2^-54 */
	crp_dd zz = crp_two_prod(dh, dh);
	double z = zz.hi, hz = 0.5 * z, hzl = 0.5 * zz.lo; /* z/2 = hz + hzl exactly */
	double cz = (z * z) * (1.0 / 24 + z * (-1.0 / 720 + z * (1.0 / 40320 + z * (-1.0 / 3628800.0)))); /* cos d_h - 1 + z/2 */
	double sdm = (dh * z) * (-1.0 / 6 + z * (1.0 / 120 + z * (-1.0 / 5040 + z * (1.0 / 362880.0)))); /* sin d_h - d_h */
	double dcross = dl * dh;   /* d_l sin d_h, to 2^-113 */
	double dsin = dl - dl * hz; /* d_l cos d_h, to 2^-120 */
	long long qi = (long long)q;
	int n = (int)((qi + shift) & 3);
	crp_dd y;
	if (j == 0) {
		if (n & 1) { /* sin a = d_h + (sdm + d_l cos d_h) */
			y = crp_fast_two_sum(dh, sdm + dsin);
		} else { /* cos a = 1 - z/2 + (cz - d_l d_h) */
			crp_dd s = crp_two_sum(1.0, -hz);
			y = crp_fast_two_sum(s.hi, s.lo + ((cz - hzl) - dcross));
		}
	} else {
		double Sh = crp_sincos_tab[4 * j], Sl = crp_sincos_tab[4 * j + 1];
		double Ch = crp_sincos_tab[4 * j + 2], Cl = crp_sincos_tab[4 * j + 3];
		/* sin a = S cos d + C sin d and cos a = C cos d - S sin d are one computation with
		 * (A, B, sg) = (S, C, +1) or (C, S, -1): A - A z/2 + sg B d_h, the small terms likewise.  Every
		 * operation is the same IEEE operation as in the two separate forms (sg only negates, which is
		 * exact, and the sums are the same operand pairs), so the result is identical, and the lanes
		 * of a wave do not split between two paths. */
		const int sn = n & 1;
		const double Ah = sn ? Sh : Ch, Al = sn ? Sl : Cl, Bh = sn ? Ch : Sh, Bl = sn ? Cl : Sl;
		const double sg = sn ? 1.0 : -1.0;
		crp_dd P = crp_two_prod(Bh, dh);
		crp_dd Q = crp_two_prod(Ah, hz);
		crp_dd s1 = crp_two_sum(Ah, sg * P.hi);
		crp_dd s2 = crp_two_sum(s1.hi, -Q.hi);
		double small = Ah * (cz - dcross) + sg * (Bh * (sdm + dsin));
		small += ((Al + sg * (Bl * dh)) - Al * hz) - Ah * hzl;
		small += (s1.lo + s2.lo) + (sg * P.lo - Q.lo);
		y = crp_fast_two_sum(s2.hi, small);
	}
	/* rounding test; y.hi > 0 here */
	uint64_t ub = crp_as_u64(y.hi);
	double u = crp_from_u64(((ub >> 52) - 52) << 52); /* ulp(y.hi) */
	double lim = ((ub & 0x000fffffffffffffULL) == 0 && y.lo < 0) ? 0.25 * u : 0.5 * u;
	double ayl = y.lo < 0 ? -y.lo : y.lo;
	if (!(ayl + y.hi * 0x1p-66 < lim))
		return 0;
	/* cos x = cos a, -sin r, -cos a, sin r for n = 0..3, with sin r = -sin a when r < 0 */
	int minus = (n == 2) || (n == 1 && !neg) || (n == 3 && neg);
	*out = minus ? -y.hi : y.hi;
	return 1;
}

CRP_MATH_FN int crp_cos_quick(double x, double *out) { return crp_cos_quick_shift(x, 0, out); }

/* cos(x), correctly rounded */
CRP_MATH_FN double crp_cos_cr(double x) {
	double v;
	if (crp_cos_quick_shift(x, 0, &v))
		return v;
	return crp_cos_accurate_shift(x, 0);
}

/* sin(x), correctly rounded (the same reduction, the quadrant shifted by one) */
CRP_MATH_FN double crp_sin_cr(double x) {
	double v;
	if (crp_cos_quick_shift(x, 3, &v))
		return v;
	return crp_cos_accurate_shift(x, 3);
}

/* log(x) for finite x > 0: the double-double logarithm (absolute error about 2^-104)
 * rounded once; the same bits on the host and the device */
CRP_MATH_FN double crp_log_dd_rn(double x) {
	crp_dd l = crp_log_dd(x);
	return l.hi + l.lo;
}

/* ---------------------------------------------------------- fast, deterministic (not correctly rounded)
 * For the random draws of DiffusionSDE's own stream, where only the same bits on the host and the
 * device matter (errors of about one ulp): plain double arithmetic in a fixed order.
 *
 * log(x), finite x > 0: x = 2^e m, m in [1/sqrt2, sqrt2), f = m - 1 (exact), s = f / (2 + f),
 * log m = 2 atanh s = 2 s (1 + s^2/3 + s^4/5 + ... + s^20/21) (|s| <= 0.172: the next term is below
 * 2^-60), plus e ln2 with ln2 in two parts (the 40-bit first part times e is exact). */
CRP_MATH_FN double crp_log_fast(double x) {
	if (!(x > 0.0) || x == 1.0 / 0.0)
		return x == 0.0 ? -1.0 / 0.0 : (x > 0.0 ? x : (x - x) / (x - x));
	const uint64_t u = crp_as_u64(x);
	int e = (int)((u >> 52) & 0x7ff) - 1023;
	double m = crp_from_u64((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
	if (e == -1023) { /* subnormal: scale up first */
		const double y = x * 0x1p54;
		const uint64_t v = crp_as_u64(y);
		e = (int)((v >> 52) & 0x7ff) - 1023 - 54;
		m = crp_from_u64((v & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
	}
	if (m > 1.4142135623730951) {
		m *= 0.5;
		e += 1;
	}
	const double f = m - 1.0;
	const double sv = f / (2.0 + f);
	const double z = sv * sv;
	const double P = z * (1.0 / 3 + z * (1.0 / 5 + z * (1.0 / 7 + z * (1.0 / 9 + z * (1.0 / 11 + z * (1.0 / 13 +
	                 z * (1.0 / 15 + z * (1.0 / 17 + z * (1.0 / 19 + z * (1.0 / 21))))))))));
	const double lm = 2.0 * sv + 2.0 * sv * P;
	const double de = (double)e;
	return de * CRP_LN2_1 + (lm + (de * CRP_LN2_2 + de * CRP_LN2_3));
}

/* sin and cos of x, |x| < 2^20: k = rint(x 2/pi), r = x - k pi/2 with the three 30-bit parts of pi/2
 * (products exact) and the rest, |r| <= pi/4, Taylor series of sin r and cos r to r^17 and r^18
 * (the next terms are below 2^-60), then the quadrant of k */
CRP_MATH_FN void crp_sincos_fast(double x, double *sin_out, double *cos_out) {
	const double k = crp_rint(x * CRP_TWO_OVER_PI);
	const double r = (((x - k * CRP_PIO2_1) - k * CRP_PIO2_2) - k * CRP_PIO2_3) - k * CRP_PIO2_4;
	const double z = r * r;
	const double sr = r + r * z * (-1.0 / 6 + z * (1.0 / 120 + z * (-1.0 / 5040 + z * (1.0 / 362880 +
	                  z * (-1.0 / 39916800 + z * (1.0 / 6227020800.0 + z * (-1.0 / 1307674368000.0 +
	                  z * (1.0 / 355687428096000.0))))))));
	const double cr = 1.0 - 0.5 * z + z * z * (1.0 / 24 + z * (-1.0 / 720 + z * (1.0 / 40320 + z * (-1.0 / 3628800 +
	                  z * (1.0 / 479001600.0 + z * (-1.0 / 87178291200.0 + z * (1.0 / 20922789888000.0 +
	                  z * (-1.0 / 6402373705728000.0))))))));
	const int q = (int)((long long)k & 3);
	*sin_out = q == 0 ? sr : (q == 1 ? cr : (q == 2 ? -sr : -cr));
	*cos_out = q == 0 ? cr : (q == 1 ? -sr : (q == 2 ? -cr : sr));
}

/* exp(x): the double-double exponential (relative error about 2^-100) rounded once, so correctly
 * rounded except for arguments whose exact value lies within 2^-100 of a rounding boundary; results
 * below 2^-1022 are rounded twice (the scaling of a subnormal result) */
CRP_MATH_COLD double crp_exp_accurate(double x) {
	int sc;
	crp_dd r = crp_exp_dd(crp_mkdd(x, 0.0), &sc);
	double v = r.hi + r.lo; /* the one rounding */
	int s1 = sc / 2, s2 = sc - s1;
	return v * crp_pow2i(s1) * crp_pow2i(s2);
}

/*
 * Quick phase of exp (Ziv): the reduction of crp_exp_dd (x = k ln2/256 + r, r = rh + rl with
 * |rh| <= 2^-9.4), expm1 r = rh + [rl (1 + rh) + rh^2 (1/2 + rh/6 + rh^2/24 + rh^3/120 + rh^4/720)] in
 * double, and 2^(j/256) (1 + expm1 r) with the one large product T.hi * rh formed exactly.  Error
 * budget relative to the result: the bracketed term (<= 2^-19.8) to a few ulps, 2^-72; the series'
 * truncation 2^-79; T.lo and the products with the bracket 2^-72; the final two additions 2^-73; the
 * table 2^-106 and the reduction 2^-100: below 2^-69 in all.  The value is returned when everything
 * within 2^-66 of it rounds the same way; otherwise the double-double phase decides.  Normal results
 * only (|x| <= 708).
 */
CRP_MATH_FN int crp_exp_quick(double x, double *out) {
	const double k = crp_rint(x * CRP_INV_LN2_256);
	const double r0 = x - k * CRP_LN2_256_1; /* exact */
	const crp_dd r = crp_two_sum(r0, -k * CRP_LN2_256_2);
	const double rh = r.hi, rl = r.lo - k * CRP_LN2_256_3;
	const double q = rh * rh * (0.5 + rh * (1.0 / 6 + rh * (1.0 / 24 + rh * (1.0 / 120 + rh * (1.0 / 720)))));
	const double em = rl + rh * rl + q; /* expm1(r) - rh */
	const long long ki = (long long)k;
	const int j = (int)(ki & 255);
	const long long m = (ki - j) / 256;
	const crp_dd T = crp_tab_dd(crp_exp2_tab, j);
	const crp_dd p = crp_two_prod(T.hi, rh);
	const double lo = p.lo + (T.hi * em + (T.lo + T.lo * (rh + em)));
	const crp_dd sh = crp_fast_two_sum(T.hi, p.hi);
	const double hi = sh.hi, l2 = sh.lo + lo;
	const double eps = hi * 0x1p-66;
	const double a = hi + (l2 + eps), b = hi + (l2 - eps);
	if (a != b)
		return 0;
	const int s1 = (int)(m / 2), s2 = (int)(m - s1);
	*out = (hi + l2) * crp_pow2i(s1) * crp_pow2i(s2);
	return 1;
}

CRP_MATH_FN double crp_exp_rn(double x) {
	if (x != x)
		return x + x;
	if (x > 709.782712893384)
		return 1.0 / 0.0;
	if (x < -745.1332191019412)
		return 0.0;
	double v;
	if (x >= -708.0 && x <= 708.0 && crp_exp_quick(x, &v))
		return v;
	return crp_exp_accurate(x);
}

/* tanh from the correctly rounded exp, (e - 1) / (e + 1) with e = exp(2|x|), odd in x, 1 where e
 * overflows.  Not correctly rounded (near 0 the subtraction loses relative accuracy, down to an absolute
 * error of about 2^-53), but the same bits on the host and the device. */
CRP_MATH_FN double crp_tanh_det(double x) {
	const double e = crp_exp_rn(2 * __builtin_fabs(x));
	const double t = (e == 1.0 / 0.0) ? 1.0 : (e - 1) / (e + 1);
	return __builtin_copysign(t, x);
}

/* a / b in double-double: three quotient digits with exact remainders (about 2^-104) */
CRP_MATH_FN crp_dd crp_dd_div(crp_dd a, crp_dd b) {
	double q1 = a.hi / b.hi;
	crp_dd r = crp_dd_add(a, crp_dd_neg(crp_dd_mul_d(b, q1)));
	double q2 = r.hi / b.hi;
	r = crp_dd_add(r, crp_dd_neg(crp_dd_mul_d(b, q2)));
	double q3 = r.hi / b.hi;
	return crp_dd_add_d(crp_fast_two_sum(q1, q2), q3);
}

/*
 * atan2(y, x) with C99's special cases: t = min(|x|,|y|) / max(|x|,|y|) in double-double,
 * atan t = atan(j/64) + atan(u), u = (t - j/64) / (1 + t j/64), |u| <= 1/128, from the table and
 * the odd series to u^17 (terms from u^9 on in double); then pi/2 - a for |y| > |x| and pi - a for
 * x < 0, the sign of y last; rounded once (about 2^-100 before the rounding).
 */
/*
 * Quick phase of atan2 for finite, nonzero x and y with t = min/max in [2^-500, 1] (Ziv): the same
 * reduction with the numerator th - t0 + tl and the denominator 1 + t t0 kept as unevaluated pairs and
 * one correction step of the division (residuals exact by FMA), u = uh + ul; atan u = uh + [ul + uh z
 * (-1/3 + z/5 - z^2/7 + ... - z^7/15)], z = uh^2, the bracket's polynomial part (<= 2^-21 relative) in
 * double; then atan(j/64) from the table, pi/2 - a and pi - a with the large parts added exactly.
 * Error budget relative to the result (>= 2^-8 or, for j = 0, about u): the polynomial 2^-73, the
 * division 2^-100, the low-part additions 2^-100, the table 2^-106: below 2^-72.  Returned when
 * everything within 2^-66 of the value rounds the same way.
 */
CRP_MATH_FN int crp_atan2_quick(double y, double x, double *out) {
	const double ax = x < 0 ? -x : x, ay = y < 0 ? -y : y;
	const int nx = x < 0, ny = y < 0;
	const int swap = ay > ax;
	const double num = swap ? ax : ay, den = swap ? ay : ax;
	const double th = num / den;
	if (!(th >= 0x1p-500) || !(den < 0x1p1000) || !(num > 0x1p-1000))
		return 0;
	const double tl = __builtin_fma(-th, den, num) / den;
	const double j = crp_rint(th * 64.0), t0 = j * (1.0 / 64.0);
	const double nh = th - t0; /* exact */
	const crp_dd pr = crp_two_prod(th, t0);
	const crp_dd dd = crp_fast_two_sum(1.0, pr.hi);
	const double dh = dd.hi, dl = dd.lo + (pr.lo + tl * t0);
	const double uh = nh / dh;
	const double ul = (__builtin_fma(-uh, dh, nh) + (tl - uh * dl)) / dh;
	const double z = uh * uh;
	const double poly = uh * z * (-1.0 / 3 + z * (1.0 / 5 + z * (-1.0 / 7 + z * (1.0 / 9 + z * (-1.0 / 11 + z * (1.0 / 13 +
	                    z * (-1.0 / 15)))))));
	const crp_dd A = crp_tab_dd(crp_atan_tab, (int)j);
	crp_dd s = crp_two_sum(A.hi, uh);
	double lo = s.lo + (A.lo + (ul + poly));
	double hi = s.hi;
	if (swap) {
		s = crp_two_sum(CRP_PIO2_DD_HI, -hi);
		lo = s.lo + (CRP_PIO2_DD_LO - lo);
		hi = s.hi;
	}
	if (nx) {
		s = crp_two_sum(CRP_PI_DD_HI, -hi);
		lo = s.lo + (CRP_PI_DD_LO - lo);
		hi = s.hi;
	}
	const double eps = hi * 0x1p-66;
	if (hi + (lo + eps) != hi + (lo - eps))
		return 0;
	const double v = hi + lo;
	*out = ny ? -v : v;
	return 1;
}

CRP_MATH_COLD double crp_atan2_accurate(double y, double x);

CRP_MATH_FN double crp_atan2_rn(double y, double x) {
	double q;
	if (x == x && y == y && x != 0.0 && y != 0.0 && crp_atan2_quick(y, x, &q))
		return q;
	return crp_atan2_accurate(y, x);
}

CRP_MATH_COLD double crp_atan2_accurate(double y, double x) {
	if (x != x || y != y)
		return x + y;
	const double ax = x < 0 ? -x : x, ay = y < 0 ? -y : y;
	const int nx = (crp_as_u64(x) >> 63) != 0, ny = (crp_as_u64(y) >> 63) != 0;
	const crp_dd PI = crp_mkdd(CRP_PI_DD_HI, CRP_PI_DD_LO), PIO2 = crp_mkdd(CRP_PIO2_DD_HI, CRP_PIO2_DD_LO);
	crp_dd a;
	const double inf = 1.0 / 0.0;
	if (ay == 0.0) {
		a = nx ? PI : crp_mkdd(0.0, 0.0);
	} else if (ax == 0.0) {
		a = PIO2;
	} else if (ax == inf && ay == inf) {
		a = crp_dd_mul_d(PIO2, 0.5);
		if (nx)
			a = crp_dd_add(PI, crp_dd_neg(a));
	} else if (ax == inf) {
		a = nx ? PI : crp_mkdd(0.0, 0.0);
	} else if (ay == inf) {
		a = PIO2;
	} else {
		const int swap = ay > ax;
		const double num = swap ? ax : ay, den = swap ? ay : ax;
		const double th = num / den;
		crp_dd at;
		if (th < 0x1p-500) {
			at = crp_mkdd(th, 0.0); /* atan t = t (t^3/3 is far below the rounding) */
		} else {
			const double tl = __builtin_fma(-th, den, num) / den;
			const double j = crp_rint(th * 64.0), t0 = j * (1.0 / 64.0);
			const crp_dd nu = crp_two_sum(th - t0, tl); /* th - t0 is exact */
			crp_dd dn = crp_two_prod(th, t0);
			dn = crp_dd_add_d(crp_dd_add_d(dn, tl * t0), 1.0);
			const crp_dd u = crp_dd_div(nu, dn);
			const crp_dd u2 = crp_dd_mul(u, u);
			const double w = u2.hi;
			const double tail = 1.0 / 9 - w * (1.0 / 11 - w * (1.0 / 13 - w * (1.0 / 15 - w * (1.0 / 17))));
			crp_dd sr = crp_dd_add(crp_dd_neg(crp_tab_dd(crp_recip_tab, 7)), crp_dd_mul_d(u2, tail));
			sr = crp_dd_add(crp_tab_dd(crp_recip_tab, 5), crp_dd_mul(u2, sr));
			sr = crp_dd_add(crp_dd_neg(crp_tab_dd(crp_recip_tab, 3)), crp_dd_mul(u2, sr));
			sr = crp_dd_add(crp_mkdd(1.0, 0.0), crp_dd_mul(u2, sr));
			at = crp_dd_add(crp_tab_dd(crp_atan_tab, (int)j), crp_dd_mul(u, sr));
		}
		a = swap ? crp_dd_add(PIO2, crp_dd_neg(at)) : at;
		if (nx)
			a = crp_dd_add(PI, crp_dd_neg(a));
	}
	const double v = a.hi + a.lo;
	return ny ? -v : v;
}

#endif
