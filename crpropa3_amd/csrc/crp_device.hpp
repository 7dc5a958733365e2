// Device-side building blocks of the batched propagator: field evaluation
// (uniform, Grid3f trilinear, plane waves) and the Boris / Cash-Karp steps.
//
// The FP64 functions follow the reference's operation order exactly (no
// contraction: the library is built with -ffp-contract=off), so that a lane
// reproduces CRPropa3's CPU result bit for bit:
//   Grid<Vector3f>::trilinearInterpolate   include/crpropa/Grid.h:540-582
//   periodicClamp / reflectiveClamp         include/crpropa/Grid.h:31-34,49-59
//   PlaneWaveTurbulence::getField           src/magneticField/turbulentField/PlaneWaveTurbulence.cpp:231-240
//   PropagationBP::dY                       src/module/PropagationBP.cpp:19-39
//   PropagationCK::dYdt / tryStep           src/module/PropagationCK.cpp:28-61
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "crp_crmath.h"

namespace crp {

constexpr double kCLight = 2.99792458e8; // include/crpropa/Units.h:64

struct V3 {
	double x, y, z;
};

__device__ __forceinline__ V3 v3(double x, double y, double z) {
	V3 r;
	r.x = x;
	r.y = y;
	r.z = z;
	return r;
}

// Cash-Karp tableau of src/module/PropagationCK.cpp:11-26 (textbook constants).
struct CK {
	static constexpr double a10 = 1. / 5.;
	static constexpr double a20 = 3. / 40., a21 = 9. / 40.;
	static constexpr double a30 = 3. / 10., a31 = -9. / 10., a32 = 6. / 5.;
	static constexpr double a40 = -11. / 54., a41 = 5. / 2., a42 = -70. / 27., a43 = 35. / 27.;
	static constexpr double a50 = 1631. / 55296., a51 = 175. / 512., a52 = 575. / 13824., a53 = 44275. / 110592.,
	                        a54 = 253. / 4096.;
	static constexpr double b0 = 37. / 378., b1 = 0., b2 = 250. / 621., b3 = 125. / 594., b4 = 0.,
	                        b5 = 512. / 1771.;
	static constexpr double s0 = 2825. / 27648., s1 = 0., s2 = 18575. / 48384., s3 = 13525. / 55296.,
	                        s4 = 277. / 14336., s5 = 1. / 4.;
};

// Vector3<T>::getR: sqrt((x*x + y*y) + z*z)
__device__ __forceinline__ double norm(const V3 &v) { return __builtin_sqrt((v.x * v.x + v.y * v.y) + v.z * v.z); }

// x / s, bit-identical to the IEEE division, from invs = RN(1/s) (one IEEE division shared by
// several quotients with the same divisor): q = x * invs is within one ulp of x/s, the FMA
// residual x - q*s is exact, and RN(q + residual * invs) is the correctly rounded quotient
// (Markstein's correction theorem; J.-M. Muller et al., Handbook of Floating-Point Arithmetic,
// division with an FMA).  The theorem assumes no underflow or overflow: magnitudes outside
// [2^-500, 2^500] (and 0, inf, NaN) take the IEEE division.
__device__ __forceinline__ double div_rn(double x, double s, double invs) {
	const double ax = __builtin_fabs(x), as = __builtin_fabs(s);
	if (!(ax > 0x1p-500 && ax < 0x1p500 && as > 0x1p-500 && as < 0x1p500))
		return x / s;
	const double q = x * invs;
	const double r = __builtin_fma(-q, s, x);
	return __builtin_fma(r, invs, q);
}

constexpr double kInvCLight = 1.0 / 299792458.0; // RN(1/c), folded by the compiler

// ---------------------------------------------------------------- fields
// Parameters as the host computes them from the reference's field objects.
struct GridDesc {
	const float *values; // AoS float3 per point, index (ix*Ny + iy)*Nz + iz
	int nx, ny, nz;
	int reflective, clip;
	double gridOrigin[3]; // origin + spacing/2 (Grid.h:225-228)
	double spacing[3];
	double origin[3], edge[3]; // clipVolume box: origin, origin + N*spacing (Grid.h:320-327)
	double invSpacing[3];      // RN(1/spacing), computed on the host
	double invN[3];            // 1/N per axis, for the division-free periodic wrap
	int fastDiv;               // spacing within [2^-500, 2^500]: div_cr may use its reciprocal path
	int interp;                // GridInterp: Grid::setInterpolationType (Grid.h:24-28, 255-264)
};

// Interpolation types of Grid::interpolate (include/crpropa/Grid.h:24-28, 318-335), numbered as
// the C ABI's crp_grid_interpolation
enum GridInterp { kInterpTrilinear = 0, kInterpTricubic = 1, kInterpNearest = 2 };

// Plane-wave mode tables, one record per mode, read through the constant address space so that
// the compiler loads them with scalar (SMEM) loads into SGPRs: every lane of a wave evaluates the
// same mode at the same time, so one scalar load serves the wave and the VALU instructions take the
// mode's values as scalar operands.  The tables are immutable while kernels run.
struct PwMode64 {    // FP64 path: k, beta, kappa, xi*A (xi*A is the first product of getField)
	double k, beta, kx, ky, kz, ax, ay, az;
};
struct PwModeMixed { // mixed path: phase in turns = p . kt + bt (FP64), amplitude xi*A in FP32
	double ktx, kty, ktz, bt;
	float ax, ay, az, pad;
};
typedef __attribute__((address_space(4))) const PwMode64 PwMode64Const;
typedef __attribute__((address_space(4))) const PwModeMixed PwModeMixedConst;

struct PlaneWaveDesc {
	int nm;
	const PwMode64 *m64;   // [nm]
	const PwModeMixed *mm; // [nm]
};

struct FieldUniform {
	double b[3];
	__device__ __forceinline__ V3 eval(const V3 &) const { return v3(b[0], b[1], b[2]); }
	__device__ __forceinline__ float3 evalf(const V3 &) const { return make_float3((float)b[0], (float)b[1], (float)b[2]); }
};

struct FieldNone {
	__device__ __forceinline__ V3 eval(const V3 &) const { return v3(0.0, 0.0, 0.0); }
	__device__ __forceinline__ float3 evalf(const V3 &) const { return make_float3(0.f, 0.f, 0.f); }
};

// ---------------------------------------------------------------- mixed-precision grid helpers
// Cell index and fraction of a periodic axis without integer division: the
// position in cells is formed with a multiply by 1/spacing (FP64), wrapped
// with floor() in FP64, then the fraction goes to FP32.
__device__ __forceinline__ void periodic_cell_fast(double x, int n, double invn, int &lo, float &frac) {
	double fl = __builtin_floor(x);
	frac = (float)(x - fl);
	double w = fl - (double)n * __builtin_floor(fl * invn);
	int i = (int)w;
	i = i >= n ? i - n : (i < 0 ? i + n : i);
	lo = i;
}

__device__ __forceinline__ float3 trilinear_f32(const float *v, float fx, float fy, float fz) {
	// v: 24 floats, corners in the order of Grid.h:572-579
	const float gx = 1.f - fx, gy = 1.f - fy, gz = 1.f - fz;
	const float w0 = gx * gy * gz, w1 = fx * gy * gz, w2 = gx * fy * gz, w3 = gx * gy * fz;
	const float w4 = fx * gy * fz, w5 = gx * fy * fz, w6 = fx * fy * gz, w7 = fx * fy * fz;
	float b[3];
#pragma unroll
	for (int a = 0; a < 3; a++) {
		float s = v[a] * w0;
		s = __builtin_fmaf(v[3 + a], w1, s);
		s = __builtin_fmaf(v[6 + a], w2, s);
		s = __builtin_fmaf(v[9 + a], w3, s);
		s = __builtin_fmaf(v[12 + a], w4, s);
		s = __builtin_fmaf(v[15 + a], w5, s);
		s = __builtin_fmaf(v[18 + a], w6, s);
		s = __builtin_fmaf(v[21 + a], w7, s);
		b[a] = s;
	}
	return make_float3(b[0], b[1], b[2]);
}

// x / s rounded exactly like the IEEE division the reference performs, without a division:
// q = x * RN(1/s) is within one ulp of x/s, the FMA residual x - q*s is exact, and
// RN(q + residual * RN(1/s)) is the correctly rounded quotient when nothing underflows or
// overflows (Markstein's correction theorem; J.-M. Muller et al., Handbook of Floating-Point
// Arithmetic, the chapter on division with an FMA).  Magnitudes outside [2^-500, 2^500] (and
// 0, inf, NaN) take the IEEE division.
__device__ __forceinline__ double div_cr(double x, double s, double invs, int fast) {
	const double ax = __builtin_fabs(x);
	if (!(fast && ax > 0x1p-500 && ax < 0x1p500))
		return x / s;
	const double q = x * invs;
	const double r = __builtin_fma(-q, s, x);
	return __builtin_fma(r, invs, q);
}

// ((i % n) + n) % n (Grid.h:31-34) without an integer division: floor(i * (1/n)) is within one
// of floor(i/n) for |i| < 2^31, the remainder is corrected into [0, n).
__device__ __forceinline__ int wrap_index(int i, int n, double invn) {
	const long long q = (long long)__builtin_floor((double)i * invn);
	long long r = (long long)i - q * (long long)n;
	r += (r < 0) ? n : 0;
	r -= (r >= n) ? n : 0;
	return (int)r;
}

__device__ __forceinline__ void periodic_clamp(double x, int n, double invn, int &lo, int &hi, double &frac) {
	double fl = __builtin_floor(x);
	lo = wrap_index((int)fl, n, invn);
	hi = (lo + 1 == n) ? 0 : lo + 1;
	frac = x - fl;
}

__device__ __forceinline__ void reflective_clamp(double x, int n, int &lo, int &hi, double &frac) {
	while ((x < -0.5) || (x > (n - 0.5)))
		x = 2 * n * (x > (n - 0.5)) - x - 1;
	double fl = __builtin_floor(x);
	lo = (int)fl;
	hi = lo + (lo < n - 1);
	if (x < 0) {
		lo = 0;
		hi = 0;
	}
	frac = x - fl;
}

// ---------------------------------------------------------------- tricubic and nearest neighbour
// crpropa::round (include/crpropa/Grid.h:61-63): r + 0.5 is rounded before the floor (so 0.5 - 2^-54
// goes up), unlike the C library's round
__device__ __forceinline__ double round_sym(double r) {
	return (r > 0.0) ? __builtin_floor(r + 0.5) : __builtin_ceil(r - 0.5);
}

// reflectiveBoundary (Grid.h:36-41) without its loop: each pass maps the index to its mirror image
// at the nearer face, and the images repeat with period 2n, so the loop ends at i mod 2n folded
// onto [0, n).  Every int maps into [0, n).
__device__ __forceinline__ int reflect_index(int i, int n) {
	const int p = 2 * n;
	int m = i % p;
	m += (m < 0) ? p : 0;
	return (m < n) ? m : p - 1 - m;
}

// Grid::closestValue (include/crpropa/Grid.h:383-405): flat index of the grid point nearest to p,
// mirrored (reflective) or wrapped (periodic) into the grid
__device__ inline size_t grid_closest(const GridDesc &g, const V3 &p) {
	const double rx = div_cr(p.x - g.gridOrigin[0], g.spacing[0], g.invSpacing[0], g.fastDiv);
	const double ry = div_cr(p.y - g.gridOrigin[1], g.spacing[1], g.invSpacing[1], g.fastDiv);
	const double rz = div_cr(p.z - g.gridOrigin[2], g.spacing[2], g.invSpacing[2], g.fastDiv);
	int ix, iy, iz;
	if (g.reflective) {
		ix = reflect_index((int)round_sym(rx), g.nx);
		iy = reflect_index((int)round_sym(ry), g.ny);
		iz = reflect_index((int)round_sym(rz), g.nz);
	} else {
		// fmod(r, n) lies in (-n, n), its rounding in [-n, n]: (i + n * (i < 0)) % n is one correction each way
		ix = (int)round_sym(__builtin_fmod(rx, (double)g.nx));
		iy = (int)round_sym(__builtin_fmod(ry, (double)g.ny));
		iz = (int)round_sym(__builtin_fmod(rz, (double)g.nz));
		ix += (ix < 0) ? g.nx : 0;
		iy += (iy < 0) ? g.ny : 0;
		iz += (iz < 0) ? g.nz : 0;
		ix -= (ix >= g.nx) ? g.nx : 0;
		iy -= (iy >= g.ny) ? g.ny : 0;
		iz -= (iz >= g.nz) ? g.nz : 0;
	}
	// a non-finite position converts to indices outside the grid: read point 0 instead of faulting
	if ((unsigned)ix >= (unsigned)g.nx || (unsigned)iy >= (unsigned)g.ny || (unsigned)iz >= (unsigned)g.nz)
		ix = iy = iz = 0;
	return ((size_t)ix * g.ny + iy) * g.nz + iz;
}

// CubicInterpolate (include/crpropa/Grid.h:436-456), one lane of the SSE code in FP32; t, t^2 and t^3
// are formed in double by the caller and rounded to float once each
__device__ __forceinline__ float cubic_f32(float p0, float p1, float p2, float p3, float t, float t2, float t3) {
	const float term = (0.5f * p2 - 0.5f * p0) * t;
	const float term2 = ((p0 + 2.f * p2) - (2.5f * p1 + 0.5f * p3)) * t2;
	const float term3 = ((1.5f * p1 + 0.5f * p3) - (0.5f * p0 + 1.5f * p2)) * t3;
	return ((term3 + term2) + term) + p1;
}

// CubicInterpolateScalar (include/crpropa/Grid.h:499-501) in FP64, left to right
__device__ __forceinline__ double cubic_f64(double p0, double p1, double p2, double p3, double t) {
	return (((((-0.5 * p0 + 1.5 * p1) - 1.5 * p2) + 0.5 * p3) * t * t * t +
	         (((p0 - 2.5 * p1) + p2 * 2) - 0.5 * p3) * t * t) +
	        (-0.5 * p0 + 0.5 * p2) * t) +
	       p1;
}

// neighbour index i of one axis through periodicGet / reflectiveGet (Grid.h:346-358)
__device__ __forceinline__ int cubic_index(int i, int n, double invn, int reflective) {
	return reflective ? reflect_index(i, n) : wrap_index(i, n, invn);
}

// unit-grid position of p: the reference's r = (pos - gridOrigin) / spacing, iX0 = floor(r),
// fX = r - iX0 (Grid.h:461-472)
__device__ __forceinline__ void cubic_origin(const GridDesc &g, const V3 &p, int i0[3], double f[3]) {
	const double r[3] = {div_cr(p.x - g.gridOrigin[0], g.spacing[0], g.invSpacing[0], g.fastDiv),
	                     div_cr(p.y - g.gridOrigin[1], g.spacing[1], g.invSpacing[1], g.fastDiv),
	                     div_cr(p.z - g.gridOrigin[2], g.spacing[2], g.invSpacing[2], g.fastDiv)};
#pragma unroll
	for (int a = 0; a < 3; a++) {
		i0[a] = (int)__builtin_floor(r[a]);
		f[a] = r[a] - (double)i0[a];
	}
}

// Grid<Vector3f>::tricubicInterpolate (include/crpropa/Grid.h:459-492): cubics along z for each of
// the 16 (x, y) neighbours, along y for each of the 4 x neighbours, then along x; the x, y and z
// components are independent lanes.  The result is FP32, widened to double as getField returns it.
__device__ inline V3 grid_tricubic(const GridDesc &g, const V3 &p) {
	int i0[3];
	double f[3];
	cubic_origin(g, p, i0, f);
	float t[3], t2[3], t3[3];
#pragma unroll
	for (int a = 0; a < 3; a++) {
		t[a] = (float)f[a];
		t2[a] = (float)(f[a] * f[a]);
		t3[a] = (float)(f[a] * f[a] * f[a]);
	}
	int Z[4];
	size_t Y[4];
#pragma unroll
	for (int d = 0; d < 4; d++) {
		Z[d] = cubic_index(i0[2] - 1 + d, g.nz, g.invN[2], g.reflective);
		Y[d] = (size_t)cubic_index(i0[1] - 1 + d, g.ny, g.invN[1], g.reflective) * g.nz;
	}
	float vx[4][3];
#pragma unroll
	for (int a = 0; a < 4; a++) {
		const size_t X = (size_t)cubic_index(i0[0] - 1 + a, g.nx, g.invN[0], g.reflective) * g.ny * g.nz;
		float vy[4][3];
#pragma unroll
		for (int b = 0; b < 4; b++) {
			const float *q0 = g.values + 3 * (X + Y[b] + Z[0]), *q1 = g.values + 3 * (X + Y[b] + Z[1]);
			const float *q2 = g.values + 3 * (X + Y[b] + Z[2]), *q3 = g.values + 3 * (X + Y[b] + Z[3]);
#pragma unroll
			for (int l = 0; l < 3; l++)
				vy[b][l] = cubic_f32(q0[l], q1[l], q2[l], q3[l], t[2], t2[2], t3[2]);
		}
#pragma unroll
		for (int l = 0; l < 3; l++)
			vx[a][l] = cubic_f32(vy[0][l], vy[1][l], vy[2][l], vy[3][l], t[1], t2[1], t3[1]);
	}
	float b[3];
#pragma unroll
	for (int l = 0; l < 3; l++)
		b[l] = cubic_f32(vx[0][l], vx[1][l], vx[2][l], vx[3][l], t[0], t2[0], t3[0]);
	return v3((double)b[0], (double)b[1], (double)b[2]);
}

// Grid<float>::tricubicInterpolate(double, ...) (include/crpropa/Grid.h:504-536): the same loops in
// FP64 over the float values; Grid::interpolate narrows the result to float
__device__ inline float scalar_grid_tricubic(const GridDesc &g, const V3 &p) {
	int i0[3];
	double f[3];
	cubic_origin(g, p, i0, f);
	int Z[4];
	size_t Y[4];
#pragma unroll
	for (int d = 0; d < 4; d++) {
		Z[d] = cubic_index(i0[2] - 1 + d, g.nz, g.invN[2], g.reflective);
		Y[d] = (size_t)cubic_index(i0[1] - 1 + d, g.ny, g.invN[1], g.reflective) * g.nz;
	}
	double vx[4];
#pragma unroll
	for (int a = 0; a < 4; a++) {
		const size_t X = (size_t)cubic_index(i0[0] - 1 + a, g.nx, g.invN[0], g.reflective) * g.ny * g.nz;
		double vy[4];
#pragma unroll
		for (int b = 0; b < 4; b++) {
			const float *q = g.values + X + Y[b];
			vy[b] = cubic_f64(q[Z[0]], q[Z[1]], q[Z[2]], q[Z[3]], f[2]);
		}
		vx[a] = cubic_f64(vy[0], vy[1], vy[2], vy[3], f[1]);
	}
	return (float)cubic_f64(vx[0], vx[1], vx[2], vx[3], f[0]);
}

// Grid::interpolate (Grid.h:329-332) for the tricubic and nearest-neighbour types, after the clip
// test: inline in the kernels of FieldGridAny (which read nothing else), out of line for the field
// types that read such grids beside others (JF12's turbulent grid, decorator leaves)
__device__ __forceinline__ V3 grid_eval_other_inline(const GridDesc &g, const V3 &p) {
	if (g.interp == kInterpNearest) {
		const float *c = g.values + 3 * grid_closest(g, p);
		return v3((double)c[0], (double)c[1], (double)c[2]);
	}
	return grid_tricubic(g, p);
}
__device__ __attribute__((noinline)) V3 grid_eval_other(const GridDesc &g, const V3 &p) {
	return grid_eval_other_inline(g, p);
}

__device__ __forceinline__ float scalar_grid_other_inline(const GridDesc &g, const V3 &p) {
	if (g.interp == kInterpNearest)
		return g.values[grid_closest(g, p)];
	return scalar_grid_tricubic(g, p);
}
__device__ __attribute__((noinline)) float scalar_grid_other(const GridDesc &g, const V3 &p) {
	return scalar_grid_other_inline(g, p);
}

struct FieldGrid {
	GridDesc g;

	__device__ __forceinline__ V3 eval(const V3 &p) const {
		if (g.clip) {
			bool in = (p.x >= g.origin[0]) && (p.x <= g.edge[0]);
			in &= (p.y >= g.origin[1]) && (p.y <= g.edge[1]);
			in &= (p.z >= g.origin[2]) && (p.z <= g.edge[2]);
			if (!in)
				return v3(0.0, 0.0, 0.0);
		}
		double rx = div_cr(p.x - g.gridOrigin[0], g.spacing[0], g.invSpacing[0], g.fastDiv);
		double ry = div_cr(p.y - g.gridOrigin[1], g.spacing[1], g.invSpacing[1], g.fastDiv);
		double rz = div_cr(p.z - g.gridOrigin[2], g.spacing[2], g.invSpacing[2], g.fastDiv);
		int x0, x1, y0, y1, z0, z1;
		double fx, fy, fz;
		if (g.reflective) {
			reflective_clamp(rx, g.nx, x0, x1, fx);
			reflective_clamp(ry, g.ny, y0, y1, fy);
			reflective_clamp(rz, g.nz, z0, z1, fz);
		} else {
			periodic_clamp(rx, g.nx, g.invN[0], x0, x1, fx);
			periodic_clamp(ry, g.ny, g.invN[1], y0, y1, fy);
			periodic_clamp(rz, g.nz, g.invN[2], z0, z1, fz);
		}
		// weights narrowed to float (Vector3<float>::operator*(float))
		const float wx0 = (float)fx, wy0 = (float)fy, wz0 = (float)fz;
		const float wx1 = (float)(1 - fx), wy1 = (float)(1 - fy), wz1 = (float)(1 - fz);
		const size_t sy = (size_t)g.nz, sx = (size_t)g.ny * g.nz;
		const size_t bx0 = (size_t)x0 * sx, bx1 = (size_t)x1 * sx;
		const size_t by0 = (size_t)y0 * sy, by1 = (size_t)y1 * sy;
		const float *v = g.values;
		// corner order of Grid.h:572-579
		const float *c0 = v + 3 * (bx0 + by0 + z0);
		const float *c1 = v + 3 * (bx1 + by0 + z0);
		const float *c2 = v + 3 * (bx0 + by1 + z0);
		const float *c3 = v + 3 * (bx0 + by0 + z1);
		const float *c4 = v + 3 * (bx1 + by0 + z1);
		const float *c5 = v + 3 * (bx0 + by1 + z1);
		const float *c6 = v + 3 * (bx1 + by1 + z0);
		const float *c7 = v + 3 * (bx1 + by1 + z1);
		float b[3];
#pragma unroll
		for (int a = 0; a < 3; a++) {
			float s = 0.0f;
			s += ((c0[a] * wx1) * wy1) * wz1;
			s += ((c1[a] * wx0) * wy1) * wz1;
			s += ((c2[a] * wx1) * wy0) * wz1;
			s += ((c3[a] * wx1) * wy1) * wz0;
			s += ((c4[a] * wx0) * wy1) * wz0;
			s += ((c5[a] * wx1) * wy0) * wz0;
			s += ((c6[a] * wx0) * wy0) * wz1;
			s += ((c7[a] * wx0) * wy0) * wz0;
			b[a] = s;
		}
		return v3((double)b[0], (double)b[1], (double)b[2]);
	}

	// mixed precision: reference index arithmetic, FP32 weights and fused sum
	__device__ __forceinline__ float3 evalf(const V3 &p) const {
		if (g.clip) {
			bool in = (p.x >= g.origin[0]) && (p.x <= g.edge[0]);
			in &= (p.y >= g.origin[1]) && (p.y <= g.edge[1]);
			in &= (p.z >= g.origin[2]) && (p.z <= g.edge[2]);
			if (!in)
				return make_float3(0.f, 0.f, 0.f);
		}
		double rx = (p.x - g.gridOrigin[0]) * g.invSpacing[0];
		double ry = (p.y - g.gridOrigin[1]) * g.invSpacing[1];
		double rz = (p.z - g.gridOrigin[2]) * g.invSpacing[2];
		int x0, x1, y0, y1, z0, z1;
		double fx, fy, fz;
		if (g.reflective) {
			reflective_clamp(rx, g.nx, x0, x1, fx);
			reflective_clamp(ry, g.ny, y0, y1, fy);
			reflective_clamp(rz, g.nz, z0, z1, fz);
		} else {
			periodic_clamp(rx, g.nx, g.invN[0], x0, x1, fx);
			periodic_clamp(ry, g.ny, g.invN[1], y0, y1, fy);
			periodic_clamp(rz, g.nz, g.invN[2], z0, z1, fz);
		}
		const size_t sy = (size_t)g.nz, sx = (size_t)g.ny * g.nz;
		const size_t X[2] = {(size_t)x0 * sx, (size_t)x1 * sx}, Yo[2] = {(size_t)y0 * sy, (size_t)y1 * sy};
		const int Z[2] = {z0, z1};
		const int cx[8] = {0, 1, 0, 0, 1, 0, 1, 1}, cy[8] = {0, 0, 1, 0, 0, 1, 1, 1}, cz[8] = {0, 0, 0, 1, 1, 1, 0, 1};
		float v[24];
#pragma unroll
		for (int k = 0; k < 8; k++) {
			const float *c = g.values + 3 * (X[cx[k]] + Yo[cy[k]] + Z[cz[k]]);
			v[3 * k] = c[0];
			v[3 * k + 1] = c[1];
			v[3 * k + 2] = c[2];
		}
		return trilinear_f32(v, (float)fx, (float)fy, (float)fz);
	}
};

// Grid<float>::interpolate + trilinearInterpolate (Grid.h:318-335, 540-582) with T = float: each
// corner value times the double weights in double, and every b += term rounds the double sum to
// float.  `g.values` holds one float per point.
__device__ __forceinline__ float scalar_grid_eval(const GridDesc &g, const V3 &p) {
	if (g.clip) {
		bool in = (p.x >= g.origin[0]) && (p.x <= g.edge[0]);
		in &= (p.y >= g.origin[1]) && (p.y <= g.edge[1]);
		in &= (p.z >= g.origin[2]) && (p.z <= g.edge[2]);
		if (!in)
			return 0.f;
	}
	const double rx = div_cr(p.x - g.gridOrigin[0], g.spacing[0], g.invSpacing[0], g.fastDiv);
	const double ry = div_cr(p.y - g.gridOrigin[1], g.spacing[1], g.invSpacing[1], g.fastDiv);
	const double rz = div_cr(p.z - g.gridOrigin[2], g.spacing[2], g.invSpacing[2], g.fastDiv);
	int x0, x1, y0, y1, z0, z1;
	double fx, fy, fz;
	if (g.reflective) {
		reflective_clamp(rx, g.nx, x0, x1, fx);
		reflective_clamp(ry, g.ny, y0, y1, fy);
		reflective_clamp(rz, g.nz, z0, z1, fz);
	} else {
		periodic_clamp(rx, g.nx, g.invN[0], x0, x1, fx);
		periodic_clamp(ry, g.ny, g.invN[1], y0, y1, fy);
		periodic_clamp(rz, g.nz, g.invN[2], z0, z1, fz);
	}
	const double gx = 1 - fx, gy = 1 - fy, gz = 1 - fz;
	const size_t sy = (size_t)g.nz, sx = (size_t)g.ny * g.nz;
	const size_t bx0 = (size_t)x0 * sx, bx1 = (size_t)x1 * sx, by0 = (size_t)y0 * sy, by1 = (size_t)y1 * sy;
	const float *v = g.values;
	float b = 0.f;
	b = (float)((double)b + (((double)v[bx0 + by0 + z0] * gx) * gy) * gz);
	b = (float)((double)b + (((double)v[bx1 + by0 + z0] * fx) * gy) * gz);
	b = (float)((double)b + (((double)v[bx0 + by1 + z0] * gx) * fy) * gz);
	b = (float)((double)b + (((double)v[bx0 + by0 + z1] * gx) * gy) * fz);
	b = (float)((double)b + (((double)v[bx1 + by0 + z1] * fx) * gy) * fz);
	b = (float)((double)b + (((double)v[bx0 + by1 + z1] * gx) * fy) * fz);
	b = (float)((double)b + (((double)v[bx1 + by1 + z0] * fx) * fy) * gz);
	b = (float)((double)b + (((double)v[bx1 + by1 + z1] * fx) * fy) * fz);
	return b;
}

// mixed precision: reciprocal spacing, FP32 weights, fused sum
__device__ __forceinline__ float scalar_grid_evalf(const GridDesc &g, const V3 &p) {
	if (g.clip) {
		bool in = (p.x >= g.origin[0]) && (p.x <= g.edge[0]);
		in &= (p.y >= g.origin[1]) && (p.y <= g.edge[1]);
		in &= (p.z >= g.origin[2]) && (p.z <= g.edge[2]);
		if (!in)
			return 0.f;
	}
	const double rx = (p.x - g.gridOrigin[0]) * g.invSpacing[0];
	const double ry = (p.y - g.gridOrigin[1]) * g.invSpacing[1];
	const double rz = (p.z - g.gridOrigin[2]) * g.invSpacing[2];
	int x0, x1, y0, y1, z0, z1;
	double fx, fy, fz;
	if (g.reflective) {
		reflective_clamp(rx, g.nx, x0, x1, fx);
		reflective_clamp(ry, g.ny, y0, y1, fy);
		reflective_clamp(rz, g.nz, z0, z1, fz);
	} else {
		periodic_clamp(rx, g.nx, g.invN[0], x0, x1, fx);
		periodic_clamp(ry, g.ny, g.invN[1], y0, y1, fy);
		periodic_clamp(rz, g.nz, g.invN[2], z0, z1, fz);
	}
	const float wx = (float)fx, wy = (float)fy, wz = (float)fz, vx = 1.f - wx, vy = 1.f - wy, vz = 1.f - wz;
	const size_t sy = (size_t)g.nz, sx = (size_t)g.ny * g.nz;
	const size_t bx0 = (size_t)x0 * sx, bx1 = (size_t)x1 * sx, by0 = (size_t)y0 * sy, by1 = (size_t)y1 * sy;
	const float *v = g.values;
	float b = v[bx0 + by0 + z0] * (vx * vy * vz);
	b = __builtin_fmaf(v[bx1 + by0 + z0], wx * vy * vz, b);
	b = __builtin_fmaf(v[bx0 + by1 + z0], vx * wy * vz, b);
	b = __builtin_fmaf(v[bx0 + by0 + z1], vx * vy * wz, b);
	b = __builtin_fmaf(v[bx1 + by0 + z1], wx * vy * wz, b);
	b = __builtin_fmaf(v[bx0 + by1 + z1], vx * wy * wz, b);
	b = __builtin_fmaf(v[bx1 + by1 + z0], wx * wy * vz, b);
	b = __builtin_fmaf(v[bx1 + by1 + z1], wx * wy * wz, b);
	return b;
}

// Periodic Grid3f read through a cell-major corner table: cell (i,j,k) holds
// the 8 corner values of the reference's corner order (Grid.h:572-579) as 24
// floats, padded to 128 bytes, so one evaluation reads one aligned line
// instead of eight scattered 12-byte points.  Same values, same arithmetic
// as FieldGrid (bit-identical results).
constexpr int kCellFloats = 32;

struct FieldGridCells {
	GridDesc g;
	const float *cells; // [nx*ny*nz][32]

	__device__ __forceinline__ V3 eval(const V3 &p) const {
		if (g.clip) {
			bool in = (p.x >= g.origin[0]) && (p.x <= g.edge[0]);
			in &= (p.y >= g.origin[1]) && (p.y <= g.edge[1]);
			in &= (p.z >= g.origin[2]) && (p.z <= g.edge[2]);
			if (!in)
				return v3(0.0, 0.0, 0.0);
		}
		double rx = div_cr(p.x - g.gridOrigin[0], g.spacing[0], g.invSpacing[0], g.fastDiv);
		double ry = div_cr(p.y - g.gridOrigin[1], g.spacing[1], g.invSpacing[1], g.fastDiv);
		double rz = div_cr(p.z - g.gridOrigin[2], g.spacing[2], g.invSpacing[2], g.fastDiv);
		int x0, x1, y0, y1, z0, z1;
		double fx, fy, fz;
		periodic_clamp(rx, g.nx, g.invN[0], x0, x1, fx);
		periodic_clamp(ry, g.ny, g.invN[1], y0, y1, fy);
		periodic_clamp(rz, g.nz, g.invN[2], z0, z1, fz);
		const float wx0 = (float)fx, wy0 = (float)fy, wz0 = (float)fz;
		const float wx1 = (float)(1 - fx), wy1 = (float)(1 - fy), wz1 = (float)(1 - fz);
		const size_t cidx = ((size_t)x0 * g.ny + y0) * g.nz + z0;
		const float4 *c = reinterpret_cast<const float4 *>(cells + kCellFloats * cidx);
		const float4 q0 = c[0], q1 = c[1], q2 = c[2], q3 = c[3], q4 = c[4], q5 = c[5];
		const float v[24] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w,
		                     q3.x, q3.y, q3.z, q3.w, q4.x, q4.y, q4.z, q4.w, q5.x, q5.y, q5.z, q5.w};
		float b[3];
#pragma unroll
		for (int a = 0; a < 3; a++) {
			float s = 0.0f;
			s += ((v[0 + a] * wx1) * wy1) * wz1;
			s += ((v[3 + a] * wx0) * wy1) * wz1;
			s += ((v[6 + a] * wx1) * wy0) * wz1;
			s += ((v[9 + a] * wx1) * wy1) * wz0;
			s += ((v[12 + a] * wx0) * wy1) * wz0;
			s += ((v[15 + a] * wx1) * wy0) * wz0;
			s += ((v[18 + a] * wx0) * wy0) * wz1;
			s += ((v[21 + a] * wx0) * wy0) * wz0;
			b[a] = s;
		}
		return v3((double)b[0], (double)b[1], (double)b[2]);
	}

	// mixed precision: reciprocal spacing, division-free periodic wrap, fused FP32 sum
	__device__ __forceinline__ float3 evalf(const V3 &p) const {
		if (g.clip) {
			bool in = (p.x >= g.origin[0]) && (p.x <= g.edge[0]);
			in &= (p.y >= g.origin[1]) && (p.y <= g.edge[1]);
			in &= (p.z >= g.origin[2]) && (p.z <= g.edge[2]);
			if (!in)
				return make_float3(0.f, 0.f, 0.f);
		}
		int x0, y0, z0;
		float fx, fy, fz;
		periodic_cell_fast((p.x - g.gridOrigin[0]) * g.invSpacing[0], g.nx, 1.0 / g.nx, x0, fx);
		periodic_cell_fast((p.y - g.gridOrigin[1]) * g.invSpacing[1], g.ny, 1.0 / g.ny, y0, fy);
		periodic_cell_fast((p.z - g.gridOrigin[2]) * g.invSpacing[2], g.nz, 1.0 / g.nz, z0, fz);
		const size_t cidx = ((size_t)x0 * g.ny + y0) * g.nz + z0;
		const float4 *c = reinterpret_cast<const float4 *>(cells + kCellFloats * cidx);
		const float4 q0 = c[0], q1 = c[1], q2 = c[2], q3 = c[3], q4 = c[4], q5 = c[5];
		const float v[24] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w,
		                     q3.x, q3.y, q3.z, q3.w, q4.x, q4.y, q4.z, q4.w, q5.x, q5.y, q5.z, q5.w};
		return trilinear_f32(v, fx, fy, fz);
	}
};

// Reduced storage for the mixed path: each periodic cell holds its 8 corner
// vectors as FP16 values normalised by the FP32 scale of the 8^3-point brick
// that contains the cell (max |component| over the brick and its +1 faces),
// followed by that scale: 48 + 4 bytes, padded to one 64-byte sector.
constexpr int kCell16Bytes = 64;

struct FieldGridCells16 {
	GridDesc g;
	const unsigned char *cells16; // [nx*ny*nz][64]

	__device__ __forceinline__ V3 eval(const V3 &p) const {
		float3 b = evalf(p);
		return v3(b.x, b.y, b.z);
	}

	__device__ __forceinline__ float3 evalf(const V3 &p) const {
		if (g.clip) {
			bool in = (p.x >= g.origin[0]) && (p.x <= g.edge[0]);
			in &= (p.y >= g.origin[1]) && (p.y <= g.edge[1]);
			in &= (p.z >= g.origin[2]) && (p.z <= g.edge[2]);
			if (!in)
				return make_float3(0.f, 0.f, 0.f);
		}
		int x0, y0, z0;
		float fx, fy, fz;
		periodic_cell_fast((p.x - g.gridOrigin[0]) * g.invSpacing[0], g.nx, 1.0 / g.nx, x0, fx);
		periodic_cell_fast((p.y - g.gridOrigin[1]) * g.invSpacing[1], g.ny, 1.0 / g.ny, y0, fy);
		periodic_cell_fast((p.z - g.gridOrigin[2]) * g.invSpacing[2], g.nz, 1.0 / g.nz, z0, fz);
		const size_t cidx = ((size_t)x0 * g.ny + y0) * g.nz + z0;
		const uint4 *c = reinterpret_cast<const uint4 *>(cells16 + (size_t)kCell16Bytes * cidx);
		const uint4 a0 = c[0], a1 = c[1], a2 = c[2], a3 = c[3];
		const unsigned int w[13] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w, a2.x, a2.y, a2.z, a2.w, a3.x};
		float v[24];
#pragma unroll
		for (int k = 0; k < 12; k++) {
			const unsigned int u = w[k];
			v[2 * k] = __half2float(__ushort_as_half((unsigned short)(u & 0xffffu)));
			v[2 * k + 1] = __half2float(__ushort_as_half((unsigned short)(u >> 16)));
		}
		const float scale = __uint_as_float(w[12]);
		float3 b = trilinear_f32(v, fx, fy, fz);
		return make_float3(b.x * scale, b.y * scale, b.z * scale);
	}
};

// ModulatedMagneticFieldGrid::getField (src/magneticField/MagneticFieldGrid.cpp:51-55): the Grid3f
// value widened to double times the Grid1f modulation widened to double
template <class G> struct FieldModulated {
	G grid;
	GridDesc mod;
	__device__ __forceinline__ V3 eval(const V3 &p) const {
		const double m = (double)scalar_grid_eval(mod, p);
		const V3 b = grid.eval(p);
		return v3(b.x * m, b.y * m, b.z * m);
	}
	__device__ __forceinline__ float3 evalf(const V3 &p) const {
		const float m = scalar_grid_evalf(mod, p);
		const float3 b = grid.evalf(p);
		return make_float3(b.x * m, b.y * m, b.z * m);
	}
};

__device__ __forceinline__ bool in_clip_box(const GridDesc &g, const V3 &p) {
	bool in = (p.x >= g.origin[0]) && (p.x <= g.edge[0]);
	in &= (p.y >= g.origin[1]) && (p.y <= g.edge[1]);
	in &= (p.z >= g.origin[2]) && (p.z <= g.edge[2]);
	return in;
}

// Grid::interpolate (include/crpropa/Grid.h:318-335) for every interpolation type: trilinear in
// line, tricubic and nearest neighbour after the clip test (in line when Inline, else out of line)
template <bool Inline = false>
__device__ __forceinline__ V3 grid_interpolate(const GridDesc &g, const V3 &p) {
	if (g.interp == kInterpTrilinear)
		return FieldGrid{g}.eval(p);
	if (g.clip && !in_clip_box(g, p))
		return v3(0.0, 0.0, 0.0);
	return Inline ? grid_eval_other_inline(g, p) : grid_eval_other(g, p);
}

template <bool Inline = false>
__device__ __forceinline__ float scalar_grid_interpolate(const GridDesc &g, const V3 &p) {
	if (g.interp == kInterpTrilinear)
		return scalar_grid_eval(g, p);
	if (g.clip && !in_clip_box(g, p))
		return 0.f;
	return Inline ? scalar_grid_other_inline(g, p) : scalar_grid_other(g, p);
}

// MagneticFieldGrid and ModulatedMagneticFieldGrid (src/magneticField/MagneticFieldGrid.cpp:5-55)
// whose Grid3f or Grid1f is read tricubic or nearest-neighbour (Grid::setInterpolationType).  Its own
// field type, so that the trilinear kernels keep their register allocation.  The mixed path reads
// trilinear grids as FieldGrid::evalf does and the other types with the FP64 path's code, whose
// result is a float already.
struct FieldGridAny {
	GridDesc g;
	GridDesc mod; // the Grid1f modulation when `modulated`
	int modulated;

	__device__ __forceinline__ V3 eval(const V3 &p) const {
		const V3 b = grid_interpolate<true>(g, p);
		if (!modulated)
			return b;
		const double m = (double)scalar_grid_interpolate<true>(mod, p);
		return v3(b.x * m, b.y * m, b.z * m);
	}
	__device__ __forceinline__ float3 evalf(const V3 &p) const {
		float3 b;
		if (g.interp == kInterpTrilinear) {
			b = FieldGrid{g}.evalf(p);
		} else {
			const V3 d = grid_interpolate<true>(g, p);
			b = make_float3((float)d.x, (float)d.y, (float)d.z);
		}
		if (!modulated)
			return b;
		const float m = mod.interp == kInterpTrilinear ? scalar_grid_evalf(mod, p) : scalar_grid_interpolate<true>(mod, p);
		return make_float3(b.x * m, b.y * m, b.z * m);
	}
};

struct FieldPlaneWave {
	PlaneWaveDesc w;

	// B = sum_i (xi_i A_i) cos(k_i (x . kappa_i) + beta_i), accumulated in mode order
	// (PlaneWaveTurbulence.cpp:231-240), with the correctly rounded cos
	__device__ __forceinline__ V3 eval(const V3 &p) const {
		double bx = 0.0, by = 0.0, bz = 0.0;
		PwMode64Const *m = (PwMode64Const *)w.m64;
		const int nm = w.nm;
#pragma unroll 2
		for (int i = 0; i < nm; i++) {
			const double kx = m[i].kx, ky = m[i].ky, kz = m[i].kz;
			double z_ = (p.x * kx + p.y * ky) + p.z * kz;
			double c = crp_cos_cr(m[i].k * z_ + m[i].beta);
			bx += m[i].ax * c;
			by += m[i].ay * c;
			bz += m[i].az * c;
		}
		return v3(bx, by, bz);
	}

	// mixed precision: phase in turns in FP64 (positions span many wavelengths), its fractional part
	// evaluated with the hardware cosine (v_cos_f32 computes cos(2 pi x)); amplitudes and the sum in FP32
	__device__ __forceinline__ float3 evalf(const V3 &p) const {
		float bx = 0.f, by = 0.f, bz = 0.f;
		PwModeMixedConst *m = (PwModeMixedConst *)w.mm;
		const int nm = w.nm;
#pragma unroll 4
		for (int i = 0; i < nm; i++) {
			const double t = __builtin_fma(p.x, m[i].ktx, __builtin_fma(p.y, m[i].kty, __builtin_fma(p.z, m[i].ktz, m[i].bt)));
			const float c = __builtin_amdgcn_cosf((float)__builtin_amdgcn_fract(t));
			bx = __builtin_fmaf(m[i].ax, c, bx);
			by = __builtin_fmaf(m[i].ay, c, by);
			bz = __builtin_fmaf(m[i].az, c, bz);
		}
		return make_float3(bx, by, bz);
	}
};

// ---------------------------------------------------------------- mixed-precision steps
__device__ __forceinline__ float3 f3(float x, float y, float z) { return make_float3(x, y, z); }
__device__ __forceinline__ float3 cross3(const float3 &a, const float3 &b) {
	return f3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
__device__ __forceinline__ float dot3(const float3 &a, const float3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 unit3(const float3 &a) {
	const float r = __builtin_amdgcn_rsqf(dot3(a, a));
	return f3(a.x * r, a.y * r, a.z * r);
}

// Boris step with FP32 field and rotation, FP64 position drift.  coef =
// (q/2/m) * step / c formed in FP64 by the caller (t = B * coef).
template <class F>
__device__ __forceinline__ void bp_dy_mixed(const F &field, V3 &pos, float3 &dir, double step, float coef) {
	const double hs = step * 0.5;
	pos = v3(pos.x + (double)dir.x * hs, pos.y + (double)dir.y * hs, pos.z + (double)dir.z * hs);
	const float3 B = field.evalf(pos);
	const float3 t = f3(B.x * coef, B.y * coef, B.z * coef);
	const float sf = 2.f / (1.f + dot3(t, t));
	const float3 s = f3(t.x * sf, t.y * sf, t.z * sf);
	const float3 c1 = cross3(dir, t);
	const float3 vh = f3(dir.x + c1.x, dir.y + c1.y, dir.z + c1.z);
	const float3 c2 = cross3(vh, s);
	dir = f3(dir.x + c2.x, dir.y + c2.y, dir.z + c2.z);
	pos = v3(pos.x + (double)dir.x * hs, pos.y + (double)dir.y * hs, pos.z + (double)dir.z * hs);
}

// Cash-Karp trial with FP32 direction stages and field, FP64 positions and
// FP64 error accumulation.  Q = q c^2 / E (per particle); h = step / c.
struct YM {
	V3 x;
	float3 u;
};

// The field at a Cash-Karp trial's start.  A rejected trial is retried from the same state with a
// smaller step (src/module/PropagationCK.cpp:99-126), so the retry's first stage reads the same field
// value again; the adaptive kernels keep it for the retries instead of reading the grid (same bits).
// CkB0Store<true> serves fields whose FP64 evaluation is an FP32 value widened (the trilinear Grid3f
// readers): it keeps the value in three floats, three registers fewer; CkB0f is the mixed path's.
template <class F> struct FloatValuedField {
	static constexpr bool value = false;
};
template <> struct FloatValuedField<FieldGridCells> {
	static constexpr bool value = true;
};
template <> struct FloatValuedField<FieldGrid> {
	static constexpr bool value = true;
};
template <bool F32> struct CkB0Store;
template <> struct CkB0Store<false> {
	V3 B;
	bool have;
	__device__ __forceinline__ V3 get() const { return B; }
	__device__ __forceinline__ void put(const V3 &b) { B = b; have = true; }
};
template <> struct CkB0Store<true> {
	float3 B;
	bool have;
	__device__ __forceinline__ V3 get() const { return v3((double)B.x, (double)B.y, (double)B.z); }
	__device__ __forceinline__ void put(const V3 &b) {
		B = make_float3((float)b.x, (float)b.y, (float)b.z);
		have = true;
	}
};
struct CkB0f {
	float3 B;
	bool have;
};

template <class F>
__device__ __forceinline__ void ck_dydt_mixed(const F &field, const YM &y, float Q, float3 &n, float3 &du,
                                              CkB0f *b0 = nullptr) {
	n = unit3(y.u);                 // velocity / c
	float3 B;
	if (b0 && b0->have) {
		B = b0->B;
	} else {
		B = field.evalf(y.x);
		if (b0) {
			b0->B = B;
			b0->have = true;
		}
	}
	const float3 c = cross3(n, B);
	du = f3(c.x * Q, c.y * Q, c.z * Q);
}

template <class F>
__device__ __forceinline__ void ck_trystep_mixed(const F &field, const YM &y, double h, float Q, YM &out, double &errU,
                                                 CkB0f *b0 = nullptr) {
	constexpr double cA[6][5] = {{0, 0, 0, 0, 0},
	                             {CK::a10, 0, 0, 0, 0},
	                             {CK::a20, CK::a21, 0, 0, 0},
	                             {CK::a30, CK::a31, CK::a32, 0, 0},
	                             {CK::a40, CK::a41, CK::a42, CK::a43, 0},
	                             {CK::a50, CK::a51, CK::a52, CK::a53, CK::a54}};
	constexpr double cB[6] = {CK::b0, CK::b1, CK::b2, CK::b3, CK::b4, CK::b5};
	constexpr double cE[6] = {CK::b0 - CK::s0, CK::b1 - CK::s1, CK::b2 - CK::s2,
	                          CK::b3 - CK::s3, CK::b4 - CK::s4, CK::b5 - CK::s5};
	const double ch = kCLight * h; // position step per unit of velocity/c
	const float hf = (float)h;
	float3 kn[6], ku[6];
	double ex = 0, ey = 0, ez = 0;
	out = y;
#pragma unroll
	for (int i = 0; i < 6; i++) {
		YM yn = y;
#pragma unroll
		for (int j = 0; j < i; j++) {
			const double ax = cA[i][j] * ch;
			const float au = (float)cA[i][j] * hf;
			yn.x = v3(yn.x.x + (double)kn[j].x * ax, yn.x.y + (double)kn[j].y * ax, yn.x.z + (double)kn[j].z * ax);
			yn.u = f3(__builtin_fmaf(ku[j].x, au, yn.u.x), __builtin_fmaf(ku[j].y, au, yn.u.y),
			          __builtin_fmaf(ku[j].z, au, yn.u.z));
		}
		ck_dydt_mixed(field, yn, Q, kn[i], ku[i], i == 0 ? b0 : nullptr);
		if (cB[i] != 0.0) {
			const double bx = cB[i] * ch;
			const float bu = (float)cB[i] * hf;
			out.x = v3(out.x.x + (double)kn[i].x * bx, out.x.y + (double)kn[i].y * bx, out.x.z + (double)kn[i].z * bx);
			out.u = f3(__builtin_fmaf(ku[i].x, bu, out.u.x), __builtin_fmaf(ku[i].y, bu, out.u.y),
			           __builtin_fmaf(ku[i].z, bu, out.u.z));
		}
		const double eh = cE[i] * h;
		ex += (double)ku[i].x * eh;
		ey += (double)ku[i].y * eh;
		ez += (double)ku[i].z * eh;
	}
	errU = __builtin_sqrt((ex * ex + ey * ey) + ez * ez);
}

// ---------------------------------------------------------------- Boris push
struct Y {
	V3 x, u;
};

// PropagationBP::dY (src/module/PropagationBP.cpp:19-39)
template <class F>
// (divisions as div_rn with shared reciprocals: invm = RN(1/m) from the caller)
__device__ __forceinline__ Y bp_dy(const F &field, const V3 &pos0, const V3 &dir, double step, double q, double m,
                                  double invm) {
	V3 pos = v3(pos0.x + (dir.x * step) / 2., pos0.y + (dir.y * step) / 2., pos0.z + (dir.z * step) / 2.);
	V3 B = field.eval(pos);
	V3 t = v3(div_rn(div_rn((B.x * q) / 2, m, invm) * step, kCLight, kInvCLight),
	          div_rn(div_rn((B.y * q) / 2, m, invm) * step, kCLight, kInvCLight),
	          div_rn(div_rn((B.z * q) / 2, m, invm) * step, kCLight, kInvCLight));
	double tt = (t.x * t.x + t.y * t.y) + t.z * t.z;
	double den = 1 + tt;
	const double invden = 1 / den;
	V3 s = v3(div_rn(t.x * 2, den, invden), div_rn(t.y * 2, den, invden), div_rn(t.z * 2, den, invden));
	V3 vh = v3(dir.x + (dir.y * t.z - t.y * dir.z), dir.y + (dir.z * t.x - t.z * dir.x),
	           dir.z + (dir.x * t.y - t.x * dir.y));
	V3 d = v3(dir.x + (vh.y * s.z - s.y * vh.z), dir.y + (vh.z * s.x - s.z * vh.x),
	          dir.z + (vh.x * s.y - s.x * vh.y));
	Y out;
	out.x = v3(pos.x + (d.x * step) / 2., pos.y + (d.y * step) / 2., pos.z + (d.z * step) / 2.);
	out.u = d;
	return out;
}

// ---------------------------------------------------------------- Cash-Karp
// PropagationCK::dYdt (src/module/PropagationCK.cpp:51-61); qcE = (q*c)/E
template <class F>
__device__ __forceinline__ Y ck_dydt(const F &field, const Y &y, double qcE,
                                     CkB0Store<FloatValuedField<F>::value> *b0 = nullptr) {
	Y d;
	double n = norm(y.u);
	const double invn = 1 / n;
	d.x = v3(div_rn(y.u.x, n, invn) * kCLight, div_rn(y.u.y, n, invn) * kCLight, div_rn(y.u.z, n, invn) * kCLight);
	V3 B;
	if (b0 && b0->have) {
		B = b0->get();
	} else {
		B = field.eval(y.x);
		if (b0)
			b0->put(B);
	}
	d.u = v3((d.x.y * B.z - B.y * d.x.z) * qcE, (d.x.z * B.x - B.z * d.x.x) * qcE,
	         (d.x.x * B.y - B.x * d.x.y) * qcE);
	return d;
}

// y_n += (k * a) * h, component-wise over x and u
__device__ __forceinline__ void axpy(Y &yn, const Y &k, double a, double h) {
	yn.x.x += (k.x.x * a) * h;
	yn.x.y += (k.x.y * a) * h;
	yn.x.z += (k.x.z * a) * h;
	yn.u.x += (k.u.x * a) * h;
	yn.u.y += (k.u.y * a) * h;
	yn.u.z += (k.u.z * a) * h;
}

// PropagationCK::tryStep (src/module/PropagationCK.cpp:28-49).  Returns the
// new phase point in `out` and the error estimate in `err`.
template <class F>
__device__ __forceinline__ void ck_trystep_err(const F &field, const Y &y, double h, double qcE, Y &out, Y &err,
                                               CkB0Store<FloatValuedField<F>::value> *b0 = nullptr) {
	err.x = v3(0.0, 0.0, 0.0);
	err.u = v3(0.0, 0.0, 0.0);
	out = y;
	const double e0 = CK::b0 - CK::s0, e1 = CK::b1 - CK::s1, e2 = CK::b2 - CK::s2, e3 = CK::b3 - CK::s3,
	             e4 = CK::b4 - CK::s4, e5 = CK::b5 - CK::s5;
	Y k0, k1, k2, k3, k4, k5, yn;
	k0 = ck_dydt(field, y, qcE, b0);
	axpy(out, k0, CK::b0, h);
	axpy(err, k0, e0, h);
	yn = y;
	axpy(yn, k0, CK::a10, h);
	k1 = ck_dydt(field, yn, qcE);
	axpy(out, k1, CK::b1, h);
	axpy(err, k1, e1, h);
	yn = y;
	axpy(yn, k0, CK::a20, h);
	axpy(yn, k1, CK::a21, h);
	k2 = ck_dydt(field, yn, qcE);
	axpy(out, k2, CK::b2, h);
	axpy(err, k2, e2, h);
	yn = y;
	axpy(yn, k0, CK::a30, h);
	axpy(yn, k1, CK::a31, h);
	axpy(yn, k2, CK::a32, h);
	k3 = ck_dydt(field, yn, qcE);
	axpy(out, k3, CK::b3, h);
	axpy(err, k3, e3, h);
	yn = y;
	axpy(yn, k0, CK::a40, h);
	axpy(yn, k1, CK::a41, h);
	axpy(yn, k2, CK::a42, h);
	axpy(yn, k3, CK::a43, h);
	k4 = ck_dydt(field, yn, qcE);
	axpy(out, k4, CK::b4, h);
	axpy(err, k4, e4, h);
	yn = y;
	axpy(yn, k0, CK::a50, h);
	axpy(yn, k1, CK::a51, h);
	axpy(yn, k2, CK::a52, h);
	axpy(yn, k3, CK::a53, h);
	axpy(yn, k4, CK::a54, h);
	k5 = ck_dydt(field, yn, qcE);
	axpy(out, k5, CK::b5, h);
	axpy(err, k5, e5, h);
}

// the same with only |err.u|, the quantity the step control reads
template <class F>
__device__ __forceinline__ void ck_trystep(const F &field, const Y &y, double h, double qcE, Y &out, double &errU,
                                           CkB0Store<FloatValuedField<F>::value> *b0 = nullptr) {
	Y err;
	ck_trystep_err(field, y, h, qcE, out, err, b0);
	errU = norm(err.u);
}

} // namespace crp
