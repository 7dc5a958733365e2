// Device kernels of the turbulent-grid generator: the k-space coefficients of
// SimpleGridTurbulence::initTurbulence
// (src/magneticField/turbulentField/SimpleGridTurbulence.cpp:18-113) and
// GridTurbulence::initTurbulence (GridTurbulence.cpp:29-124) from the caller's
// MT19937 draws, and the layout steps around the complex-to-real inverse FFT
// (GridTurbulence.cpp:147-185).
//
// Modes are visited in the reference's order (ix, iy, iz < n/2 + 1, iz
// fastest); a mode inside [spacing/lMax, spacing/lMin] consumes a fixed
// number of draws (4 for the simple spectrum: rand, randNorm = randDblExc +
// randExc, rand; 2 for the general one: rand, rand), so the draws of mode m
// start at draws-per-mode x (inside modes before m).  One thread owns one
// (ix, iy) row and walks it in order; the row's first inside-mode index comes
// from an exclusive scan of the per-row counts.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace crp {

enum TurbKind { TURB_SIMPLE = 0, TURB_GENERAL = 1, TURB_GENERAL_POWERLAW = 2 };

struct TurbDesc {
	int n;
	int n2;           // n/2 + 1 complex points along z
	int kind;
	double kMin, kMax; // spacing/lMax, spacing/lMin
	double alpha;      // simple: amplitude randNorm() * k^(alpha/2)
	double lambda;     // general: E is evaluated at k * lambda, lambda = 1/spacing * 2 pi
	double lbo, sidx, qidx;
};

// the discrete wave number of index i: i/n - i/(n/2) (integer division), narrowed to float
// as Vector3f::setXYZ does
__device__ __forceinline__ float turb_K(int i, int n) { return (float)((double)i / n - (double)(i / (n / 2))); }

// |ek| in float (Vector3f::getR) and the range test in double (SimpleGridTurbulence.cpp:55-60)
__device__ __forceinline__ bool turb_inside(const TurbDesc &D, float kx, float ky, float kz, double &k) {
	k = (double)__fsqrt_rn(__fadd_rn(__fadd_rn(__fmul_rn(kx, kx), __fmul_rn(ky, ky)), __fmul_rn(kz, kz)));
	return !((k < D.kMin) || (k > D.kMax));
}

__global__ void k_turb_row_count(TurbDesc D, uint32_t *rowCount) {
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= D.n * D.n)
		return;
	const float kx = turb_K(r / D.n, D.n), ky = turb_K(r % D.n, D.n);
	uint32_t c = 0;
	for (int iz = 0; iz < D.n2; iz++) {
		double k;
		c += turb_inside(D, kx, ky, turb_K(iz, D.n), k) ? 1u : 0u;
	}
	rowCount[r] = c;
}

// Vector3f::getAngleTo(1,1,1) < 1e-3 (include/crpropa/Vector3.h:172-187), in float
__device__ __forceinline__ bool turb_parallel_111(float x, float y, float z) {
	const float dot = __fadd_rn(__fadd_rn(x, y), z);
	const float r3 = __fsqrt_rn(3.f);
	const float r = __fsqrt_rn(__fadd_rn(__fadd_rn(__fmul_rn(x, x), __fmul_rn(y, y)), __fmul_rn(z, z)));
	const float c = __fdiv_rn(__fdiv_rn(dot, r3), r);
	float ang;
	if (c >= 1.)
		ang = 0.f;
	else if (c <= -1.)
		ang = (float)M_PI;
	else
		ang = acosf(c);
	return ang < 1e-3f;
}

// Random::rand / randExc / randDblExc (src/Random.cpp:80-98) of one draw
__device__ __forceinline__ double turb_rand(uint32_t u) { return (double)u * (1.0 / 4294967295.0); }
__device__ __forceinline__ double turb_randexc(uint32_t u) { return (double)u * (1.0 / 4294967296.0); }
__device__ __forceinline__ double turb_randdblexc(uint32_t u) { return ((double)u + 0.5) * (1.0 / 4294967296.0); }

// B holds 3 components x n x n x n2 complex floats (re, im)
__global__ void k_turb_coeffs(TurbDesc D, const uint64_t *rowOffset, const uint32_t *__restrict__ draws,
                              float2 *B) {
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= D.n * D.n)
		return;
	const size_t plane = (size_t)D.n * D.n * D.n2;
	const int perMode = D.kind == TURB_SIMPLE ? 4 : 2;
	const float kx = turb_K(r / D.n, D.n), ky = turb_K(r % D.n, D.n);
	uint64_t m = rowOffset[r];
	for (int iz = 0; iz < D.n2; iz++) {
		const size_t i = (size_t)r * D.n2 + iz;
		const float kz = turb_K(iz, D.n);
		double k;
		if (!turb_inside(D, kx, ky, kz, k)) {
			B[i] = make_float2(0.f, 0.f);
			B[plane + i] = make_float2(0.f, 0.f);
			B[2 * plane + i] = make_float2(0.f, 0.f);
			continue;
		}
		const uint32_t *u = draws + m * perMode;
		m++;
		// orthogonal base (SimpleGridTurbulence.cpp:63-76), float arithmetic
		float e1x, e1y, e1z, e2x, e2y, e2z;
		if (turb_parallel_111(kx, ky, kz)) {
			e1x = -1.f; e1y = 1.f; e1z = 0.f;
			e2x = 1.f; e2y = 1.f; e2z = -2.f;
		} else {
			// e1 = (1,1,1) x ek, e2 = ek x e1 (Vector3::cross, include/crpropa/Vector3.h:268-271)
			e1x = __fsub_rn(kz, ky);
			e1y = __fsub_rn(kx, kz);
			e1z = __fsub_rn(ky, kx);
			e2x = __fsub_rn(__fmul_rn(ky, e1z), __fmul_rn(e1y, kz));
			e2y = __fsub_rn(__fmul_rn(kz, e1x), __fmul_rn(e1z, kx));
			e2z = __fsub_rn(__fmul_rn(kx, e1y), __fmul_rn(e1x, ky));
		}
		const float r1 = __fsqrt_rn(__fadd_rn(__fadd_rn(__fmul_rn(e1x, e1x), __fmul_rn(e1y, e1y)), __fmul_rn(e1z, e1z)));
		const float r2 = __fsqrt_rn(__fadd_rn(__fadd_rn(__fmul_rn(e2x, e2x), __fmul_rn(e2y, e2y)), __fmul_rn(e2z, e2z)));
		e1x = __fdiv_rn(e1x, r1); e1y = __fdiv_rn(e1y, r1); e1z = __fdiv_rn(e1z, r1);
		e2x = __fdiv_rn(e2x, r2); e2y = __fdiv_rn(e2y, r2); e2z = __fdiv_rn(e2z, r2);
		// random orientation perpendicular to k: b = e1 cos(theta) + e2 sin(theta), the
		// double cos/sin narrowed to float (Vector3<float>::operator*(T))
		const double theta = 2 * M_PI * turb_rand(u[0]);
		const float ct = (float)::cos(theta), st = (float)::sin(theta);
		float bx = __fadd_rn(__fmul_rn(e1x, ct), __fmul_rn(e2x, st));
		float by = __fadd_rn(__fmul_rn(e1y, ct), __fmul_rn(e2y, st));
		float bz = __fadd_rn(__fmul_rn(e1z, ct), __fmul_rn(e2z, st));
		float amp;
		int next;
		if (D.kind == TURB_SIMPLE) {
			// Random::randNorm(0, 1) (src/Random.cpp:111-115) times k^(alpha/2)
			const double rr = ::sqrt(-2.0 * ::log(1.0 - turb_randdblexc(u[1]))) * 1.0;
			const double phi = 2.0 * 3.14159265358979323846264338328 * turb_randexc(u[2]);
			const double g = 0.0 + rr * ::cos(phi);
			amp = (float)(g * ::pow(k, D.alpha / 2));
			next = 3;
		} else if (D.kind == TURB_GENERAL_POWERLAW) {
			// SimpleTurbulenceSpectrum::energySpectrum (SimpleGridTurbulence.h:35-37)
			amp = (float)::sqrt(::pow(k * D.lambda, -D.sidx - 2));
			next = 1;
		} else {
			// TurbulenceSpectrum::energySpectrum (TurbulentField.h:75-80)
			const double kh = k * D.lambda * D.lbo;
			amp = (float)::sqrt(::pow(kh, D.qidx) / ::pow(1.0 + kh * kh, (D.sidx + D.qidx) / 2.0 + 1.0));
			next = 1;
		}
		bx = __fmul_rn(bx, amp);
		by = __fmul_rn(by, amp);
		bz = __fmul_rn(bz, amp);
		// uniform random phase; re = b cos(phase), im = b sin(phase) in double, stored as float
		const double phase = 2 * M_PI * turb_rand(u[next]);
		const double cp = ::cos(phase), sp = ::sin(phase);
		B[i] = make_float2((float)((double)bx * cp), (float)((double)bx * sp));
		B[plane + i] = make_float2((float)((double)by * cp), (float)((double)by * sp));
		B[2 * plane + i] = make_float2((float)((double)bz * cp), (float)((double)bz * sp));
	}
}

// The complex-to-real transform uses only the Hermitian part of the planes kz = 0 and, for even n,
// kz = n/2 (the imaginary parts of their DC and Nyquist terms drop out after the transforms along x
// and y).  Forming that part explicitly, X(kx,ky) <- (X(kx,ky) + conj X(-kx,-ky)) / 2, makes the
// result independent of how an FFT library treats non-Hermitian input.  One thread per pair.
__global__ void k_turb_hermitize(float2 *B, int n, int n2, int iz) {
	const int q = blockIdx.x * blockDim.x + threadIdx.x; // (ix, iy) in the plane
	if (q >= n * n)
		return;
	const int ix = q / n, iy = q % n;
	const int jx = (n - ix) % n, jy = (n - iy) % n;
	const int p = jx * n + jy;
	if (p < q)
		return; // the partner handles the pair
	const size_t plane = (size_t)n * n * n2;
	for (int c = 0; c < 3; c++) {
		float2 *a = B + c * plane + (size_t)q * n2 + iz;
		float2 *b = B + c * plane + (size_t)p * n2 + iz;
		if (p == q) {
			a->y = 0.f;
		} else {
			const float re = __fmul_rn(__fadd_rn(a->x, b->x), 0.5f);
			const float im = __fmul_rn(__fsub_rn(a->y, b->y), 0.5f);
			*a = make_float2(re, im);
			*b = make_float2(re, -im);
		}
	}
}

// real output of the in-place transform (padded rows of 2 n2 floats) -> Grid3f AoS (x, y, z)
// per point, index (ix n + iy) n + iz (GridTurbulence.cpp:174-183)
__global__ void k_turb_pack(const float *B, int n, int n2, float *grid) {
	const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	const size_t np = (size_t)n * n * n;
	if (p >= np)
		return;
	const size_t ix = p / ((size_t)n * n), iy = (p / n) % n, iz = p % n;
	const size_t src = ix * n * 2 * n2 + iy * 2 * n2 + iz;
	const size_t plane = (size_t)n * n * n2 * 2; // floats per component
	grid[3 * p] = B[src];
	grid[3 * p + 1] = B[plane + src];
	grid[3 * p + 2] = B[2 * plane + src];
}

} // namespace crp
