#define _GNU_SOURCE /* sincos */
/*
 * ORACLE -- test infrastructure, not product code (see crp_oracle.h).
 *
 * CPU restatement of the reference hot path.  Vector operations are written
 * out component by component in the reference's order of evaluation:
 *   Vector3<T>::getR   sqrt((x*x + y*y) + z*z)           include/crpropa/Vector3.h:126-128
 *   Vector3<T>::dot    (x*v.x + y*v.y) + z*v.z           include/crpropa/Vector3.h:263-265
 *   Vector3<T>::cross  (y*v.z - v.y*z, z*v.x - v.z*x, x*v.y - v.x*y)  include/crpropa/Vector3.h:268-271
 *   v * s, v / s       per component                     include/crpropa/Vector3.h:305-316
 * Build: oracle/Makefile (gcc -O3 -ffp-contract=off, no -march: no FMA).
 */
#include "crp_oracle.h"
#include "crp_crmath.h"
#include "crp_philox.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* c_light, include/crpropa/Units.h:64 */
static const double C_LIGHT = 2.99792458e8;

/* ------------------------------------------------------------------ libm */
double orc_cos(double x, int libm) { return libm == ORC_LIBM_CR ? crp_cos_cr(x) : cos(x); }
double orc_pow(double x, double y, int libm) { return libm == ORC_LIBM_CR ? crp_pow_cr(x, y) : pow(x, y); }
static double orc_exp(double x, int libm) { return libm == ORC_LIBM_CR ? crp_exp_rn(x) : exp(x); }
static double orc_atan2(double y, double x, int libm) { return libm == ORC_LIBM_CR ? crp_atan2_rn(y, x) : atan2(y, x); }

void orc_cos_batch(const double *x, size_t n, double *out, int libm) {
	for (size_t i = 0; i < n; i++)
		out[i] = orc_cos(x[i], libm);
}

void orc_pow_batch(const double *x, double y, size_t n, double *out, int libm) {
	for (size_t i = 0; i < n; i++)
		out[i] = orc_pow(x[i], y, libm);
}

/* ------------------------------------------------------------------ MT19937 (src/Random.cpp) */
#define ORC_N 624
#define ORC_M 397

/* Random::initialize, src/Random.cpp:370-379 */
static void orc_rng_initialize(orc_rng *r, uint32_t seed) {
	r->state[0] = seed;
	for (int i = 1; i < ORC_N; i++)
		r->state[i] = 1812433253U * (r->state[i - 1] ^ (r->state[i - 1] >> 30)) + (uint32_t)i;
}

/* Random::twist, include/crpropa/Random.h:214-225 */
static inline uint32_t orc_twist(uint32_t m, uint32_t s0, uint32_t s1) {
	uint32_t mix = (s0 & 0x80000000U) | (s1 & 0x7fffffffU);
	return m ^ (mix >> 1) ^ ((0U - (s1 & 1U)) & 0x9908b0dfU);
}

/* Random::reload, src/Random.cpp:381-391 */
static void orc_rng_reload(orc_rng *r) {
	uint32_t *p = r->state;
	int i;
	for (i = ORC_N - ORC_M; i--; ++p)
		*p = orc_twist(p[ORC_M], p[0], p[1]);
	for (i = ORC_M; --i; ++p)
		*p = orc_twist(p[ORC_M - ORC_N], p[0], p[1]);
	*p = orc_twist(p[ORC_M - ORC_N], p[0], r->state[0]);
	r->left = ORC_N;
	r->pNext = r->state;
}

/* Random::seed(uint32_t), src/Random.cpp:309-314 */
void orc_rng_seed(orc_rng *r, uint32_t seed) {
	orc_rng_initialize(r, seed);
	orc_rng_reload(r);
}

/* Random::randInt, src/Random.cpp:249-260 */
uint32_t orc_rng_randint(orc_rng *r) {
	if (r->left == 0)
		orc_rng_reload(r);
	--r->left;
	uint32_t s1 = *r->pNext++;
	s1 ^= (s1 >> 11);
	s1 ^= (s1 << 7) & 0x9d2c5680U;
	s1 ^= (s1 << 15) & 0xefc60000U;
	return s1 ^ (s1 >> 18);
}

/* Random::rand, src/Random.cpp:80-82 */
double orc_rng_rand(orc_rng *r) { return (double)orc_rng_randint(r) * (1.0 / 4294967295.0); }

/* Random::randExc / randDblExc, src/Random.cpp:88-98 */
static double orc_rng_randexc(orc_rng *r) { return (double)orc_rng_randint(r) * (1.0 / 4294967296.0); }
static double orc_rng_randdblexc(orc_rng *r) { return ((double)orc_rng_randint(r) + 0.5) * (1.0 / 4294967296.0); }

/* Random::randUniform, src/Random.cpp:118-120 */
double orc_rng_uniform(orc_rng *r, double min, double max) { return min + (max - min) * orc_rng_rand(r); }

/* Random::randNorm, src/Random.cpp:111-116 */
double orc_rng_norm(orc_rng *r, double mean, double variance) {
	double rr = sqrt(-2.0 * log(1.0 - orc_rng_randdblexc(r))) * variance;
	double phi = 2.0 * 3.14159265358979323846264338328 * orc_rng_randexc(r);
	return mean + rr * cos(phi);
}

/* Random::randVector, src/Random.cpp:141-146 */
void orc_rng_vector(orc_rng *r, double out[3]) {
	double z = orc_rng_uniform(r, -1.0, 1.0);
	double t = orc_rng_uniform(r, -1.0 * M_PI, M_PI);
	double rr = sqrt(1 - z * z);
	out[0] = rr * cos(t);
	out[1] = rr * sin(t);
	out[2] = z;
}

/* Random::randPowerLaw, src/Random.cpp:189-205 */
double orc_rng_powerlaw(orc_rng *r, double index, double min, double max) {
	if (fabs(index + 1.0) < 2.220446049250313e-16) {
		double part1 = log(max);
		double part2 = log(min);
		return exp((part1 - part2) * orc_rng_rand(r) + part2);
	}
	double part1 = pow(max, index + 1);
	double part2 = pow(min, index + 1);
	double ex = 1 / (index + 1);
	return pow((part1 - part2) * orc_rng_rand(r) + part2, ex);
}

/* Random::randBin(vector<double>), src/Random.cpp:135-139 (std::lower_bound) */
size_t orc_rng_bin(orc_rng *r, const double *cdf, size_t n) {
	double v = orc_rng_rand(r) * cdf[n - 1];
	size_t lo = 0, hi = n;
	while (lo < hi) {
		size_t mid = lo + (hi - lo) / 2;
		if (cdf[mid] < v)
			lo = mid + 1;
		else
			hi = mid;
	}
	return lo;
}

void orc_rng_randint_batch(uint32_t seed, size_t n, uint32_t *out) {
	orc_rng r;
	orc_rng_seed(&r, seed);
	for (size_t i = 0; i < n; i++)
		out[i] = orc_rng_randint(&r);
}

/* Noise grid used by the golden fixtures: per point (z fastest), components
 * x, y, z each randUniform(-1, 1) * amp, stored as float. */
void orc_noise_grid(uint32_t seed, int nx, int ny, int nz, double amp, float *out) {
	orc_rng r;
	orc_rng_seed(&r, seed);
	size_t n = (size_t)nx * ny * nz;
	for (size_t i = 0; i < 3 * n; i++)
		out[i] = (float)(orc_rng_uniform(&r, -1, 1) * amp);
}

/* Per particle: species (randBin over cumulative weights, when nw > 1),
 * energy randPowerLaw(-1, emin, emax), direction randVector -- the draw order
 * of the golden fixtures. */
void orc_sample_particles(uint32_t seed, size_t n, double emin, double emax, const double *weights, size_t nw,
                          int32_t *species, double *energy, double *dir) {
	orc_rng r;
	orc_rng_seed(&r, seed);
	double cdf[64];
	double acc = 0;
	for (size_t i = 0; i < nw && i < 64; i++) {
		acc += weights[i];
		cdf[i] = acc;
	}
	for (size_t i = 0; i < n; i++) {
		species[i] = nw > 1 ? (int32_t)orc_rng_bin(&r, cdf, nw) : 0;
		energy[i] = orc_rng_powerlaw(&r, -1, emin, emax);
		orc_rng_vector(&r, dir + 3 * i);
	}
}

/* ------------------------------------------------------------------ grid */
/* periodicClamp, include/crpropa/Grid.h:31-34 */
static inline void orc_periodic_clamp(double x, int n, int *lo, int *hi) {
	*lo = (((int)floor(x) % n) + n) % n;
	*hi = (*lo + 1) % n;
}

/* reflectiveClamp, include/crpropa/Grid.h:49-59 */
static inline void orc_reflective_clamp(double x, int n, int *lo, int *hi, double *res) {
	while ((x < -0.5) || (x > (n - 0.5)))
		x = 2 * n * (x > (n - 0.5)) - x - 1;
	*res = x;
	*lo = (int)floor(x);
	*hi = *lo + (*lo < n - 1);
	if (x < 0) {
		*lo = 0;
		*hi = 0;
	}
}

/* periodicBoundary, include/crpropa/Grid.h:43-46 */
int orc_periodic_boundary(int index, int n) { return ((index % n) + n) % n; }

/* reflectiveBoundary, include/crpropa/Grid.h:36-41 */
int orc_reflective_boundary(int index, int n) {
	while ((index < -0.5) || (index > (n - 0.5)))
		index = 2 * n * (index > (n - 0.5)) - index - 1;
	return index;
}

/* crpropa::round, include/crpropa/Grid.h:61-63 (not the C library's round: r + 0.5 is rounded
 * before the floor, so 0.5 - 2^-54 goes up) */
static inline double orc_round_sym(double r) { return (r > 0.0) ? floor(r + 0.5) : ceil(r - 0.5); }

/* Grid::closestValue (include/crpropa/Grid.h:383-405): flat index of the nearest grid point of a
 * grid with n points per axis, origin and spacing */
static size_t orc_closest_index(const int n[3], const double origin[3], const double spacing[3], int reflective,
                                const double p[3]) {
	int ix[3];
	for (int a = 0; a < 3; a++) {
		const double go = origin[a] + spacing[a] / 2;
		const double r = (p[a] - go) / spacing[a];
		if (reflective) {
			int i = (int)orc_round_sym(r);
			while ((i < -0.5) || (i > (n[a] - 0.5)))
				i = 2 * n[a] * (i > (n[a] - 0.5)) - i - 1;
			ix[a] = i;
		} else {
			int i = (int)orc_round_sym(fmod(r, n[a]));
			ix[a] = (i + n[a] * (i < 0)) % n[a];
		}
	}
	return ((size_t)ix[0] * n[1] + ix[1]) * n[2] + ix[2];
}

/* CubicInterpolate, include/crpropa/Grid.h:436-456: one lane of the SSE code, in float; position,
 * position^2 and position^3 formed in double and rounded to float once each (_mm_set1_ps) */
static float orc_cubic_f32(float p0, float p1, float p2, float p3, double position) {
	const float c1 = 0.5f, c2 = 1.5f, c3 = 2.f, c4 = 2.5f;
	const float pos = (float)position;
	const float pos2 = (float)(position * position);
	const float pos3 = (float)(position * position * position);
	const float term = (c1 * p2 - c1 * p0) * pos;
	const float term2 = ((p0 + c3 * p2) - (c4 * p1 + c1 * p3)) * pos2;
	const float term3 = ((c2 * p1 + c1 * p3) - (c1 * p0 + c2 * p2)) * pos3;
	return ((term3 + term2) + term) + p1;
}

/* CubicInterpolateScalar, include/crpropa/Grid.h:499-501, in double, left to right */
static double orc_cubic_f64(double p0, double p1, double p2, double p3, double pos) {
	return ((-0.5 * p0 + 3 / 2. * p1 - 3 / 2. * p2 + 0.5 * p3) * pos * pos * pos +
	        (p0 - 5 / 2. * p1 + p2 * 2 - 0.5 * p3) * pos * pos + (-0.5 * p0 + 0.5 * p2) * pos + p1);
}

/* Neighbour indices iX0-1 .. iX0+2 of one axis through periodicGet / reflectiveGet (Grid.h:346-358):
 * the int offset goes through size_t and back to int unchanged */
static void orc_cubic_axis(int i0, int n, int reflective, int out[4]) {
	for (int d = 0; d < 4; d++)
		out[d] = reflective ? orc_reflective_boundary(i0 - 1 + d, n) : orc_periodic_boundary(i0 - 1 + d, n);
}

/* Grid<Vector3f>::tricubicInterpolate (include/crpropa/Grid.h:459-492, SIMD build): cubic along z
 * for each (x, y) neighbour, then along y for each x, then along x; x, y, z lanes independent */
static void orc_grid_tricubic3(const orc_field *f, const double p[3], double B[3]) {
	double r[3], fr[3];
	int i0[3];
	for (int a = 0; a < 3; a++) {
		const double go = f->origin[a] + f->spacing[a] / 2;
		r[a] = (p[a] - go) / f->spacing[a];
		i0[a] = (int)floor(r[a]);
		fr[a] = r[a] - i0[a];
	}
	int X[4], Y[4], Z[4];
	orc_cubic_axis(i0[0], f->nx, f->reflective, X);
	orc_cubic_axis(i0[1], f->ny, f->reflective, Y);
	orc_cubic_axis(i0[2], f->nz, f->reflective, Z);
	float vx[4][3];
	for (int a = 0; a < 4; a++) {
		float vy[4][3];
		for (int b = 0; b < 4; b++) {
			const float *q[4];
			for (int c = 0; c < 4; c++)
				q[c] = f->grid + 3 * (((size_t)X[a] * f->ny + Y[b]) * f->nz + Z[c]);
			for (int l = 0; l < 3; l++)
				vy[b][l] = orc_cubic_f32(q[0][l], q[1][l], q[2][l], q[3][l], fr[2]);
		}
		for (int l = 0; l < 3; l++)
			vx[a][l] = orc_cubic_f32(vy[0][l], vy[1][l], vy[2][l], vy[3][l], fr[1]);
	}
	for (int l = 0; l < 3; l++)
		B[l] = orc_cubic_f32(vx[0][l], vx[1][l], vx[2][l], vx[3][l], fr[0]);
}

/* Grid<float>::tricubicInterpolate(double, ...) (include/crpropa/Grid.h:504-536): the same loops in
 * double over the modulation grid's float values; Grid::interpolate narrows the result to float */
static float orc_grid1_tricubic(const orc_field *f, const double p[3]) {
	const int n[3] = {f->mnx, f->mny, f->mnz};
	double fr[3];
	int idx[3][4];
	for (int a = 0; a < 3; a++) {
		const double go = f->morigin[a] + f->mspacing[a] / 2;
		const double r = (p[a] - go) / f->mspacing[a];
		const int i0 = (int)floor(r);
		fr[a] = r - i0;
		orc_cubic_axis(i0, n[a], f->mreflective, idx[a]);
	}
	double vx[4];
	for (int a = 0; a < 4; a++) {
		double vy[4];
		for (int b = 0; b < 4; b++) {
			double vz[4];
			for (int c = 0; c < 4; c++)
				vz[c] = f->mod[((size_t)idx[0][a] * f->mny + idx[1][b]) * f->mnz + idx[2][c]];
			vy[b] = orc_cubic_f64(vz[0], vz[1], vz[2], vz[3], fr[2]);
		}
		vx[a] = orc_cubic_f64(vy[0], vy[1], vy[2], vy[3], fr[1]);
	}
	return (float)orc_cubic_f64(vx[0], vx[1], vx[2], vx[3], fr[0]);
}

/* Grid<Vector3f>::interpolate + trilinearInterpolate, include/crpropa/Grid.h:318-335,540-582.
 * Weights are narrowed to float (Vector3<float>::operator*(float)), products
 * and the 8-term sum run in float, the result widens to double
 * (src/magneticField/MagneticFieldGrid.cpp:17-19). */
static void orc_grid_eval1(const orc_field *f, const double p[3], double B[3]) {
	if (f->clip) {
		double ex = f->origin[0] + (double)f->nx * f->spacing[0];
		double ey = f->origin[1] + (double)f->ny * f->spacing[1];
		double ez = f->origin[2] + (double)f->nz * f->spacing[2];
		int in = (p[0] >= f->origin[0]) && (p[0] <= ex);
		in &= (p[1] >= f->origin[1]) && (p[1] <= ey);
		in &= (p[2] >= f->origin[2]) && (p[2] <= ez);
		if (!in) {
			B[0] = B[1] = B[2] = 0;
			return;
		}
	}
	if (f->interp == CRP_GRID_TRICUBIC) {
		orc_grid_tricubic3(f, p, B);
		return;
	}
	if (f->interp == CRP_GRID_NEAREST_NEIGHBOUR) {
		const int n[3] = {f->nx, f->ny, f->nz};
		const float *v = f->grid + 3 * orc_closest_index(n, f->origin, f->spacing, f->reflective, p);
		B[0] = v[0];
		B[1] = v[1];
		B[2] = v[2];
		return;
	}
	/* gridOrigin = origin + spacing/2 (Grid.h:225-228); r = (pos - gridOrigin)/spacing */
	double r[3];
	for (int a = 0; a < 3; a++) {
		double go = f->origin[a] + f->spacing[a] / 2;
		r[a] = (p[a] - go) / f->spacing[a];
	}
	int n[3] = {f->nx, f->ny, f->nz};
	int lo[3], hi[3];
	double f0[3];
	for (int a = 0; a < 3; a++) {
		if (f->reflective) {
			double res;
			orc_reflective_clamp(r[a], n[a], &lo[a], &hi[a], &res);
			f0[a] = res - floor(res);
		} else {
			orc_periodic_clamp(r[a], n[a], &lo[a], &hi[a]);
			f0[a] = r[a] - floor(r[a]);
		}
	}
	float wx0 = (float)f0[0], wy0 = (float)f0[1], wz0 = (float)f0[2];
	float wx1 = (float)(1 - f0[0]), wy1 = (float)(1 - f0[1]), wz1 = (float)(1 - f0[2]);
	const int X[8] = {lo[0], hi[0], lo[0], lo[0], hi[0], lo[0], hi[0], hi[0]};
	const int Y[8] = {lo[1], lo[1], hi[1], lo[1], lo[1], hi[1], hi[1], hi[1]};
	const int Z[8] = {lo[2], lo[2], lo[2], hi[2], hi[2], hi[2], lo[2], hi[2]};
	const float WX[8] = {wx1, wx0, wx1, wx1, wx0, wx1, wx0, wx0};
	const float WY[8] = {wy1, wy1, wy0, wy1, wy1, wy0, wy0, wy0};
	const float WZ[8] = {wz1, wz1, wz1, wz0, wz0, wz0, wz1, wz0};
	float b[3] = {0.f, 0.f, 0.f};
	for (int c = 0; c < 8; c++) {
		size_t idx = ((size_t)X[c] * f->ny + Y[c]) * f->nz + Z[c];
		const float *v = f->grid + 3 * idx;
		for (int a = 0; a < 3; a++) {
			float t = ((v[a] * WX[c]) * WY[c]) * WZ[c];
			b[a] += t;
		}
	}
	B[0] = b[0];
	B[1] = b[1];
	B[2] = b[2];
}

/* Grid<float>::interpolate + trilinearInterpolate (include/crpropa/Grid.h:318-335,540-582) with
 * T = float: each corner value times the double weights in double, and every b += term
 * rounds the double sum to float. */
static float orc_grid1_eval(const orc_field *f, const double p[3]) {
	if (f->mclip) {
		for (int a = 0; a < 3; a++) {
			const double n = a == 0 ? f->mnx : (a == 1 ? f->mny : f->mnz);
			if (!((p[a] >= f->morigin[a]) && (p[a] <= f->morigin[a] + n * f->mspacing[a])))
				return 0.f;
		}
	}
	if (f->minterp == CRP_GRID_TRICUBIC)
		return orc_grid1_tricubic(f, p);
	if (f->minterp == CRP_GRID_NEAREST_NEIGHBOUR) {
		const int n[3] = {f->mnx, f->mny, f->mnz};
		return f->mod[orc_closest_index(n, f->morigin, f->mspacing, f->mreflective, p)];
	}
	double r[3];
	for (int a = 0; a < 3; a++) {
		double go = f->morigin[a] + f->mspacing[a] / 2;
		r[a] = (p[a] - go) / f->mspacing[a];
	}
	int n[3] = {f->mnx, f->mny, f->mnz};
	int lo[3], hi[3];
	double f0[3];
	for (int a = 0; a < 3; a++) {
		if (f->mreflective) {
			double res;
			orc_reflective_clamp(r[a], n[a], &lo[a], &hi[a], &res);
			f0[a] = res - floor(res);
		} else {
			orc_periodic_clamp(r[a], n[a], &lo[a], &hi[a]);
			f0[a] = r[a] - floor(r[a]);
		}
	}
	const double f1[3] = {1 - f0[0], 1 - f0[1], 1 - f0[2]};
	const int X[8] = {lo[0], hi[0], lo[0], lo[0], hi[0], lo[0], hi[0], hi[0]};
	const int Y[8] = {lo[1], lo[1], hi[1], lo[1], lo[1], hi[1], hi[1], hi[1]};
	const int Z[8] = {lo[2], lo[2], lo[2], hi[2], hi[2], hi[2], lo[2], hi[2]};
	const double WX[8] = {f1[0], f0[0], f1[0], f1[0], f0[0], f1[0], f0[0], f0[0]};
	const double WY[8] = {f1[1], f1[1], f0[1], f1[1], f1[1], f0[1], f0[1], f0[1]};
	const double WZ[8] = {f1[2], f1[2], f1[2], f0[2], f0[2], f0[2], f1[2], f0[2]};
	float b = 0.f;
	for (int c = 0; c < 8; c++) {
		const float v = f->mod[((size_t)X[c] * f->mny + Y[c]) * f->mnz + Z[c]];
		b = (float)((double)b + (((double)v * WX[c]) * WY[c]) * WZ[c]);
	}
	return b;
}

/* ModulatedMagneticFieldGrid::getField (src/magneticField/MagneticFieldGrid.cpp:51-55):
 * the Grid3f value widened to double, times the modulation widened to double */
static void orc_modulated_eval1(const orc_field *f, const double p[3], double B[3]) {
	const float m = orc_grid1_eval(f, p);
	orc_grid_eval1(f, p, B);
	for (int a = 0; a < 3; a++)
		B[a] = B[a] * (double)m;
}

/* PlaneWaveTurbulence::getField (default build), src/magneticField/turbulentField/PlaneWaveTurbulence.cpp:231-240 */
static void orc_planewave_eval1(const orc_field *f, const double p[3], double B[3], int libm) {
	double bx = 0, by = 0, bz = 0;
	for (int i = 0; i < f->nm; i++) {
		const double *kap = f->kappa + 3 * i;
		const double *xi = f->xi + 3 * i;
		double z_ = (p[0] * kap[0] + p[1] * kap[1]) + p[2] * kap[2];
		double c = orc_cos(f->k[i] * z_ + f->beta[i], libm);
		double A = f->Ak[i];
		bx += (xi[0] * A) * c;
		by += (xi[1] * A) * c;
		bz += (xi[2] * A) * c;
	}
	B[0] = bx;
	B[1] = by;
	B[2] = bz;
}

/* ---------------------------------------------------------------- JF12 */
static const double J_KPC = 1e3 * (648000 / M_PI * 149597870700.0);

void orc_jf12_defaults(int planck, crp_jf12_params *p) {
	const double muG = 1e-6 * (1e-4 * 1.0);
	memset(p, 0, sizeof(*p));
	p->use_regular = p->use_disk = p->use_toroidal_halo = p->use_x = 1;
	const double pitch = 11.5 * M_PI / 180;
	p->sin_pitch = sin(pitch);
	p->cos_pitch = cos(pitch);
	p->tan90_minus_pitch = tan(M_PI / 2 - pitch);
	const double arms[8] = {5.1, 6.3, 7.1, 8.3, 9.8, 11.4, 12.7, 15.5};
	const double disk[8] = {0.1, 3.0, -0.9, -0.8, -2.0, -4.2, 0.0, 2.7};
	const double turb[8] = {10.81, 6.96, 9.59, 6.96, 1.96, 16.34, 37.29, 10.35};
	for (int i = 0; i < 8; i++) {
		p->r_arms[i] = arms[i] * J_KPC;
		p->b_disk[i] = disk[i] * muG;
		p->b_disk_turb[i] = turb[i] * muG;
	}
	p->b_ring = 0.1 * muG; p->h_disk = 0.40 * J_KPC; p->w_disk = 0.27 * J_KPC;
	p->b_north = 1.4 * muG; p->b_south = -1.1 * muG; p->r_north = 9.22 * J_KPC; p->r_south = 17 * J_KPC;
	p->w_halo = 0.20 * J_KPC; p->z0 = 5.3 * J_KPC;
	p->b_x = 4.6 * muG;
	const double thetaX0 = 49.0 * M_PI / 180;
	p->sin_theta_x0 = sin(thetaX0); p->cos_theta_x0 = cos(thetaX0); p->tan_theta_x0 = tan(thetaX0);
	p->r_xc = 4.8 * J_KPC; p->r_x = 2.9 * J_KPC;
	p->sqrt_beta = sqrt(1.36);
	p->b_disk_turb5 = 7.63 * muG; p->z_disk_turb = 0.61 * J_KPC;
	p->b_halo_turb = 4.68 * muG; p->r_halo_turb = 10.97 * J_KPC; p->z_halo_turb = 2.84 * J_KPC;
	if (planck) {
		const double pt[8] = {3.12, 6.24, 3.12, 6.24, 3.12, 6.24, 3.12, 6.24};
		p->b_disk[5] = -3.5 * muG;
		p->b_x = 1.8 * muG;
		for (int i = 0; i < 8; i++)
			p->b_disk_turb[i] = pt[i] * muG;
		p->b_disk_turb5 = 3.90 * muG;
		p->b_halo_turb = 7.332 * muG;
	}
}

/* Grid<float>::closestValue (include/crpropa/Grid.h:383-405) on the modulation-grid slots */
static float orc_grid1_closest(const orc_field *f, const double p[3]) {
	const int n[3] = {f->mnx, f->mny, f->mnz};
	return f->mod[orc_closest_index(n, f->morigin, f->mspacing, f->mreflective, p)];
}

/* pow(x, 2.) as the reference's build evaluates it: GCC rewrites the call as x * x (the correctly
 * rounded square) */
static double jf12_sq(double x) { return x * x; }

/* sin and cos of one angle: GCC merges the reference's two calls into one sincos() */
static void jf12_sincos(double x, double *s, double *c, int libm) {
	if (libm == ORC_LIBM_CR) {
		*s = crp_sin_cr(x);
		*c = crp_cos_cr(x);
	} else {
		sincos(x, s, c);
	}
}

static double jf12_logistic(double x, double x0, double w, int libm) { /* JF12Field.cpp:188-190 */
	return 1. / (1. + orc_exp(-2. * (fabs(x) - x0) / w, libm));
}

static double jf12_phi(const double p[3], int libm) { /* Vector3::getPhi, include/crpropa/Vector3.h:136-142 */
	const double eps = 2.2250738585072014e-308;
	if ((fabs(p[0]) < eps) && (fabs(p[1]) < eps))
		return 0.0;
	return orc_atan2(p[1], p[0], libm);
}

static double jf12_rnegx(const crp_jf12_params *J, double r, double phi, int libm) {
	double rn = r * orc_exp(-(phi - M_PI) / J->tan90_minus_pitch, libm);
	if (rn > J->r_arms[7])
		rn = r * orc_exp(-(phi + M_PI) / J->tan90_minus_pitch, libm);
	if (rn > J->r_arms[7])
		rn = r * orc_exp(-(phi + 3 * M_PI) / J->tan90_minus_pitch, libm);
	return rn;
}

/* JF12Field::getRegularField, getDiskField, getToroidalHaloField, getXField (JF12Field.cpp:192-300) */
static void jf12_regular(const crp_jf12_params *J, const double p[3], double B[3], int libm) {
	double b[3] = {0, 0, 0};
	const double d = sqrt((p[0] * p[0] + p[1] * p[1]) + p[2] * p[2]);
	if (d < 20 * J_KPC) {
		const double r = sqrt(p[0] * p[0] + p[1] * p[1]);
		const double phi = jf12_phi(p, libm);
		double sinPhi, cosPhi;
		jf12_sincos(phi, &sinPhi, &cosPhi, libm);
		const double z = p[2];
		double dk[3] = {0, 0, 0};
		if (J->use_disk) {
			const double lfDisk = jf12_logistic(z, J->h_disk, J->w_disk, libm);
			if (r > 3 * J_KPC) {
				double bMag = 0;
				if (r < 5 * J_KPC) {
					bMag = J->b_ring * (5 * J_KPC / r) * (1 - lfDisk);
					dk[0] += -bMag * sinPhi;
					dk[1] += bMag * cosPhi;
				} else {
					const double rn = jf12_rnegx(J, r, phi, libm);
					for (int i = 7; i >= 0; i--)
						if (rn < J->r_arms[i])
							bMag = J->b_disk[i];
					bMag *= (5 * J_KPC / r) * (1 - lfDisk);
					dk[0] += bMag * (J->sin_pitch * cosPhi - J->cos_pitch * sinPhi);
					dk[1] += bMag * (J->sin_pitch * sinPhi + J->cos_pitch * cosPhi);
				}
			}
		}
		for (int a = 0; a < 3; a++)
			b[a] += dk[a];
		double h[3] = {0, 0, 0};
		if (J->use_toroidal_halo && (r * r + z * z > 1 * J_KPC * J_KPC)) {
			const double lfDisk = jf12_logistic(z, J->h_disk, J->w_disk, libm);
			double bMagH = orc_exp(-fabs(z) / J->z0, libm) * lfDisk;
			if (z >= 0)
				bMagH *= J->b_north * (1 - jf12_logistic(r, J->r_north, J->w_halo, libm));
			else
				bMagH *= J->b_south * (1 - jf12_logistic(r, J->r_south, J->w_halo, libm));
			h[0] += -bMagH * sinPhi;
			h[1] += bMagH * cosPhi;
		}
		for (int a = 0; a < 3; a++)
			b[a] += h[a];
		double x[3] = {0, 0, 0};
		if (J->use_x && (r * r + z * z > 1 * J_KPC * J_KPC)) {
			double bMagX, sinThetaX, cosThetaX, rp;
			const double rc = J->r_xc + fabs(z) / J->tan_theta_x0;
			if (r < rc) {
				rp = r * J->r_xc / rc;
				bMagX = J->b_x * orc_exp(-1 * rp / J->r_x, libm) * jf12_sq(J->r_xc / rc);
				double thetaX = orc_atan2(fabs(z), (r - rp), libm);
				if (z == 0)
					thetaX = M_PI / 2.;
				jf12_sincos(thetaX, &sinThetaX, &cosThetaX, libm);
			} else {
				rp = r - fabs(z) / J->tan_theta_x0;
				bMagX = J->b_x * orc_exp(-rp / J->r_x, libm) * (rp / r);
				sinThetaX = J->sin_theta_x0;
				cosThetaX = J->cos_theta_x0;
			}
			const double zsign = z < 0 ? -1 : 1;
			x[0] += zsign * bMagX * cosThetaX * cosPhi;
			x[1] += zsign * bMagX * cosThetaX * sinPhi;
			x[2] += bMagX * sinThetaX;
		}
		for (int a = 0; a < 3; a++)
			b[a] += x[a];
	}
	B[0] = b[0];
	B[1] = b[1];
	B[2] = b[2];
}

/* JF12Field::getTurbulentStrength (JF12Field.cpp:307-331) */
static double jf12_turb_strength(const crp_jf12_params *J, const double p[3], int libm) {
	if (sqrt((p[0] * p[0] + p[1] * p[1]) + p[2] * p[2]) > 20 * J_KPC)
		return 0;
	const double r = sqrt(p[0] * p[0] + p[1] * p[1]);
	const double phi = jf12_phi(p, libm);
	double bDisk = 0;
	if (r < 5 * J_KPC) {
		bDisk = J->b_disk_turb5;
	} else {
		const double rn = jf12_rnegx(J, r, phi, libm);
		for (int i = 7; i >= 0; i--)
			if (rn < J->r_arms[i])
				bDisk = J->b_disk_turb[i];
		bDisk *= (5 * J_KPC) / r;
	}
	bDisk *= orc_exp(-0.5 * jf12_sq(p[2] / J->z_disk_turb), libm);
	const double bHalo = J->b_halo_turb * orc_exp(-r / J->r_halo_turb, libm) *
	                     orc_exp(-0.5 * jf12_sq(p[2] / J->z_halo_turb), libm);
	return sqrt(jf12_sq(bDisk) + jf12_sq(bHalo));
}

/* JF12Field::getField (JF12Field.cpp:337-346): turbulent (Vector3f * float(strength) in float),
 * then striated (regular * (1 + sqrt(beta) * closest value)) or regular */
static void orc_jf12_eval1(const orc_field *f, const double p[3], double B[3], int libm) {
	const crp_jf12_params *J = f->jf12;
	double b[3] = {0, 0, 0};
	if (J->use_turbulent) {
		double g[3];
		orc_grid_eval1(f, p, g);
		const float s = (float)jf12_turb_strength(J, p, libm);
		for (int a = 0; a < 3; a++)
			b[a] += (double)((float)g[a] * s);
	}
	if (J->use_striated) {
		double rg[3];
		jf12_regular(J, p, rg, libm);
		const double m = 1. + J->sqrt_beta * (double)orc_grid1_closest(f, p);
		for (int a = 0; a < 3; a++)
			b[a] += rg[a] * m;
	} else if (J->use_regular) {
		double rg[3];
		jf12_regular(J, p, rg, libm);
		for (int a = 0; a < 3; a++)
			b[a] += rg[a];
	}
	B[0] = b[0];
	B[1] = b[1];
	B[2] = b[2];
}

/* MagneticField::getField(position), the one-argument call (include/crpropa/magneticField/MagneticField.h:26-28):
 * a missing field gives B = 0 (src/module/PropagationBP.cpp:137-149). */
static void orc_field_eval1(const orc_field *f, const double p[3], double B[3], int libm) {
	switch (f ? f->type : ORC_FIELD_NONE) {
	case ORC_FIELD_UNIFORM: B[0] = f->B0[0]; B[1] = f->B0[1]; B[2] = f->B0[2]; break;
	case ORC_FIELD_GRID: orc_grid_eval1(f, p, B); break;
	case ORC_FIELD_PLANEWAVE: orc_planewave_eval1(f, p, B, libm); break;
	case ORC_FIELD_MODULATED: orc_modulated_eval1(f, p, B); break;
	case ORC_FIELD_JF12: orc_jf12_eval1(f, p, B, libm); break;
	case ORC_FIELD_PERIODIC: {
		/* PeriodicMagneticField::getField, src/magneticField/MagneticField.cpp:41-58:
		 * n = floor((position - origin) / extends), p = position - origin - n * extends,
		 * reflected in the odd periods */
		double q[3];
		for (int a = 0; a < 3; a++) {
			double n = floor((p[a] - f->porigin[a]) / f->extends[a]);
			q[a] = (p[a] - f->porigin[a]) - n * f->extends[a];
			if (f->preflective && ((long)fabs(n) % 2) == 1)
				q[a] = f->extends[a] - q[a];
		}
		orc_field_eval1(f->children[0], q, B, libm);
		break;
	}
	case ORC_FIELD_LIST: {
		/* MagneticFieldList::getField, src/magneticField/MagneticField.cpp:64-69: b = 0; b += each */
		double b[3] = {0, 0, 0};
		for (int i = 0; i < f->nchildren; i++) {
			double c[3];
			orc_field_eval1(f->children[i], p, c, libm);
			b[0] += c[0];
			b[1] += c[1];
			b[2] += c[2];
		}
		B[0] = b[0]; B[1] = b[1]; B[2] = b[2];
		break;
	}
	/* MagneticFieldEvolution only overrides getField(position, z): its one-argument getField is the
	 * base class's, zero (MagneticField.h:26-28, 102) */
	default: B[0] = B[1] = B[2] = 0; break;
	}
}

/* MagneticField::getField(position, z), the call of PropagationBP/CK::getFieldAtPosition
 * (src/module/PropagationBP.cpp:137-149): the base class forwards to getField(position);
 * MagneticFieldEvolution::getField (src/magneticField/MagneticField.cpp:76-79) scales its field's
 * getField(position) by pow(1 + z, m). */
static inline void orc_field_eval2(const orc_field *f, const double p[3], double z, double B[3], int libm) {
	if (f && f->type == ORC_FIELD_EVOLUTION) {
		double c[3];
		orc_field_eval1(f->children[0], p, c, libm);
		const double s = orc_pow(1 + z, f->evo_m, libm);
		B[0] = c[0] * s;
		B[1] = c[1] * s;
		B[2] = c[2] * s;
		return;
	}
	orc_field_eval1(f, p, B, libm);
}

void orc_field_eval(const orc_field *f, const double *pos, size_t m, double *B, int libm) {
	for (size_t i = 0; i < m; i++)
		orc_field_eval1(f, pos + 3 * i, B + 3 * i, libm);
}

void orc_field_eval_z(const orc_field *f, const double *pos, const double *z, size_t m, double *B, int libm) {
	for (size_t i = 0; i < m; i++)
		orc_field_eval2(f, pos + 3 * i, z[i], B + 3 * i, libm);
}

/* PlaneWaveTurbulence ctor, src/magneticField/turbulentField/PlaneWaveTurbulence.cpp:99-178 (seed != 0).
 * simple_spectrum selects E(k) = k^(-s-2) (SimpleTurbulenceSpectrum,
 * include/crpropa/magneticField/turbulentField/SimpleGridTurbulence.h:36-38; lBendover is
 * then 1000*lMax) instead of TurbulenceSpectrum::energySpectrum
 * (include/crpropa/magneticField/turbulentField/TurbulentField.h:72-77). */
void orc_planewave_init(double brms, double lmin, double lmax, double lbendover, double sindex, double qindex,
                        int simple_spectrum, int nm, int seed, double *k, double *Ak, double *xi, double *kappa,
                        double *beta) {
	orc_rng r;
	orc_rng_seed(&r, (uint32_t)seed);
	double kmax = 2 * M_PI / lmin;
	double kmin = 2 * M_PI / lmax;
	double delta = log10(kmax / kmin);
	for (int i = 0; i < nm; i++)
		k[i] = pow(10, log10(kmin) + ((double)i) / ((double)(nm - 1)) * delta);
	double delta_k0 = (k[1] - k[0]) / k[1];
	double Ak2_sum = 0;
	for (int i = 0; i < nm; i++) {
		double kk = k[i];
		double kHat = kk * lbendover;
		double E;
		if (simple_spectrum) {
			E = pow(kk, -sindex - 2);
		} else {
			double kh = kk * lbendover;
			E = pow(kh, qindex) / pow(1.0 + kh * kh, (sindex + qindex) / 2.0 + 1.0);
		}
		double Gk = E * (1 + kHat * kHat);
		Ak[i] = Gk * delta_k0 * kk;
		Ak2_sum += Ak[i];
		double phi = orc_rng_uniform(&r, -M_PI, M_PI);
		double costheta = orc_rng_uniform(&r, -1., 1.);
		double sintheta = sqrt(1 - costheta * costheta);
		double alpha = orc_rng_uniform(&r, 0, 2 * M_PI);
		double b = orc_rng_uniform(&r, 0, 2 * M_PI);
		kappa[3 * i] = sintheta * cos(phi);
		kappa[3 * i + 1] = sintheta * sin(phi);
		kappa[3 * i + 2] = costheta;
		xi[3 * i] = costheta * cos(phi) * cos(alpha) + sin(phi) * sin(alpha);
		xi[3 * i + 1] = costheta * sin(phi) * cos(alpha) - cos(phi) * sin(alpha);
		xi[3 * i + 2] = -sintheta * cos(alpha);
		beta[i] = b;
	}
	for (int i = 0; i < nm; i++)
		Ak[i] = sqrt(2 * Ak[i] / Ak2_sum) * brms;
}

/* ------------------------------------------------------------------ propagators */
typedef struct {
	double x[3], u[3];
} orc_y;

/* PropagationBP::dY, src/module/PropagationBP.cpp:19-39 */
static inline orc_y orc_bp_dy(const orc_field *f, const double pos0[3], const double dir0[3], double step, double q,
                              double m, double z, int libm) {
	double pos[3], dir[3], B[3], t[3], s[3], v[3];
	for (int a = 0; a < 3; a++)
		pos[a] = pos0[a] + (dir0[a] * step) / 2.;
	orc_field_eval2(f, pos, z, B, libm);
	for (int a = 0; a < 3; a++)
		t[a] = ((((B[a] * q) / 2) / m) * step) / C_LIGHT;
	double tt = (t[0] * t[0] + t[1] * t[1]) + t[2] * t[2];
	for (int a = 0; a < 3; a++)
		s[a] = (t[a] * 2) / (1 + tt);
	/* v_help = dir + dir x t */
	v[0] = dir0[0] + (dir0[1] * t[2] - t[1] * dir0[2]);
	v[1] = dir0[1] + (dir0[2] * t[0] - t[2] * dir0[0]);
	v[2] = dir0[2] + (dir0[0] * t[1] - t[0] * dir0[1]);
	/* dir = dir + v_help x s */
	dir[0] = dir0[0] + (v[1] * s[2] - s[1] * v[2]);
	dir[1] = dir0[1] + (v[2] * s[0] - s[2] * v[0]);
	dir[2] = dir0[2] + (v[0] * s[1] - s[0] * v[1]);
	orc_y y;
	for (int a = 0; a < 3; a++) {
		y.x[a] = pos[a] + (dir[a] * step) / 2.;
		y.u[a] = dir[a];
	}
	return y;
}

static inline double orc_norm(const double v[3]) { return sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]); }

/* PropagationCK::dYdt, src/module/PropagationCK.cpp:51-61 */
static inline orc_y orc_ck_dydt(const orc_field *f, const orc_y *y, double q, double E, double z, int libm) {
	orc_y d;
	double n = orc_norm(y->u);
	double B[3];
	for (int a = 0; a < 3; a++)
		d.x[a] = (y->u[a] / n) * C_LIGHT;
	orc_field_eval2(f, y->x, z, B, libm);
	double fac = (q * C_LIGHT) / E;
	d.u[0] = (d.x[1] * B[2] - B[1] * d.x[2]) * fac;
	d.u[1] = (d.x[2] * B[0] - B[2] * d.x[0]) * fac;
	d.u[2] = (d.x[0] * B[1] - B[0] * d.x[1]) * fac;
	return d;
}

/* Cash-Karp tableau, src/module/PropagationCK.cpp:11-26 (textbook constants) */
static const double CK_A[36] = {0., 0., 0., 0., 0., 0.,
                                1. / 5., 0., 0., 0., 0., 0.,
                                3. / 40., 9. / 40., 0., 0., 0., 0.,
                                3. / 10., -9. / 10., 6. / 5., 0., 0., 0.,
                                -11. / 54., 5. / 2., -70. / 27., 35. / 27., 0., 0.,
                                1631. / 55296., 175. / 512., 575. / 13824., 44275. / 110592., 253. / 4096., 0.};
static const double CK_B[6] = {37. / 378., 0, 250. / 621., 125. / 594., 0., 512. / 1771.};
static const double CK_BS[6] = {2825. / 27648., 0., 18575. / 48384., 13525. / 55296., 277. / 14336., 1. / 4.};

/* PropagationCK::tryStep, src/module/PropagationCK.cpp:28-49 */
static inline void orc_ck_trystep(const orc_field *f, const orc_y *y, orc_y *out, orc_y *err, double h, double q,
                                  double E, double z, int libm) {
	orc_y k[6];
	*out = *y;
	for (int a = 0; a < 3; a++)
		err->x[a] = err->u[a] = 0;
	for (int i = 0; i < 6; i++) {
		orc_y yn = *y;
		for (int j = 0; j < i; j++) {
			double aij = CK_A[i * 6 + j];
			for (int a = 0; a < 3; a++) {
				yn.x[a] += (k[j].x[a] * aij) * h;
				yn.u[a] += (k[j].u[a] * aij) * h;
			}
		}
		k[i] = orc_ck_dydt(f, &yn, q, E, z, libm);
		double bi = CK_B[i], ei = CK_B[i] - CK_BS[i];
		for (int a = 0; a < 3; a++) {
			out->x[a] += (k[i].x[a] * bi) * h;
			out->u[a] += (k[i].u[a] * bi) * h;
			err->x[a] += (k[i].x[a] * ei) * h;
			err->u[a] += (k[i].u[a] * ei) * h;
		}
	}
}

/* The reference's single-step functions on n phase points (tests): PropagationBP::dY and tryStep
 * (src/module/PropagationBP.cpp:8-39,152-159), PropagationCK::dYdt and tryStep
 * (src/module/PropagationCK.cpp:28-61).  x, u and the outputs AoS [3n]; m: mass (BP) or energy (CK). */
void orc_step_batch(const orc_field *f, int op, size_t n, const double *x, const double *u, const double *h,
                    const double *z, const double *q, const double *m, double *ox, double *ou, double *ex, double *eu,
                    int libm) {
	for (size_t i = 0; i < n; i++) {
		orc_y y, o, e;
		for (int a = 0; a < 3; a++) {
			y.x[a] = x[3 * i + a];
			y.u[a] = u[3 * i + a];
			e.x[a] = e.u[a] = 0;
		}
		const double zi = z ? z[i] : 0.0;
		if (op == 0 || op == 1) {
			o = orc_bp_dy(f, y.x, y.u, h[i], q[i], m[i], zi, libm);
			if (op == 1) {
				orc_y half = orc_bp_dy(f, y.x, y.u, h[i] / 2, q[i], m[i], zi, libm);
				orc_y cmp = orc_bp_dy(f, half.x, half.u, h[i] / 2, q[i], m[i], zi, libm);
				double diff[3] = {o.x[0] - cmp.x[0], o.x[1] - cmp.x[1], o.x[2] - cmp.x[2]};
				double S = orc_norm(diff) / (h[i] * (1 - 1 / 4.));
				for (int a = 0; a < 3; a++)
					e.x[a] = e.u[a] = S; /* Y(double) */
			}
		} else if (op == 2) {
			o = orc_ck_dydt(f, &y, q[i], m[i], zi, libm);
		} else {
			orc_ck_trystep(f, &y, &o, &e, h[i], q[i], m[i], zi, libm);
		}
		for (int a = 0; a < 3; a++) {
			ox[3 * i + a] = o.x[a];
			ou[3 * i + a] = o.u[a];
			if (ex) {
				ex[3 * i + a] = e.x[a];
				eu[3 * i + a] = e.u[a];
			}
		}
	}
}

static inline double orc_clip(double x, double lo, double hi) {
	/* std::max(lower, std::min(x, upper)), include/crpropa/Common.h:29-32 */
	double m = (hi < x) ? hi : x;
	return (m < lo) ? lo : m;
}
static inline double orc_max(double a, double b) { return (a < b) ? b : a; } /* std::max */
static inline double orc_min(double a, double b) { return (b < a) ? b : a; } /* std::min */

/* One process() call of PropagationBP (src/module/PropagationBP.cpp:62-124) or
 * PropagationCK (src/module/PropagationCK.cpp:77-136), followed by
 * MaximumTrajectoryLength::process (src/module/BreakCondition.cpp:42-64). */
/* PeriodicBox::process (src/module/Boundary.cpp:16-27): n = floor((pos - origin) / size);
 * current, previous, source and created positions move by -n * size */
static void orc_periodic_box(const orc_propagator *p, orc_state *s, size_t i, double *pos) {
	if (!p->box)
		return;
	double nb[3];
	for (int a = 0; a < 3; a++)
		nb[a] = floor((pos[a] - p->boxOrigin[a]) / p->boxSize[a]);
	if ((nb[0] == 0) && (nb[1] == 0) && (nb[2] == 0))
		return;
	for (int a = 0; a < 3; a++) {
		const double sh = nb[a] * p->boxSize[a];
		pos[a] = pos[a] - sh;
		if (s->prev_pos)
			s->prev_pos[3 * i + a] = s->prev_pos[3 * i + a] - sh;
		if (s->source_pos)
			s->source_pos[3 * i + a] = s->source_pos[3 * i + a] - sh;
		if (s->created_pos)
			s->created_pos[3 * i + a] = s->created_pos[3 * i + a] - sh;
	}
}

/* SphericalBoundary::process (src/module/Boundary.cpp:216-223): d = |pos - center|
 * (Vector3::getR, include/crpropa/Vector3.h:126-128); reject when d >= radius;
 * limitNextStep(radius - d + margin) */
static void orc_spherical_boundary(const orc_propagator *p, orc_state *s, size_t i, const double *pos) {
	if (!p->sphere)
		return;
	const double dx = pos[0] - p->sphereCenter[0], dy = pos[1] - p->sphereCenter[1], dz = pos[2] - p->sphereCenter[2];
	const double d = sqrt(dx * dx + dy * dy + dz * dz);
	if (d >= p->sphereRadius)
		s->active[i] = 0;
	if (p->sphereLimitStep)
		s->next_step[i] = orc_min(s->next_step[i], p->sphereRadius - d + p->sphereMargin);
}

/* MaximumTrajectoryLength::process (src/module/BreakCondition.cpp:42-64) without observer positions */
static void orc_max_length(const orc_propagator *p, orc_state *s, size_t i) {
	if (!(p->maxTrajectoryLength > 0))
		return;
	const double len = s->trajectory_length[i];
	if (len >= p->maxTrajectoryLength)
		s->active[i] = 0;
	else
		s->next_step[i] = orc_min(s->next_step[i], p->maxTrajectoryLength - len);
}

/* Observer::process with ObserverTimeEvolution::checkDetection (src/module/Observer.cpp:26-57,
 * 279-306): one detection per call at most; limitNextStep towards the next time; the snapshot
 * holds the state the observer sees (nextStep limited by the modules before it only) */
static void orc_observer(const orc_propagator *p, orc_state *s, size_t i, const double *pos, const double *dir) {
	if (p->nObs <= 0)
		return;
	const int k = s->det_index[i];
	if (k >= p->nObs)
		return;
	const double len = s->trajectory_length[i];
	const double distance = len - p->obsTimes[k];
	if (distance < 0.) {
		s->next_step[i] = orc_min(s->next_step[i], -distance);
		return;
	}
	if (k < p->nObs - 2)
		s->next_step[i] = orc_min(s->next_step[i], p->obsTimes[k + 1] - len);
	s->det_index[i] = k + 1;
	double *r = s->snap + ((size_t)k * s->n + i) * 9;
	for (int a = 0; a < 3; a++) {
		r[a] = pos[a];
		r[3 + a] = dir[a];
	}
	r[6] = len;
	r[7] = s->time[i];
	r[8] = s->next_step[i];
	if (p->obsDeactivate)
		s->active[i] = 0;
}

static inline void orc_process(const orc_field *f, const orc_propagator *p, orc_state *s, size_t i, int libm,
                               long long *trials) {
	double *pos = s->pos + 3 * i, *dir = s->dir + 3 * i;
	if (s->prev_pos) {
		for (int a = 0; a < 3; a++) {
			s->prev_pos[3 * i + a] = pos[a];
			s->prev_dir[3 * i + a] = dir[a];
		}
	}
	double q = s->charge[i];
	const double zr = s->redshift ? s->redshift[i] : 0.0; /* candidate->getRedshift() */
	double step = p->maxStep;
	double newStep;
	orc_y yIn, yOut, yErr;
	for (int a = 0; a < 3; a++) {
		yIn.x[a] = pos[a];
		yIn.u[a] = dir[a];
	}
	if (q == 0) {
		step = orc_clip(s->next_step[i], p->minStep, p->maxStep);
		for (int a = 0; a < 3; a++)
			pos[a] = yIn.x[a] + yIn.u[a] * step;
		newStep = p->maxStep;
	} else {
		newStep = step;
		double E = s->energy[i];
		if (p->minStep == p->maxStep) {
			if (p->type == ORC_PROP_BP) {
				double m = E / (C_LIGHT * C_LIGHT);
				yOut = orc_bp_dy(f, yIn.x, yIn.u, step, q, m, zr, libm);
			} else {
				orc_ck_trystep(f, &yIn, &yOut, &yErr, step / C_LIGHT, q, E, zr, libm);
			}
		} else {
			step = orc_clip(s->next_step[i], p->minStep, p->maxStep);
			newStep = step;
			double r;
			double m = E / (C_LIGHT * C_LIGHT);
			for (;;) {
				if (p->type == ORC_PROP_BP) {
					/* tryStep + errorEstimation, src/module/PropagationBP.cpp:8-16,152-159 */
					yOut = orc_bp_dy(f, yIn.x, yIn.u, step, q, m, zr, libm);
					orc_y half = orc_bp_dy(f, yIn.x, yIn.u, step / 2, q, m, zr, libm);
					orc_y cmp = orc_bp_dy(f, half.x, half.u, step / 2, q, m, zr, libm);
					double diff[3] = {yOut.x[0] - cmp.x[0], yOut.x[1] - cmp.x[1], yOut.x[2] - cmp.x[2]};
					double S = orc_norm(diff) / (step * (1 - 1 / 4.));
					/* Y(double) sets all three u components to S (PropagationBP.h:40-42) */
					r = sqrt((S * S + S * S) + S * S) / p->tolerance;
				} else {
					orc_ck_trystep(f, &yIn, &yOut, &yErr, step / C_LIGHT, q, E, zr, libm);
					r = orc_norm(yErr.u) / p->tolerance;
				}
				if (r > 1) {
					if (step == p->minStep)
						break;
					(*trials)++;
					newStep = step * 0.95 * orc_pow(r, -0.2, libm);
					newStep = orc_max(newStep, 0.1 * step);
					newStep = orc_max(newStep, p->minStep);
					step = newStep;
				} else {
					if (step != p->maxStep) {
						newStep = step * 0.95 * orc_pow(r, -0.2, libm);
						newStep = orc_min(newStep, 5 * step);
						newStep = orc_min(newStep, p->maxStep);
					}
					break;
				}
			}
		}
		/* setPosition; setDirection(getUnitVector()) normalises twice
		 * (src/module/PropagationBP.cpp:120-121, src/ParticleState.cpp:30-32) */
		double n1 = orc_norm(yOut.u);
		double u1[3] = {yOut.u[0] / n1, yOut.u[1] / n1, yOut.u[2] / n1};
		double n2 = orc_norm(u1);
		for (int a = 0; a < 3; a++) {
			pos[a] = yOut.x[a];
			dir[a] = u1[a] / n2;
		}
	}
	/* Candidate::setCurrentStep / setNextStep, src/Candidate.cpp:92-100 */
	s->current_step[i] = step;
	s->trajectory_length[i] += step;
	s->time[i] += step / C_LIGHT;
	s->next_step[i] = newStep;
	static const int defaultOrder[4] = {ORC_MOD_BOX, ORC_MOD_SPHERE, ORC_MOD_MAXLEN, ORC_MOD_OBSERVER};
	const int *order = p->nOrder > 0 ? p->order : defaultOrder;
	const int nOrder = p->nOrder > 0 ? p->nOrder : 4;
	for (int m = 0; m < nOrder; m++) {
		switch (order[m]) {
		case ORC_MOD_BOX: orc_periodic_box(p, s, i, pos); break;
		case ORC_MOD_SPHERE: orc_spherical_boundary(p, s, i, pos); break;
		case ORC_MOD_MAXLEN: orc_max_length(p, s, i); break;
		case ORC_MOD_OBSERVER: orc_observer(p, s, i, pos, dir); break;
		}
	}
}

long long orc_propagate(const orc_field *f, const orc_propagator *p, orc_state *s, size_t n, long long maxSteps,
                        int libm, int nthreads, long long *trials) {
	long long total = 0, rej = 0;
#ifdef _OPENMP
	if (nthreads <= 0)
		nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total, rej) num_threads(nthreads)
#endif
	for (size_t i = 0; i < n; i++) {
		long long local = 0;
		for (long long k = 0; k < maxSteps && s->active[i]; k++) {
			orc_process(f, p, s, i, libm, &rej);
			local++;
		}
		total += local;
	}
	if (trials)
		*trials = rej;
	return total;
}

/* ------------------------------------------------------------------ differentiable Boris model */
typedef struct {
	const float *g;
	int n[3];
	double go[3], sp[3], scale;
} orc_agrid;

/* corner c of the cell containing p: flat point index, weight, d(weight)/d(p) */
static void orc_acell(const orc_agrid *G, const double p[3], size_t idx[8], double w[8], double dw[8][3]) {
	int lo[3], hi[3];
	double f[3];
	for (int a = 0; a < 3; a++) {
		double r = (p[a] - G->go[a]) / G->sp[a];
		double fl = floor(r);
		f[a] = r - fl;
		long long i = (long long)fl % G->n[a];
		if (i < 0)
			i += G->n[a];
		lo[a] = (int)i;
		hi[a] = (int)((i + 1) % G->n[a]);
	}
	for (int c = 0; c < 8; c++) {
		/* corner order of Grid.h:572-579 */
		static const int bx[8] = {0, 1, 0, 0, 1, 0, 1, 1}, by[8] = {0, 0, 1, 0, 0, 1, 1, 1}, bz[8] = {0, 0, 0, 1, 1, 1, 0, 1};
		int b[3] = {bx[c], by[c], bz[c]};
		double ww[3], sg[3];
		for (int a = 0; a < 3; a++) {
			ww[a] = b[a] ? f[a] : 1.0 - f[a];
			sg[a] = b[a] ? 1.0 : -1.0;
		}
		w[c] = ww[0] * ww[1] * ww[2];
		dw[c][0] = sg[0] * ww[1] * ww[2] / G->sp[0];
		dw[c][1] = ww[0] * sg[1] * ww[2] / G->sp[1];
		dw[c][2] = ww[0] * ww[1] * sg[2] / G->sp[2];
		int X = b[0] ? hi[0] : lo[0], Yy = b[1] ? hi[1] : lo[1], Z = b[2] ? hi[2] : lo[2];
		idx[c] = ((size_t)X * G->n[1] + Yy) * G->n[2] + Z;
	}
}

static void orc_cross3(const double a[3], const double b[3], double o[3]) {
	o[0] = a[1] * b[2] - a[2] * b[1];
	o[1] = a[2] * b[0] - a[0] * b[2];
	o[2] = a[0] * b[1] - a[1] * b[0];
}
static double orc_dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

typedef struct {
	double x1[3], T[3], t[3], D, s2[3], v[3], up[3], r1, n1[3], r2, un[3];
	size_t idx[8];
	double w[8], dw[8][3];
} orc_astep;

static void orc_astep_fwd(const orc_agrid *G, double h, double kq, const double x[3], const double u[3], orc_astep *S,
                          double xo[3], double uo[3]) {
	for (int a = 0; a < 3; a++)
		S->x1[a] = x[a] + u[a] * (h / 2);
	orc_acell(G, S->x1, S->idx, S->w, S->dw);
	for (int a = 0; a < 3; a++) {
		double s = 0;
		for (int c = 0; c < 8; c++)
			s += S->w[c] * (double)G->g[3 * S->idx[c] + a];
		S->T[a] = s;
		S->t[a] = s * G->scale * kq;
	}
	S->D = 1 + orc_dot3(S->t, S->t);
	double ut[3], vs[3];
	for (int a = 0; a < 3; a++)
		S->s2[a] = 2 * S->t[a] / S->D;
	orc_cross3(u, S->t, ut);
	for (int a = 0; a < 3; a++)
		S->v[a] = u[a] + ut[a];
	orc_cross3(S->v, S->s2, vs);
	for (int a = 0; a < 3; a++)
		S->up[a] = u[a] + vs[a];
	S->r1 = sqrt(orc_dot3(S->up, S->up));
	for (int a = 0; a < 3; a++)
		S->n1[a] = S->up[a] / S->r1;
	S->r2 = sqrt(orc_dot3(S->n1, S->n1));
	for (int a = 0; a < 3; a++) {
		S->un[a] = S->n1[a] / S->r2;
		xo[a] = S->x1[a] + S->up[a] * (h / 2);
		uo[a] = S->un[a];
	}
}

void orc_adjoint_forward(const float *grid, int nx, int ny, int nz, const double origin[3], const double spacing[3],
                         double scale, double h, int nsteps, size_t n, const double *x0, const double *u0,
                         const double *kq, double *x_out, double *u_out) {
	orc_agrid G = {grid, {nx, ny, nz}, {0, 0, 0}, {spacing[0], spacing[1], spacing[2]}, scale};
	for (int a = 0; a < 3; a++)
		G.go[a] = origin[a] + spacing[a] / 2;
	for (size_t i = 0; i < n; i++) {
		double x[3] = {x0[3 * i], x0[3 * i + 1], x0[3 * i + 2]}, u[3] = {u0[3 * i], u0[3 * i + 1], u0[3 * i + 2]};
		orc_astep S;
		for (int s = 0; s < nsteps; s++)
			orc_astep_fwd(&G, h, kq[i], x, u, &S, x, u);
		for (int a = 0; a < 3; a++) {
			x_out[3 * i + a] = x[a];
			u_out[3 * i + a] = u[a];
		}
	}
}

void orc_adjoint_backward(const float *grid, int nx, int ny, int nz, const double origin[3], const double spacing[3],
                          double scale, double h, int nsteps, size_t n, const double *x0, const double *u0,
                          const double *kq, const double *gx_out, const double *gu_out, double *grad_grid,
                          double *grad_x0, double *grad_u0, double *grad_scale) {
	orc_agrid G = {grid, {nx, ny, nz}, {0, 0, 0}, {spacing[0], spacing[1], spacing[2]}, scale};
	for (int a = 0; a < 3; a++)
		G.go[a] = origin[a] + spacing[a] / 2;
	double *tape = (double *)malloc(sizeof(double) * 6 * ((size_t)nsteps + 1));
	double gs = 0;
	for (size_t i = 0; i < n; i++) {
		double x[3] = {x0[3 * i], x0[3 * i + 1], x0[3 * i + 2]}, u[3] = {u0[3 * i], u0[3 * i + 1], u0[3 * i + 2]};
		orc_astep S;
		for (int s = 0; s < nsteps; s++) {
			memcpy(tape + 6 * s, x, sizeof x);
			memcpy(tape + 6 * s + 3, u, sizeof u);
			orc_astep_fwd(&G, h, kq[i], x, u, &S, x, u);
		}
		double ax[3], au[3];
		for (int a = 0; a < 3; a++) {
			ax[a] = gx_out ? gx_out[3 * i + a] : 0.0;
			au[a] = gu_out[3 * i + a];
		}
		for (int s = nsteps - 1; s >= 0; s--) {
			double xo[3], uo[3];
			orc_astep_fwd(&G, h, kq[i], tape + 6 * s, tape + 6 * s + 3, &S, xo, uo);
			const double *us = tape + 6 * s + 3;
			/* u_next = n1/r2: adjoint of a/|a| is (g - n (n.g))/|a| */
			double an1[3], aup[3], c;
			c = orc_dot3(S.un, au);
			for (int a = 0; a < 3; a++)
				an1[a] = (au[a] - S.un[a] * c) / S.r2;
			c = orc_dot3(S.n1, an1);
			for (int a = 0; a < 3; a++)
				aup[a] = (an1[a] - S.n1[a] * c) / S.r1 + ax[a] * (h / 2);
			/* u' = u + v x s2 ; v = u + u x t ; s2 = 2 t / D */
			double av[3], as2[3], tav[3], at[3], aU[3];
			orc_cross3(S.s2, aup, av);
			orc_cross3(aup, S.v, as2);
			orc_cross3(S.t, av, tav);
			for (int a = 0; a < 3; a++)
				aU[a] = aup[a] + av[a] + tav[a];
			orc_cross3(av, us, at);
			double ts2 = orc_dot3(S.t, as2);
			for (int a = 0; a < 3; a++)
				at[a] += 2.0 / S.D * as2[a] - 4.0 / (S.D * S.D) * ts2 * S.t[a];
			/* t = kq * scale * T */
			double aT[3];
			for (int a = 0; a < 3; a++) {
				aT[a] = at[a] * kq[i] * scale;
				gs += S.T[a] * at[a] * kq[i];
			}
			double ax1[3] = {ax[0], ax[1], ax[2]};
			for (int cc = 0; cc < 8; cc++) {
				const float *gv = grid + 3 * S.idx[cc];
				double d = (double)gv[0] * aT[0] + (double)gv[1] * aT[1] + (double)gv[2] * aT[2];
				for (int a = 0; a < 3; a++) {
					ax1[a] += S.dw[cc][a] * d;
					grad_grid[3 * S.idx[cc] + a] += S.w[cc] * aT[a];
				}
			}
			for (int a = 0; a < 3; a++) {
				ax[a] = ax1[a];
				au[a] = aU[a] + ax1[a] * (h / 2);
			}
		}
		for (int a = 0; a < 3; a++) {
			grad_x0[3 * i + a] = ax[a];
			grad_u0[3 * i + a] = au[a];
		}
	}
	*grad_scale += gs;
	free(tape);
}

/* ------------------------------------------------------------------ turbulent grids */
/* Vector3<float>::getAngleTo / isParallelTo (include/crpropa/Vector3.h:172-187), all in float */
static int orc_parallel_to_111(const float e[3], float maxAngle) {
	const float n0[3] = {1.f, 1.f, 1.f};
	const float dot = e[0] * n0[0] + e[1] * n0[1] + e[2] * n0[2];
	const float rv = sqrtf(n0[0] * n0[0] + n0[1] * n0[1] + n0[2] * n0[2]);
	const float re = sqrtf(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
	const float cosd = dot / rv / re;
	float ang;
	if (cosd >= 1.)
		ang = 0;
	else if (cosd <= -1.)
		ang = (float)M_PI;
	else
		ang = acosf(cosd);
	return ang < maxAngle;
}

/* SimpleGridTurbulence::initTurbulence (src/magneticField/turbulentField/SimpleGridTurbulence.cpp:18-113)
 * and GridTurbulence::initTurbulence (GridTurbulence.cpp:29-124) up to the inverse FFT */
size_t orc_turbulence_coeffs(int n, double spacing, double lmin, double lmax, int kind, double alpha, double lbendover,
                             double sindex, double qindex, uint32_t seed, float *bk) {
	const size_t nn = (size_t)n, n2 = nn / 2 + 1, plane = nn * nn * n2;
	orc_rng rng;
	orc_rng_seed(&rng, seed);
	double *K = (double *)malloc(nn * sizeof(double));
	for (size_t i = 0; i < nn; i++)
		K[i] = (double)i / nn - (double)(i / (nn / 2));
	const double kMin = spacing / lmax, kMax = spacing / lmin;
	const double lambda = 1 / spacing * 2 * M_PI;
	size_t inside = 0;
	for (size_t ix = 0; ix < nn; ix++)
		for (size_t iy = 0; iy < nn; iy++)
			for (size_t iz = 0; iz < n2; iz++) {
				const size_t i = ix * nn * n2 + iy * n2 + iz;
				float ek[3] = {(float)K[ix], (float)K[iy], (float)K[iz]};
				const double k = sqrtf(ek[0] * ek[0] + ek[1] * ek[1] + ek[2] * ek[2]);
				float *o[3] = {bk + 2 * i, bk + 2 * (plane + i), bk + 2 * (2 * plane + i)};
				if ((k < kMin) || (k > kMax)) {
					for (int c = 0; c < 3; c++)
						o[c][0] = o[c][1] = 0;
					continue;
				}
				inside++;
				float e1[3], e2[3];
				if (orc_parallel_to_111(ek, 1e-3f)) {
					e1[0] = -1.f; e1[1] = 1.f; e1[2] = 0.f;
					e2[0] = 1.f; e2[1] = 1.f; e2[2] = -2.f;
				} else {
					/* e1 = n0 x ek, e2 = ek x e1 (Vector3::cross, Vector3.h:268-271) */
					const float n0[3] = {1.f, 1.f, 1.f};
					e1[0] = n0[1] * ek[2] - ek[1] * n0[2];
					e1[1] = n0[2] * ek[0] - ek[2] * n0[0];
					e1[2] = n0[0] * ek[1] - ek[0] * n0[1];
					e2[0] = ek[1] * e1[2] - e1[1] * ek[2];
					e2[1] = ek[2] * e1[0] - e1[2] * ek[0];
					e2[2] = ek[0] * e1[1] - e1[0] * ek[1];
				}
				const float r1 = sqrtf(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
				const float r2 = sqrtf(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
				for (int a = 0; a < 3; a++) {
					e1[a] /= r1;
					e2[a] /= r2;
				}
				const double theta = 2 * M_PI * orc_rng_rand(&rng);
				const float ct = (float)cos(theta), st = (float)sin(theta);
				float b[3];
				for (int a = 0; a < 3; a++)
					b[a] = e1[a] * ct + e2[a] * st;
				float amp;
				if (kind == ORC_TURB_SIMPLE) {
					amp = (float)(orc_rng_norm(&rng, 0.0, 1.0) * pow(k, alpha / 2));
				} else if (kind == ORC_TURB_GENERAL_POWERLAW) {
					/* SimpleTurbulenceSpectrum::energySpectrum (SimpleGridTurbulence.h:35-37) */
					amp = (float)sqrt(pow(k * lambda, -sindex - 2));
				} else {
					/* TurbulenceSpectrum::energySpectrum (TurbulentField.h:75-80) */
					const double kHat = k * lambda * lbendover;
					amp = (float)sqrt(pow(kHat, qindex) / pow(1.0 + kHat * kHat, (sindex + qindex) / 2.0 + 1.0));
				}
				for (int a = 0; a < 3; a++)
					b[a] *= amp;
				const double phase = 2 * M_PI * orc_rng_rand(&rng);
				const double cp = cos(phase), sp = sin(phase);
				for (int c = 0; c < 3; c++) {
					o[c][0] = (float)(b[c] * cp);
					o[c][1] = (float)(b[c] * sp);
				}
			}
	free(K);
	return inside;
}

void orc_cos_phases(const double *x, size_t n, double *quick, int32_t *decided, double *accurate) {
	for (size_t i = 0; i < n; i++) {
		decided[i] = crp_cos_quick(x[i], &quick[i]);
		accurate[i] = crp_cos_accurate(x[i]);
	}
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
	crp_philox_ctr c;
	for (int i = 0; i < 4; i++)
		c.v[i] = ctr[i];
	crp_philox_ctr o = crp_philox4x32_10(c, key[0], key[1]);
	for (int i = 0; i < 4; i++)
		out[i] = o.v[i];
}

/* the quick phase of exp: decided flag, quick value, accurate value (tests) */
void orc_exp_phases(const double *x, size_t n, int *decided, double *quick, double *accurate) {
	for (size_t i = 0; i < n; i++) {
		double q = 0;
		decided[i] = crp_exp_quick(x[i], &q);
		quick[i] = q;
		accurate[i] = crp_exp_accurate(x[i]);
	}
}

/* the quick phase of atan2: decided flag, quick value, accurate value (tests) */
void orc_atan2_phases(const double *y, const double *x, size_t n, int *decided, double *quick, double *accurate) {
	for (size_t i = 0; i < n; i++) {
		double q = 0;
		decided[i] = crp_atan2_quick(y[i], x[i], &q);
		quick[i] = q;
		accurate[i] = crp_atan2_accurate(y[i], x[i]);
	}
}

void orc_fast_batch(const double *x, size_t n, double *sin_out, double *cos_out, double *log_out) {
	for (size_t i = 0; i < n; i++) {
		crp_sincos_fast(x[i], sin_out + i, cos_out + i);
		log_out[i] = crp_log_fast(x[i]);
	}
}

void orc_exp_atan2_batch(const double *x, const double *y, size_t n, double *exp_out, double *atan2_out) {
	for (size_t i = 0; i < n; i++) {
		exp_out[i] = crp_exp_rn(x[i]);
		atan2_out[i] = crp_atan2_rn(y[i], x[i]);
	}
}

void orc_sin_log_batch(const double *x, size_t n, double *sin_out, double *log_out) {
	for (size_t i = 0; i < n; i++) {
		sin_out[i] = crp_sin_cr(x[i]);
		log_out[i] = x[i] > 0 ? crp_log_dd_rn(x[i]) : 0.0;
	}
}

/* ------------------------------------------------------------------ DiffusionSDE */
static const double KPC = 1e3 * (648000 / M_PI * 149597870700.0);

typedef struct {
	int kind;
	orc_rng *mt;
	crp_sde_rng ph;
} orc_sde_rng;

/* Random::rand / randExc / randDblExc (src/Random.cpp:80-98); one 53-bit uniform in [0, 1) each on
 * the Philox stream */
static double sde_rand(orc_sde_rng *r) { return r->kind == ORC_SDE_MT ? orc_rng_rand(r->mt) : crp_sde_uniform(&r->ph); }
static double sde_randexc(orc_sde_rng *r) {
	return r->kind == ORC_SDE_MT ? (double)orc_rng_randint(r->mt) * (1.0 / 4294967296.0) : crp_sde_uniform(&r->ph);
}
static double sde_randdblexc(orc_sde_rng *r) {
	return r->kind == ORC_SDE_MT ? ((double)orc_rng_randint(r->mt) + 0.5) * (1.0 / 4294967296.0)
	                             : crp_sde_uniform(&r->ph);
}
/* the reference's stream: glibc; the device's: the fast deterministic functions of crp_crmath.h */
static double sde_log(int k, double x) { return k == ORC_SDE_MT ? log(x) : crp_log_fast(x); }
static double sde_cos(int k, double x) {
	double s, c;
	if (k == ORC_SDE_MT)
		return cos(x);
	crp_sincos_fast(x, &s, &c);
	return c;
}
static double sde_sin(int k, double x) {
	double s, c;
	if (k == ORC_SDE_MT)
		return sin(x);
	crp_sincos_fast(x, &s, &c);
	return s;
}
static double sde_pow(int k, double x, double y) { return k == ORC_SDE_MT ? pow(x, y) : crp_pow_cr(x, y); }

/* Random::randNorm(0, 1) (src/Random.cpp:111-115) */
static double sde_randnorm(orc_sde_rng *r) {
	double rr = sqrt(-2.0 * sde_log(r->kind, 1.0 - sde_randdblexc(r))) * 1.0;
	double phi = 2.0 * 3.14159265358979323846264338328 * sde_randexc(r);
	return 0.0 + rr * sde_cos(r->kind, phi);
}

/* Random::randVector (src/Random.cpp:141-146) */
static void sde_randvector(orc_sde_rng *r, double v[3]) {
	double z = -1.0 + (1.0 - -1.0) * sde_rand(r);
	double t = -1.0 * M_PI + (M_PI - -1.0 * M_PI) * sde_rand(r);
	double rr = sqrt(1 - z * z);
	v[0] = rr * sde_cos(r->kind, t);
	v[1] = rr * sde_sin(r->kind, t);
	v[2] = z;
}

static void sde_cross(const double a[3], const double b[3], double o[3]) { /* Vector3.h:268-271 */
	o[0] = a[1] * b[2] - b[1] * a[2];
	o[1] = a[2] * b[0] - b[2] * a[0];
	o[2] = a[0] * b[1] - b[0] * a[1];
}
static double sde_r(const double v[3]) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
static void sde_unit(const double v[3], double o[3]) { /* getUnitVector: v / |v| */
	double r = sde_r(v);
	for (int a = 0; a < 3; a++)
		o[a] = v[a] / r;
}

/* Random::randConeVector(mean, pi/2) (src/Random.cpp:148-162) and Vector3::getRotated
 * (include/crpropa/Vector3.h:213-226).  On the reference's stream theta = acos(w) and cos/sin of
 * theta are taken; on the Philox stream cos theta = w, sin theta = sqrt(1 - w^2) directly. */
static void sde_randcone(orc_sde_rng *r, const double mean[3], double o[3]) {
	const int k = r->kind;
	double w = 1 + (sde_cos(k, M_PI / 2.) - 1) * sde_rand(r);
	double c, s;
	if (k == ORC_SDE_MT) {
		double theta = acos(w);
		c = cos(theta);
		s = sin(theta);
	} else {
		c = w;
		s = sqrt((1 - w) * (1 + w));
	}
	double rv[3], u[3];
	sde_randvector(r, rv);
	sde_cross(mean, rv, u);
	double ur = sde_r(u);
	if (ur != 0.)
		for (int a = 0; a < 3; a++)
			u[a] = u[a] / ur;
	double Rx[3] = {c + u[0] * u[0] * (1 - c), u[0] * u[1] * (1 - c) - u[2] * s, u[0] * u[2] * (1 - c) + u[1] * s};
	double Ry[3] = {u[1] * u[0] * (1 - c) + u[2] * s, c + u[1] * u[1] * (1 - c), u[1] * u[2] * (1 - c) - u[0] * s};
	double Rz[3] = {u[2] * u[0] * (1 - c) - u[1] * s, u[2] * u[1] * (1 - c) + u[0] * s, c + u[2] * u[2] * (1 - c)};
	o[0] = mean[0] * Rx[0] + mean[1] * Rx[1] + mean[2] * Rx[2];
	o[1] = mean[0] * Ry[0] + mean[1] * Ry[1] + mean[2] * Ry[2];
	o[2] = mean[0] * Rz[0] + mean[1] * Rz[1] + mean[2] * Rz[2];
}

/* DiffusionSDE::tryStep (src/module/DiffusionSDE.cpp:233-254): Cash-Karp along the unit field direction */
static void sde_trystep(const orc_field *f, const double in[3], double out[3], double err[3], double t, double z,
                        int libm) {
	double k[6][3];
	for (int a = 0; a < 3; a++)
		out[a] = in[a];
	for (int i = 0; i < 6; i++) {
		double y[3] = {in[0], in[1], in[2]};
		for (int j = 0; j < i; j++)
			for (int a = 0; a < 3; a++)
				y[a] += (k[j][a] * CK_A[i * 6 + j]) * t;
		double B[3], ub[3];
		orc_field_eval2(f, y, z, B, libm);
		sde_unit(B, ub);
		const double bi = CK_B[i], ei = CK_B[i] - CK_BS[i];
		for (int a = 0; a < 3; a++) {
			k[i][a] = ub[a] * C_LIGHT;
			out[a] += (k[i][a] * bi) * t;
			err[a] += ((k[i][a] * ei) * t) / KPC;
		}
	}
}

/* tanh: glibc's for the reference's stream, the deterministic one of crp_crmath.h for the device's */
static double orc_tanh(double x, int libm) { return libm == ORC_LIBM_CR ? crp_tanh_det(x) : tanh(x); }

/* One advection field's getField (src/advectionField/AdvectionField.cpp), par in the C ABI's order */
static void orc_adv_term(int kind, const double *par, const double pos[3], int libm, double w[3]) {
	w[0] = w[1] = w[2] = 0;
	if (kind == CRP_ADV_UNIFORM) { /* :33-35 */
		w[0] = par[0];
		w[1] = par[1];
		w[2] = par[2];
	} else if (kind == CRP_ADV_CONSTANT_SPHERICAL || kind == CRP_ADV_SPHERICAL) { /* :54-57, :101-109 */
		const double P[3] = {pos[0] - par[0], pos[1] - par[1], pos[2] - par[2]};
		const double R = sde_r(P);
		if (kind == CRP_ADV_SPHERICAL && R > par[3])
			return;
		double vr = par[3];
		if (kind == CRP_ADV_SPHERICAL) /* getV (:120-123) */
			vr = par[4] * (1 - orc_exp(-(orc_pow(R, par[6], libm) / par[5]), libm));
		for (int a = 0; a < 3; a++)
			w[a] = (P[a] / R) * vr;
	} else if (kind == CRP_ADV_CARTESIAN_SHOCK || kind == CRP_ADV_OBLIQUE_SHOCK) { /* :188-199, :360-372 */
		const int oblique = kind == CRP_ADV_OBLIQUE_SHOCK;
		const double x = pos[0];
		const double vUp = par[1];
		const double vDown = vUp / par[0];
		const double a = (vUp + vDown) * 0.5;
		const double b = (vUp - vDown) * 0.5;
		w[0] = a - b * orc_tanh(x / par[oblique ? 3 : 2], libm);
		if (oblique)
			w[1] = par[2];
	} else if (kind == CRP_ADV_SPHERICAL_SHOCK_1D) { /* :254-275 */
		const double r = sde_r(pos);
		const double rShock = par[0], vUp = par[1], lShock = par[3];
		const int coolUpstream = par[4] != 0;
		const double vDown = vUp / par[2];
		const double a = (vUp + vDown) * 0.5;
		const double b = (vUp - vDown) * 0.5;
		double v;
		if (coolUpstream) {
			if (r <= rShock)
				v = a - b * orc_tanh((r - rShock) / lShock, libm);
			else
				v = (a - b * orc_tanh((r - rShock) / lShock, libm)) * (rShock / r) * (rShock / r);
		} else
			v = (a - b * orc_tanh((r - rShock) / lShock, libm)) * (rShock / r) * (rShock / r);
		for (int k = 0; k < 3; k++)
			w[k] = (pos[k] / r) * v;
	} else if (kind == CRP_ADV_SPHERICAL_SHOCK) { /* :442-452 with g (:463-466) */
		const double R[3] = {pos[0] - par[0], pos[1] - par[1], pos[2] - par[2]};
		const double r = sde_r(R);
		const double eps = 2.2250738585072014e-308; /* Vector3::getPhi */
		const double phi = (fabs(R[0]) < eps && fabs(R[1]) < eps) ? 0.0 : orc_atan2(R[1], R[0], libm);
		/* getUnitVectorPhi (include/crpropa/Vector3.h:167-169): separate sin and cos calls in the
		 * reference's build (its samples show it); called through volatile pointers so that this
		 * compiler does not merge them into sincos either */
		double sphi, cphi;
		if (libm == ORC_LIBM_CR) { /* the device's e_phi: (-y, x) / sqrt(x^2 + y^2), no atan2/sin/cos */
			const int axis = fabs(R[0]) < eps && fabs(R[1]) < eps;
			const double rho = sqrt(R[0] * R[0] + R[1] * R[1]);
			sphi = axis ? 0.0 : R[1] / rho;
			cphi = axis ? 1.0 : R[0] / rho;
		} else {
			double (*volatile sinf_)(double) = sin, (*volatile cosf_)(double) = cos;
			sphi = sinf_(phi);
			cphi = cosf_(phi);
		}
		const double r_0 = par[3], v_0 = par[4], lambda = par[5], r_rot = par[6], v_phi = par[7];
		const double ga = (r - r_0) / lambda;
		const double g = 1. / (1 + orc_exp(-ga, libm));
		const double q = r_0 / (2 * r); /* pow(q, 2.): GCC writes q * q */
		const double v_r = v_0 * (1 + (q * q - 1) * g);
		const double v_p = v_phi * (r_rot / r);
		const double e_phi[3] = {-sphi, cphi, 0};
		for (int k = 0; k < 3; k++)
			w[k] = (R[k] / r) * v_r + e_phi[k] * v_p;
	}
}

/* the advection velocity of p's term list at n positions (AdvectionFieldList::getField) */
void orc_adv_eval(const orc_sde *p, const double *pos, size_t n, double *v, int libm) {
	for (size_t i = 0; i < n; i++) {
		double s[3] = {0, 0, 0};
		for (int t = 0; t < p->adv_n && t < 8; t++) {
			double w[3];
			orc_adv_term(p->adv_list_kind[t], p->adv_list_par[t], pos + 3 * i, libm, w);
			for (int a = 0; a < 3; a++)
				s[a] += w[a];
		}
		for (int a = 0; a < 3; a++)
			v[3 * i + a] = s[a];
	}
}

/* DiffusionSDE::process (src/module/DiffusionSDE.cpp:48-209) */
static void orc_sde_process(const orc_field *f, const orc_sde *p, orc_state *s, size_t i, orc_sde_rng *r) {
	const int k = r->kind, libm = k == ORC_SDE_MT ? ORC_LIBM_GLIBC : ORC_LIBM_CR;
	const double zr = s->redshift ? s->redshift[i] : 0.0; /* candidate->getRedshift() (DiffusionSDE.cpp:77) */
	double *pos = s->pos + 3 * i, *dir = s->dir + 3 * i;
	if (s->prev_pos)
		for (int a = 0; a < 3; a++) {
			s->prev_pos[3 * i + a] = pos[a];
			s->prev_dir[3 * i + a] = dir[a];
		}
	const double h = orc_clip(s->next_step[i], p->minStep, p->maxStep) / C_LIGHT;
	double lin[3] = {0, 0, 0};
	if (p->advection) {
		/* driftStep (DiffusionSDE.cpp:212-216): LinProp(0) += advField(pos) * h */
		double v[3] = {p->adv[0], p->adv[1], p->adv[2]};
		if (p->adv_n > 0) { /* AdvectionFieldList::getField: b(0.); b += field_i(pos) */
			v[0] = v[1] = v[2] = 0;
			for (int t = 0; t < p->adv_n && t < 8; t++) {
				double w[3];
				orc_adv_term(p->adv_list_kind[t], p->adv_list_par[t], pos, libm, w);
				for (int a = 0; a < 3; a++)
					v[a] += w[a];
			}
		} else if (p->adv_kind == 2 || p->adv_kind == 3) {
			const double P[3] = {pos[0] - p->adv_origin[0], pos[1] - p->adv_origin[1], pos[2] - p->adv_origin[2]};
			const double R = sde_r(P);
			double vr = p->adv_par[0]; /* ConstantSphericalAdvectionField: vWind */
			if (p->adv_kind == 3) /* SphericalAdvectionField::getV, zero beyond the radius */
				vr = R > p->adv_par[0] ? 0.0 : p->adv_par[1] * (1 - orc_exp(-(orc_pow(R, p->adv_par[3], libm) / p->adv_par[2]), libm));
			for (int a = 0; a < 3; a++)
				v[a] = (p->adv_kind == 3 && R > p->adv_par[0]) ? 0.0 : (P[a] / R) * vr;
		}
		for (int a = 0; a < 3; a++)
			lin[a] = 0 + v[a] * h;
	}
	if (s->charge[i] == 0) { /* rectilinear, with the advection drift */
		for (int a = 0; a < 3; a++)
			pos[a] = (pos[a] + lin[a]) + (dir[a] * h) * C_LIGHT;
		s->current_step[i] = h * C_LIGHT;
		s->trajectory_length[i] += h * C_LIGHT;
		s->time[i] += (h * C_LIGHT) / C_LIGHT;
		s->next_step[i] = p->maxStep;
		return;
	}
	const double rig = s->energy[i] / s->charge[i];
	/* calculateBTensor (:258-266) */
	const double D = p->scale * 6.1e24 * sde_pow(k, fabs(rig) / 4.0e9, p->alpha);
	const double B0 = sde_pow(k, 2 * D, 0.5), B4 = sde_pow(k, 2 * p->epsilon * D, 0.5), B8 = B4;
	double eta[3];
	for (int a = 0; a < 3; a++)
		eta[a] = sde_randnorm(r);
	const double TStep = B0 * eta[0], NStep = B4 * eta[1], BStep = B8 * eta[2];
	double in[3] = {pos[0], pos[1], pos[2]}, out[3], err[3];
	double propTime = TStep * sqrt(h) / C_LIGHT;
	int counter = 0;
	double rr;
	do {
		err[0] = err[1] = err[2] = 0;
		sde_trystep(f, in, out, err, propTime, zr, libm);
		rr = sde_r(err) / p->tolerance;
		propTime *= 0.5;
		counter += 1;
	} while (rr > 1 && fabs(propTime) >= p->minStep / C_LIGHT);
	const double stepNumber = ldexp(1.0, counter - 1); /* pow(2, counter - 1) */
	const double allowedTime = TStep * sqrt(h) / C_LIGHT / stepNumber;
	double start[3] = {in[0], in[1], in[2]};
	err[0] = err[1] = err[2] = 0;
	for (double j = 0; j < stepNumber; j++) {
		sde_trystep(f, start, out, err, allowedTime, zr, libm);
		start[0] = out[0];
		start[1] = out[1];
		start[2] = out[2];
	}
	double d[3] = {out[0] - in[0], out[1] - in[1], out[2] - in[2]}, T[3];
	sde_unit(d, T);
	const double tTest = sde_r(T);
	if (tTest != tTest) { /* no field: advection only, or rectilinear */
		for (int a = 0; a < 3; a++)
			pos[a] = p->advection ? in[a] + lin[a] : in[a] + (dir[a] * h) * C_LIGHT;
		s->current_step[i] = h * C_LIGHT;
		s->trajectory_length[i] += h * C_LIGHT;
		s->time[i] += (h * C_LIGHT) / C_LIGHT;
		s->next_step[i] = orc_clip(5 * h * C_LIGHT, p->minStep, p->maxStep);
		return;
	}
	double N[3] = {0, 0, 0}, Bv[3], tmp[3];
	while (sde_r(N) == 0.) {
		double rv[3];
		sde_randvector(r, rv);
		sde_cross(T, rv, N);
	}
	sde_unit(N, tmp);
	N[0] = tmp[0]; N[1] = tmp[1]; N[2] = tmp[2];
	sde_cross(T, N, tmp);
	sde_unit(tmp, Bv);
	double PO[3];
	const double sh = sqrt(h);
	for (int a = 0; a < 3; a++)
		PO[a] = (out[a] + lin[a]) + (N[a] * NStep + Bv[a] * BStep) * sh;
	if (isnan(sde_r(PO))) {
		s->active[i] = 0;
		return;
	}
	double D2[3];
	sde_randcone(r, T, D2);
	double dr = sde_r(D2);
	for (int a = 0; a < 3; a++) {
		pos[a] = PO[a];
		dir[a] = D2[a] / dr; /* ParticleState::setDirection */
	}
	s->current_step[i] = h * C_LIGHT;
	s->trajectory_length[i] += h * C_LIGHT;
	s->time[i] += (h * C_LIGHT) / C_LIGHT;
	s->next_step[i] = stepNumber > 1 ? h * sde_pow(k, stepNumber, -2.) * C_LIGHT : 4 * h * C_LIGHT;
}

/* The CPU baseline's form: ModuleList::run's OpenMP loop over the candidates with one MT19937 per
 * thread (Random::instance() is per thread, src/Random.cpp), seeded seed + thread number; glibc libm.
 * The numbers depend on the thread schedule, as the reference's do. */
long long orc_sde_propagate_threads(const orc_field *f, const orc_sde *p, orc_state *s, size_t n, long long maxSteps,
                                    int nthreads) {
	if (nthreads <= 0)
		nthreads = omp_get_max_threads();
	long long total = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : total)
	{
		orc_rng mt;
		orc_rng_seed(&mt, p->seed + (uint32_t)omp_get_thread_num());
		orc_sde_rng r;
		r.kind = ORC_SDE_MT;
		r.mt = &mt;
#pragma omp for schedule(dynamic, 16)
		for (size_t i = 0; i < n; i++) {
			for (long long k = 0; k < maxSteps && s->active[i]; k++) {
				orc_sde_process(f, p, s, i, &r);
				if (p->maxTrajectoryLength > 0) {
					const double len = s->trajectory_length[i];
					if (len >= p->maxTrajectoryLength)
						s->active[i] = 0;
					else
						s->next_step[i] = orc_min(s->next_step[i], p->maxTrajectoryLength - len);
				}
				total++;
			}
		}
	}
	return total;
}

long long orc_sde_propagate(const orc_field *f, const orc_sde *p, orc_state *s, size_t n, long long maxSteps,
                            uint32_t *sde_step, const uint64_t *sde_stream) {
	orc_rng mt;
	if (p->rng == ORC_SDE_MT)
		orc_rng_seed(&mt, p->seed);
	long long total = 0;
	for (size_t i = 0; i < n; i++) {
		for (long long k = 0; k < maxSteps && s->active[i]; k++) {
			orc_sde_rng r;
			r.kind = p->rng;
			r.mt = &mt;
			if (p->rng == ORC_SDE_PHILOX)
				crp_sde_rng_init(&r.ph, p->seed, sde_stream ? sde_stream[i] : (uint64_t)i,
				                 sde_step ? sde_step[i] : (uint32_t)k);
			orc_sde_process(f, p, s, i, &r);
			if (sde_step)
				sde_step[i]++;
			if (p->maxTrajectoryLength > 0) { /* MaximumTrajectoryLength::process */
				const double len = s->trajectory_length[i];
				if (len >= p->maxTrajectoryLength)
					s->active[i] = 0;
				else
					s->next_step[i] = orc_min(s->next_step[i], p->maxTrajectoryLength - len);
			}
			total++;
		}
	}
	return total;
}
