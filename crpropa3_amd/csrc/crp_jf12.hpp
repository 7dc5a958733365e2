// JF12Field on the device (included by crp_kernels.hip): the Jansson-Farrar 2012 Galactic field,
// src/magneticField/JF12Field.cpp:230-346, in the reference's FP64 operation order with the
// correctly rounded exp, atan2, sin and cos of crp_crmath.h; pow(x, 2.) is x * x, as GCC compiles
// the reference.
#pragma once

#include "crp_device.hpp"
#include "crpropa3_amd.h"

namespace crp {

// include/crpropa/Units.h: parsec = 648000 / pi * au, kiloparsec = 1e3 parsec
constexpr double kJKpc = 1e3 * (648000 / 3.14159265358979323846 * 149597870700.0);
constexpr double kJPi = 3.14159265358979323846;

// Grid<float>::closestValue (include/crpropa/Grid.h:383-405): nearest point, reflective or periodic
__device__ __forceinline__ float scalar_grid_closest(const GridDesc &g, const V3 &p) { return g.values[grid_closest(g, p)]; }

struct FieldJF12 {
	crp_jf12_params p;
	GridDesc striated; // Grid1f, one float per point (use_striated)
	GridDesc turbulent; // Grid3f (use_turbulent)

	// pow(x, 2.) as the reference's build evaluates it: GCC rewrites the call as x * x
	__device__ __forceinline__ static double sq(double x) { return x * x; }

	// JF12Field::logisticFunction (JF12Field.cpp:188-190)
	__device__ __forceinline__ double logistic(double x, double x0, double w) const {
		return 1. / (1. + crp_exp_rn(-2. * (__builtin_fabs(x) - x0) / w));
	}

	// radius where the arm through (r, phi) crosses the negative x-axis, tried for three windings
	__device__ __forceinline__ double r_negx(double r, double phi) const {
		double rn = r * crp_exp_rn(-(phi - kJPi) / p.tan90_minus_pitch);
		if (rn > p.r_arms[7])
			rn = r * crp_exp_rn(-(phi + kJPi) / p.tan90_minus_pitch);
		if (rn > p.r_arms[7])
			rn = r * crp_exp_rn(-(phi + 3 * kJPi) / p.tan90_minus_pitch);
		return rn;
	}

	// What getRegularField and getTurbulentStrength both compute from the position, once per evaluation
	// (the same operations, so the same bits): |pos|, the in-plane radius, the azimuth (Vector3::getPhi,
	// include/crpropa/Vector3.h:136-142) and the arm radius r_negx when r >= 5 kpc.
	struct Geo {
		double d, r, phi, rn;
		bool haveRn;
	};
	__device__ __forceinline__ Geo geo(const V3 &pos) const {
		Geo g;
		g.d = norm(pos);
		g.r = __builtin_sqrt(pos.x * pos.x + pos.y * pos.y);
		const double eps = 2.2250738585072014e-308; // numeric_limits<double>::min()
		g.phi = (__builtin_fabs(pos.x) < eps && __builtin_fabs(pos.y) < eps) ? 0.0 : crp_atan2_rn(pos.y, pos.x);
		g.haveRn = false;
		g.rn = 0;
		return g;
	}
	__device__ __forceinline__ double arm_radius(Geo &g) const {
		if (!g.haveRn) {
			g.rn = r_negx(g.r, g.phi);
			g.haveRn = true;
		}
		return g.rn;
	}

	// JF12Field::getRegularField with getDiskField, getToroidalHaloField, getXField (:192-300)
	__device__ V3 regular(const V3 &pos, Geo &gm) const {
		double bx = 0, by = 0, bz = 0;
		const double d = gm.d;
		if (d < 20 * kJKpc) {
			const double r = gm.r;
			const double phi = gm.phi;
			const double sinPhi = crp_sin_cr(phi), cosPhi = crp_cos_cr(phi);
			const double z = pos.z;
			// the disk's logistic factor, shared by the disk and the toroidal halo
			const bool needLf = p.use_disk || (p.use_toroidal_halo && (r * r + z * z > 1 * kJKpc * kJKpc));
			const double lfShared = needLf ? logistic(z, p.h_disk, p.w_disk) : 0.0;
			// disk
			double dx = 0, dy = 0;
			if (p.use_disk) {
				const double lfDisk = lfShared;
				if (r > 3 * kJKpc) {
					double bMag = 0;
					if (r < 5 * kJKpc) {
						bMag = p.b_ring * (5 * kJKpc / r) * (1 - lfDisk);
						dx += -bMag * sinPhi;
						dy += bMag * cosPhi;
					} else {
						const double rn = arm_radius(gm);
						for (int i = 7; i >= 0; i--)
							if (rn < p.r_arms[i])
								bMag = p.b_disk[i];
						bMag *= (5 * kJKpc / r) * (1 - lfDisk);
						dx += bMag * (p.sin_pitch * cosPhi - p.cos_pitch * sinPhi);
						dy += bMag * (p.sin_pitch * sinPhi + p.cos_pitch * cosPhi);
					}
				}
			}
			bx += dx;
			by += dy;
			bz += 0.0;
			// toroidal halo
			double hx = 0, hy = 0;
			if (p.use_toroidal_halo && (r * r + z * z > 1 * kJKpc * kJKpc)) {
				const double lfDisk = lfShared;
				double bMagH = crp_exp_rn(-__builtin_fabs(z) / p.z0) * lfDisk;
				if (z >= 0)
					bMagH *= p.b_north * (1 - logistic(r, p.r_north, p.w_halo));
				else
					bMagH *= p.b_south * (1 - logistic(r, p.r_south, p.w_halo));
				hx += -bMagH * sinPhi;
				hy += bMagH * cosPhi;
			}
			bx += hx;
			by += hy;
			bz += 0.0;
			// X field
			double xx = 0, xy = 0, xz = 0;
			if (p.use_x && (r * r + z * z > 1 * kJKpc * kJKpc)) {
				double bMagX, sinThetaX, cosThetaX, rp;
				const double rc = p.r_xc + __builtin_fabs(z) / p.tan_theta_x0;
				if (r < rc) {
					rp = r * p.r_xc / rc;
					bMagX = p.b_x * crp_exp_rn(-1 * rp / p.r_x) * sq(p.r_xc / rc);
					double thetaX = crp_atan2_rn(__builtin_fabs(z), (r - rp));
					if (z == 0)
						thetaX = kJPi / 2.;
					sinThetaX = crp_sin_cr(thetaX);
					cosThetaX = crp_cos_cr(thetaX);
				} else {
					rp = r - __builtin_fabs(z) / p.tan_theta_x0;
					bMagX = p.b_x * crp_exp_rn(-rp / p.r_x) * (rp / r);
					sinThetaX = p.sin_theta_x0;
					cosThetaX = p.cos_theta_x0;
				}
				const double zsign = z < 0 ? -1 : 1;
				xx += zsign * bMagX * cosThetaX * cosPhi;
				xy += zsign * bMagX * cosThetaX * sinPhi;
				xz += bMagX * sinThetaX;
			}
			bx += xx;
			by += xy;
			bz += xz;
		}
		return v3(bx, by, bz);
	}

	// JF12Field::getTurbulentStrength (:307-331)
	__device__ double turbulent_strength(const V3 &pos, Geo &gm) const {
		if (gm.d > 20 * kJKpc)
			return 0;
		const double r = gm.r;
		double bDisk = 0;
		if (r < 5 * kJKpc) {
			bDisk = p.b_disk_turb5;
		} else {
			const double rn = arm_radius(gm);
			for (int i = 7; i >= 0; i--)
				if (rn < p.r_arms[i])
					bDisk = p.b_disk_turb[i];
			bDisk *= (5 * kJKpc) / r;
		}
		bDisk *= crp_exp_rn(-0.5 * sq(pos.z / p.z_disk_turb));
		const double bHalo = p.b_halo_turb * crp_exp_rn(-r / p.r_halo_turb) *
		                     crp_exp_rn(-0.5 * sq(pos.z / p.z_halo_turb));
		return __builtin_sqrt(sq(bDisk) + sq(bHalo));
	}

	// JF12Field::getField (:337-346); the turbulent term is Vector3f * (float)strength in float.
	// AnyInterp: the turbulent grid may be read tricubic or nearest-neighbour (Grid::interpolate).
	template <bool AnyInterp> __device__ __forceinline__ V3 eval_impl(const V3 &pos) const {
		double bx = 0, by = 0, bz = 0;
		Geo gm = geo(pos);
		if (p.use_turbulent) {
			const V3 g = AnyInterp ? grid_interpolate(turbulent, pos) : FieldGrid{turbulent}.eval(pos);
			const float s = (float)turbulent_strength(pos, gm);
			bx += (double)((float)g.x * s);
			by += (double)((float)g.y * s);
			bz += (double)((float)g.z * s);
		}
		if (p.use_striated) {
			const V3 b = regular(pos, gm);
			const double f = 1. + p.sqrt_beta * (double)scalar_grid_closest(striated, pos);
			bx += b.x * f;
			by += b.y * f;
			bz += b.z * f;
		} else if (p.use_regular) {
			const V3 b = regular(pos, gm);
			bx += b.x;
			by += b.y;
			bz += b.z;
		}
		return v3(bx, by, bz);
	}
	// Out of line: one copy serves the six stages of a Cash-Karp step.
	__device__ __attribute__((noinline)) V3 eval(const V3 &pos) const { return eval_impl<false>(pos); }
	__device__ float3 evalf(const V3 &pos) const {
		const V3 b = eval(pos);
		return make_float3((float)b.x, (float)b.y, (float)b.z);
	}
	// the same with a turbulent grid of any interpolation type; its own copy, so that the trilinear
	// one keeps its register allocation
	__device__ __attribute__((noinline)) V3 eval_any(const V3 &pos) const { return eval_impl<true>(pos); }
};

// JF12Field whose turbulent Grid3f is read tricubic or nearest-neighbour
struct FieldJF12Any {
	FieldJF12 f;
	__device__ __forceinline__ V3 eval(const V3 &pos) const { return f.eval_any(pos); }
	__device__ __forceinline__ float3 evalf(const V3 &pos) const {
		const V3 b = f.eval_any(pos);
		return make_float3((float)b.x, (float)b.y, (float)b.z);
	}
};

} // namespace crp
