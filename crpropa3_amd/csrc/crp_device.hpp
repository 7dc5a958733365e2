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

// Vector3<T>::getR: sqrt((x*x + y*y) + z*z)
__device__ __forceinline__ double norm(const V3 &v) { return __builtin_sqrt((v.x * v.x + v.y * v.y) + v.z * v.z); }

// ---------------------------------------------------------------- fields
// Parameters as the host computes them from the reference's field objects.
struct GridDesc {
	const float *values; // AoS float3 per point, index (ix*Ny + iy)*Nz + iz
	int nx, ny, nz;
	int reflective, clip;
	double gridOrigin[3]; // origin + spacing/2 (Grid.h:225-228)
	double spacing[3];
	double origin[3], edge[3]; // clipVolume box: origin, origin + N*spacing (Grid.h:320-327)
};

struct PlaneWaveDesc {
	int nm;
	const double *k;     // [nm]
	const double *beta;  // [nm]
	const double *kappa; // [3*nm]
	const double *xiA;   // [3*nm]: xi[i] * Ak[i], per component (the first product of getField)
};

struct FieldUniform {
	double b[3];
	__device__ __forceinline__ V3 eval(const V3 &) const { return v3(b[0], b[1], b[2]); }
};

struct FieldNone {
	__device__ __forceinline__ V3 eval(const V3 &) const { return v3(0.0, 0.0, 0.0); }
};

__device__ __forceinline__ void periodic_clamp(double x, int n, int &lo, int &hi, double &frac) {
	double fl = __builtin_floor(x);
	lo = (((int)fl % n) + n) % n;
	hi = (lo + 1) % n;
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
		double rx = (p.x - g.gridOrigin[0]) / g.spacing[0];
		double ry = (p.y - g.gridOrigin[1]) / g.spacing[1];
		double rz = (p.z - g.gridOrigin[2]) / g.spacing[2];
		int x0, x1, y0, y1, z0, z1;
		double fx, fy, fz;
		if (g.reflective) {
			reflective_clamp(rx, g.nx, x0, x1, fx);
			reflective_clamp(ry, g.ny, y0, y1, fy);
			reflective_clamp(rz, g.nz, z0, z1, fz);
		} else {
			periodic_clamp(rx, g.nx, x0, x1, fx);
			periodic_clamp(ry, g.ny, y0, y1, fy);
			periodic_clamp(rz, g.nz, z0, z1, fz);
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
};

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
		double rx = (p.x - g.gridOrigin[0]) / g.spacing[0];
		double ry = (p.y - g.gridOrigin[1]) / g.spacing[1];
		double rz = (p.z - g.gridOrigin[2]) / g.spacing[2];
		int x0, x1, y0, y1, z0, z1;
		double fx, fy, fz;
		periodic_clamp(rx, g.nx, x0, x1, fx);
		periodic_clamp(ry, g.ny, y0, y1, fy);
		periodic_clamp(rz, g.nz, z0, z1, fz);
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
};

struct FieldPlaneWave {
	PlaneWaveDesc w;

	__device__ __forceinline__ V3 eval(const V3 &p) const {
		double bx = 0.0, by = 0.0, bz = 0.0;
		for (int i = 0; i < w.nm; i++) {
			const double kx = w.kappa[3 * i], ky = w.kappa[3 * i + 1], kz = w.kappa[3 * i + 2];
			double z_ = (p.x * kx + p.y * ky) + p.z * kz;
			double c = crp_cos_cr(w.k[i] * z_ + w.beta[i]);
			bx += w.xiA[3 * i] * c;
			by += w.xiA[3 * i + 1] * c;
			bz += w.xiA[3 * i + 2] * c;
		}
		return v3(bx, by, bz);
	}
};

// ---------------------------------------------------------------- Boris push
struct Y {
	V3 x, u;
};

// PropagationBP::dY (src/module/PropagationBP.cpp:19-39)
template <class F>
__device__ __forceinline__ Y bp_dy(const F &field, const V3 &pos0, const V3 &dir, double step, double q, double m) {
	V3 pos = v3(pos0.x + (dir.x * step) / 2., pos0.y + (dir.y * step) / 2., pos0.z + (dir.z * step) / 2.);
	V3 B = field.eval(pos);
	V3 t = v3(((((B.x * q) / 2) / m) * step) / kCLight, ((((B.y * q) / 2) / m) * step) / kCLight,
	          ((((B.z * q) / 2) / m) * step) / kCLight);
	double tt = (t.x * t.x + t.y * t.y) + t.z * t.z;
	double den = 1 + tt;
	V3 s = v3((t.x * 2) / den, (t.y * 2) / den, (t.z * 2) / den);
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
// Tableau of src/module/PropagationCK.cpp:11-26 (textbook Cash-Karp constants).
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

// PropagationCK::dYdt (src/module/PropagationCK.cpp:51-61); qcE = (q*c)/E
template <class F>
__device__ __forceinline__ Y ck_dydt(const F &field, const Y &y, double qcE) {
	Y d;
	double n = norm(y.u);
	d.x = v3((y.u.x / n) * kCLight, (y.u.y / n) * kCLight, (y.u.z / n) * kCLight);
	V3 B = field.eval(y.x);
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
// new phase point in `out` and |err.u| in *errU.
template <class F>
__device__ __forceinline__ void ck_trystep(const F &field, const Y &y, double h, double qcE, Y &out, double &errU) {
	Y err;
	err.x = v3(0.0, 0.0, 0.0);
	err.u = v3(0.0, 0.0, 0.0);
	out = y;
	const double e0 = CK::b0 - CK::s0, e1 = CK::b1 - CK::s1, e2 = CK::b2 - CK::s2, e3 = CK::b3 - CK::s3,
	             e4 = CK::b4 - CK::s4, e5 = CK::b5 - CK::s5;
	Y k0, k1, k2, k3, k4, k5, yn;
	k0 = ck_dydt(field, y, qcE);
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
	errU = norm(err.u);
}

} // namespace crp
