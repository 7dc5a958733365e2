// DiffusionSDE on the device (included by crp_kernels.hip): stochastic
// transport of cosmic rays along the magnetic field lines,
// src/module/DiffusionSDE.cpp:48-266.
//
// One particle per lane, persistent, with the queue refill of k_propagate.
// A step of DiffusionSDE::process is a variable number of Cash-Karp steps
// along the unit field direction (tryStep): trials with the time halved until
// the error is below the tolerance, then 2^(trials-1) sub-steps.  Each loop
// iteration of the kernel runs ONE tryStep for every busy lane, whichever
// phase (trial or sub-step) the lane is in, so the field evaluations stay
// converged across the wave; the random draws at the start and end of a step
// are the divergent part.
//
// Bit-exact shortcuts (the same floating-point operations, not repeated):
//  - every trial of a step starts at the same point: the first stage's field
//    value is evaluated once per step;
//  - the first sub-step uses the last trial's time (propTime halved k-1 times
//    equals propTime / 2^(k-1) unless it underflows; the bits are compared)
//    and start point, so its result is the last trial's: it is not recomputed.
//
// Random numbers: Philox4x32-10 per (particle stream, step) (crp_philox.h);
// randNorm, randVector and randConeVector as src/Random.cpp:111-162 with the
// fast deterministic log and sin/cos of crp_crmath.h (about one ulp; the same
// bits in the oracle), and cos/sin of acos(w) taken as w and sqrt(1 - w^2).
#pragma once

#include "crp_philox.h"
#include "crp_propagate.hpp"

namespace crp {

// include/crpropa/Units.h: au, parsec = 648000 / pi * au, kpc = 1e3 parsec
constexpr double kKpc = 1e3 * (648000 / 3.14159265358979323846 * 149597870700.0);
constexpr double kInvKpc = 1.0 / kKpc;
// a step whose trials end past 2^kSdeMaxHalvings sub-steps stops the particle (CTR_SDE_LIMIT);
// Params::sdeMaxHalvings holds the limit of a run (this value unless a test lowers it)
constexpr int kSdeMaxHalvings = 20;

// B / |B| * c_light (Vector3::getUnitVector, then the product)
__device__ __forceinline__ V3 sde_unit_c(const V3 &b) {
	const double r = norm(b), ir = 1 / r;
	return v3(div_rn(b.x, r, ir) * kCLight, div_rn(b.y, r, ir) * kCLight, div_rn(b.z, r, ir) * kCLight);
}

__device__ __forceinline__ V3 sde_unit(const V3 &b) {
	const double r = norm(b), ir = 1 / r;
	return v3(div_rn(b.x, r, ir), div_rn(b.y, r, ir), div_rn(b.z, r, ir));
}

// Vector3::cross (include/crpropa/Vector3.h:268-271)
__device__ __forceinline__ V3 sde_cross(const V3 &a, const V3 &b) {
	return v3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}

// DiffusionSDE::tryStep (src/module/DiffusionSDE.cpp:233-254) from x with time t; k0 = the first
// stage's unit field times c at x.  out = x + sum k_i b_i t, err = sum (k_i (b_i - b*_i)) t / kpc.
template <class F>
__device__ __forceinline__ void sde_trystep(const F &field, const V3 &x, const V3 &k0, double t, V3 &out, V3 &err) {
	const double e0 = CK::b0 - CK::s0, e1 = CK::b1 - CK::s1, e2 = CK::b2 - CK::s2, e3 = CK::b3 - CK::s3,
	             e4 = CK::b4 - CK::s4, e5 = CK::b5 - CK::s5;
	out = x;
	err = v3(0.0, 0.0, 0.0);
	auto acc = [&](const V3 &k, double b, double e) {
		out.x += (k.x * b) * t;
		out.y += (k.y * b) * t;
		out.z += (k.z * b) * t;
		err.x += div_rn((k.x * e) * t, kKpc, kInvKpc);
		err.y += div_rn((k.y * e) * t, kKpc, kInvKpc);
		err.z += div_rn((k.z * e) * t, kKpc, kInvKpc);
	};
	// y_n = x + sum_j (k_j a_ij) t, in the reference's order
	auto ax = [&](V3 &y, const V3 &k, double a) {
		y.x += (k.x * a) * t;
		y.y += (k.y * a) * t;
		y.z += (k.z * a) * t;
	};
	acc(k0, CK::b0, e0);
	V3 y = x;
	ax(y, k0, CK::a10);
	const V3 k1 = sde_unit_c(field.eval(y));
	acc(k1, CK::b1, e1);
	y = x;
	ax(y, k0, CK::a20);
	ax(y, k1, CK::a21);
	const V3 k2 = sde_unit_c(field.eval(y));
	acc(k2, CK::b2, e2);
	y = x;
	ax(y, k0, CK::a30);
	ax(y, k1, CK::a31);
	ax(y, k2, CK::a32);
	const V3 k3 = sde_unit_c(field.eval(y));
	acc(k3, CK::b3, e3);
	y = x;
	ax(y, k0, CK::a40);
	ax(y, k1, CK::a41);
	ax(y, k2, CK::a42);
	ax(y, k3, CK::a43);
	const V3 k4 = sde_unit_c(field.eval(y));
	acc(k4, CK::b4, e4);
	y = x;
	ax(y, k0, CK::a50);
	ax(y, k1, CK::a51);
	ax(y, k2, CK::a52);
	ax(y, k3, CK::a53);
	ax(y, k4, CK::a54);
	const V3 k5 = sde_unit_c(field.eval(y));
	acc(k5, CK::b5, e5);
}

// Random::randNorm(0, 1) (src/Random.cpp:111-115): randDblExc, then randExc
__device__ __forceinline__ double sde_randnorm(crp_sde_rng &r) {
	const double rr = __builtin_sqrt(-2.0 * crp_log_fast(1.0 - crp_sde_uniform(&r))) * 1.0;
	const double phi = 2.0 * 3.14159265358979323846264338328 * crp_sde_uniform(&r);
	double sp, cp;
	crp_sincos_fast(phi, &sp, &cp);
	return 0.0 + rr * cp;
}

// Random::randVector (src/Random.cpp:141-146)
__device__ __forceinline__ V3 sde_randvector(crp_sde_rng &r) {
	const double pi = 3.14159265358979323846;
	const double z = -1.0 + (1.0 - -1.0) * crp_sde_uniform(&r);
	const double t = -1.0 * pi + (pi - -1.0 * pi) * crp_sde_uniform(&r);
	const double rr = __builtin_sqrt(1 - z * z);
	double st, ct;
	crp_sincos_fast(t, &st, &ct);
	return v3(rr * ct, rr * st, z);
}

// Random::randConeVector(mean, pi/2) and Vector3::getRotated (src/Random.cpp:148-162,
// include/crpropa/Vector3.h:213-226); cosHalfPi = cos(pi/2) rounded
__device__ __forceinline__ V3 sde_randcone(crp_sde_rng &r, const V3 &mean, double cosHalfPi) {
	const double w = 1 + (cosHalfPi - 1) * crp_sde_uniform(&r);
	const double c = w, s = __builtin_sqrt((1 - w) * (1 + w));
	const V3 rv = sde_randvector(r);
	V3 u = sde_cross(mean, rv);
	const double ur = norm(u);
	if (ur != 0.) {
		const double iu = 1 / ur;
		u = v3(div_rn(u.x, ur, iu), div_rn(u.y, ur, iu), div_rn(u.z, ur, iu));
	}
	const double oc = 1 - c;
	const V3 rx = v3(c + u.x * u.x * oc, u.x * u.y * oc - u.z * s, u.x * u.z * oc + u.y * s);
	const V3 ry = v3(u.y * u.x * oc + u.z * s, c + u.y * u.y * oc, u.y * u.z * oc - u.x * s);
	const V3 rz = v3(u.z * u.x * oc - u.y * s, u.z * u.y * oc + u.x * s, c + u.z * u.z * oc);
	return v3(mean.x * rx.x + mean.y * rx.y + mean.z * rx.z, mean.x * ry.x + mean.y * ry.y + mean.z * ry.z,
	          mean.x * rz.x + mean.y * rz.y + mean.z * rz.z);
}

// The advection velocity at pos: the terms' getField (src/advectionField/AdvectionField.cpp) summed
// in order from 0, as AdvectionFieldList::getField does (:12-17); for one term the sum is the field
// itself (0 + v differs from v only in the sign of a zero, which the drift 0 + v h removes).  The
// terms are read as scalar loads (every lane reads the same term).  Out of line: the transcendental
// functions of the shock profiles stay out of the transport loop's registers.
typedef __attribute__((address_space(4))) const SdeAdvTerm SdeAdvTermConst;

__device__ __forceinline__ V3 sde_advection_velocity(const SdeAdvTerm *terms, int n, const V3 pos) {
	SdeAdvTermConst *T = (SdeAdvTermConst *)terms;
	double vx = 0, vy = 0, vz = 0;
	for (int t = 0; t < n; t++) {
		const int kind = T[t].kind;
		double w0 = 0, w1 = 0, w2 = 0;
		if (kind == CRP_ADV_UNIFORM) { // UniformAdvectionField (:33-35)
			w0 = T[t].par[0];
			w1 = T[t].par[1];
			w2 = T[t].par[2];
		} else if (kind == CRP_ADV_CONSTANT_SPHERICAL || kind == CRP_ADV_SPHERICAL) {
			// ConstantSphericalAdvectionField (:54-57): unit vector of pos - origin times vWind;
			// SphericalAdvectionField (:101-109, 120-123): vMax (1 - exp(-R^alpha / tau)), 0 beyond the radius
			const V3 d = v3(pos.x - T[t].par[0], pos.y - T[t].par[1], pos.z - T[t].par[2]);
			const double R = norm(d);
			double vr = T[t].par[3];
			const bool outside = kind == CRP_ADV_SPHERICAL && R > T[t].par[3];
			if (kind == CRP_ADV_SPHERICAL && !outside)
				vr = T[t].par[4] * (1 - crp_exp_rn(-(crp_pow_cr(R, T[t].par[6]) / T[t].par[5])));
			w0 = outside ? 0.0 : (d.x / R) * vr;
			w1 = outside ? 0.0 : (d.y / R) * vr;
			w2 = outside ? 0.0 : (d.z / R) * vr;
		} else if (kind == CRP_ADV_CARTESIAN_SHOCK || kind == CRP_ADV_OBLIQUE_SHOCK) {
			// OneDimensionalCartesianShock (:188-199), ObliqueAdvectionShock (:360-372):
			// v_x = a - b tanh(x / lShock), a, b = (vUp +- vUp / compressionRatio) / 2; v_y = vY (oblique)
			const bool oblique = kind == CRP_ADV_OBLIQUE_SHOCK;
			const double vUp = T[t].par[1];
			const double vDown = vUp / T[t].par[0];
			const double a = (vUp + vDown) * 0.5;
			const double b = (vUp - vDown) * 0.5;
			w0 = a - b * crp_tanh_det(pos.x / T[t].par[oblique ? 3 : 2]);
			w1 = oblique ? T[t].par[2] : 0.0;
		} else if (kind == CRP_ADV_SPHERICAL_SHOCK_1D) {
			// OneDimensionalSphericalShock (:254-275): radial, a - b tanh((r - rShock) / lShock), times
			// (rShock / r)^2 outside the shock or without upstream cooling
			const double r = norm(pos);
			const double rS = T[t].par[0], vUp = T[t].par[1];
			const double vDown = vUp / T[t].par[2];
			const double a = (vUp + vDown) * 0.5;
			const double b = (vUp - vDown) * 0.5;
			const double th = a - b * crp_tanh_det((r - rS) / T[t].par[3]);
			const double v = (T[t].par[4] != 0 && r <= rS) ? th : (th * (rS / r)) * (rS / r);
			w0 = (pos.x / r) * v;
			w1 = (pos.y / r) * v;
			w2 = (pos.z / r) * v;
		} else if (kind == CRP_ADV_SPHERICAL_SHOCK) {
			// SphericalAdvectionShock (:442-452, 463-466): v_r e_r + v_phi (r_rot / r) e_phi with
			// v_r = v0 (1 + ((r0 / 2r)^2 - 1) g(r)), g = 1 / (1 + exp(-(r - r0) / lambda)); pow(x, 2.) is x * x
			// e_phi = (-sin phi, cos phi, 0) with phi = atan2(y, x) (Vector3::getUnitVectorPhi) is taken as
			// (-y, x, 0) / sqrt(x^2 + y^2): within an ulp or two of the reference's, without atan2, sin and cos
			const V3 R = v3(pos.x - T[t].par[0], pos.y - T[t].par[1], pos.z - T[t].par[2]);
			const double r = norm(R);
			const double eps = 2.2250738585072014e-308; // Vector3::getPhi: numeric_limits<double>::min()
			const bool axis = __builtin_fabs(R.x) < eps && __builtin_fabs(R.y) < eps;
			const double rho = __builtin_sqrt(R.x * R.x + R.y * R.y);
			const double ex = axis ? -0.0 : -(R.y / rho), ey = axis ? 1.0 : R.x / rho;
			const double r0 = T[t].par[3], v0 = T[t].par[4], lambda = T[t].par[5];
			const double g = 1. / (1 + crp_exp_rn(-((r - r0) / lambda)));
			const double q = r0 / (2 * r);
			const double vr = v0 * (1 + (q * q - 1) * g);
			const double vp = T[t].par[7] * (T[t].par[6] / r);
			w0 = (R.x / r) * vr + ex * vp;
			w1 = (R.y / r) * vr + ey * vp;
			w2 = (R.z / r) * vr + 0.0 * vp;
		}
		vx += w0;
		vy += w1;
		vz += w2;
	}
	return v3(vx, vy, vz);
}

// driftStep (src/module/DiffusionSDE.cpp:212-216): lin = 0 + v(pos) h with the advection field's
// velocity at the step's start
__device__ __forceinline__ void sde_drift(const Params &P, const V3 &pos, double h, double lin[3]) {
	const V3 v = sde_advection_velocity(P.sdeAdvTerms, P.sdeAdvection, pos);
	lin[0] = 0 + v.x * h;
	lin[1] = 0 + v.y * h;
	lin[2] = 0 + v.z * h;
}

enum SdePhase { SDE_BEGIN = 0, SDE_TRIAL = 1, SDE_SUB = 2 };

// MODS: the periodic box, spherical boundary and observer are compiled in (the length limit always is);
// without them the kernel keeps fewer registers
template <class F, bool MODS>
__global__ void __launch_bounds__(256) k_sde(F field0, Params P, Soa S, unsigned long long n,
                                             unsigned long long *__restrict__ ctr) {
	const int lane = __lane_id();
	F field = field0; // the lane's copy: a field that reads the candidate's redshift binds it at refill
	const double cosHalfPi = crp_cos_cr(3.14159265358979323846 / 2.);
	const double minTime = P.minStep / kCLight; // minStep / c_light of the trial loop's test

	bool have = false, exhausted = false;
	unsigned idx = 0;
	long long steps = 0;
	int phase = SDE_BEGIN;
	V3 pos = v3(0, 0, 0), dir = v3(0, 0, 0);
	double q = 0, next = 0, B0 = 0, B4 = 0;
	uint32_t rstep = 0;
	uint64_t stream = 0;
	// the step in progress
	double h = 0, sh = 0, TStep = 0, NStep = 0, BStep = 0, propTime = 0, lastT = 0, allowed = 0;
	double stepNumber = 0, j = 0;
	int counter = 0;
	V3 k0 = v3(0, 0, 0), start = v3(0, 0, 0), out = v3(0, 0, 0);
	unsigned nAcc = 0, nRej = 0, nEval = 0, nFin = 0, nPause = 0, nLimit = 0;
	extern __shared__ double crp_cold[];
	double *const cw = crp_cold + (threadIdx.x >> 6) * (kColdDoubles * 64) + (threadIdx.x & 63);
	double *const cPPx = cw, *const cPPy = cw + 64, *const cPPz = cw + 128;
	double *const cPDx = cw + 192, *const cPDy = cw + 256, *const cPDz = cw + 320;
	double *const cCur = cw + 384, *const cLen = cw + 448, *const cTim = cw + 512;
	double *const cDet = cw + 576; // observer detection count (small integer held as a double)

	for (;;) {
		if (nEval > 0x80000000u) {
			atomicAdd(&ctr[CTR_ACCEPTED], (unsigned long long)nAcc);
			atomicAdd(&ctr[CTR_REJECTED], (unsigned long long)nRej);
			atomicAdd(&ctr[CTR_EVALS], (unsigned long long)nEval);
			nAcc = nRej = nEval = 0;
		}
		// ---- refill lanes without a particle from the queue
		const bool need = !have && !exhausted;
		const unsigned long long m = __ballot(need);
		if (m) {
			const int leader = __ffsll((long long)m) - 1;
			unsigned long long base = 0;
			if (lane == leader)
				base = atomicAdd(&ctr[CTR_QUEUE], (unsigned long long)__popcll(m));
			base = __shfl(base, leader);
			if (need) {
				const unsigned long long my = base + (unsigned long long)__popcll(m & ((1ull << lane) - 1ull));
				if (my < n) {
					idx = S.order[my];
					have = S.active[idx] != 0;
					if (have) {
						pos = v3(S.px[idx], S.py[idx], S.pz[idx]);
						dir = v3(S.dx[idx], S.dy[idx], S.dz[idx]);
						q = S.q[idx];
						if (UsesRedshift<F>::value)
							bind_redshift(field, S.z[idx]); // DiffusionSDE.cpp:77, 245
						next = S.next[idx];
						*cCur = S.cur[idx];
						*cLen = S.len[idx];
						*cTim = S.time[idx];
						rstep = S.sdeStep[idx];
						if (MODS && P.nObs > 0)
							*cDet = (double)S.det[idx];
						stream = S.sdeStream[idx];
						if (q != 0) {
							// calculateBTensor (src/module/DiffusionSDE.cpp:258-266): constant per particle;
							// pow(x, 0.5) correctly rounded is sqrt(x)
							const double rig = S.E[idx] / q;
							const double D = P.sdeScale * 6.1e24 * crp_pow_cr(__builtin_fabs(rig) / 4.0e9, P.sdeAlpha);
							B0 = __builtin_sqrt(2 * D);
							B4 = __builtin_sqrt(2 * P.sdeEpsilon * D);
						}
						steps = 0;
						phase = SDE_BEGIN;
					}
				} else {
					exhausted = true;
				}
			}
		}
		if (__ballot(have) == 0) {
			if (__ballot(!exhausted) == 0)
				break;
			continue;
		}
		if (!have)
			continue;

		bool done = false;   // the step is complete (commit below)
		bool alive = true;   // the particle stays active after this step
		bool moved = true;   // the step sets position and step state (false: deactivated with NaN)
		V3 nPos = pos, nDir = dir;
		double lin[3] = {0, 0, 0};

		if (phase == SDE_BEGIN) {
			// candidate->previous = current; h = clip(nextStep, minStep, maxStep) / c_light
			*cPPx = pos.x; *cPPy = pos.y; *cPPz = pos.z;
			*cPDx = dir.x; *cPDy = dir.y; *cPDz = dir.z;
			h = div_rn(dclip(next, P.minStep, P.maxStep), kCLight, kInvCLight);
			if (q == 0) {
				// rectilinear, with the advection drift (DiffusionSDE.cpp:62-75)
				if (P.sdeAdvection)
					sde_drift(P, pos, h, lin);
				nPos = v3((pos.x + lin[0]) + (dir.x * h) * kCLight, (pos.y + lin[1]) + (dir.y * h) * kCLight,
				          (pos.z + lin[2]) + (dir.z * h) * kCLight);
				next = P.maxStep;
				done = true;
			} else {
				crp_sde_rng r;
				crp_sde_rng_init(&r, P.sdeSeed, stream, rstep);
				const double e0 = sde_randnorm(r), e1 = sde_randnorm(r), e2 = sde_randnorm(r);
				TStep = B0 * e0;
				NStep = B4 * e1;
				BStep = B4 * e2;
				sh = __builtin_sqrt(h);
				propTime = div_rn(TStep * sh, kCLight, kInvCLight);
				counter = 0;
				k0 = sde_unit_c(field.eval(pos));
				nEval += 1;
				phase = SDE_TRIAL;
			}
		}

		if (phase != SDE_BEGIN) {
			// one tryStep: a trial from pos with propTime, or a sub-step from start with allowed
			const bool trial = phase == SDE_TRIAL;
			const V3 x = trial ? pos : start;
			const double t = trial ? propTime : allowed;
			V3 kx = k0;
			if (!trial) {
				kx = sde_unit_c(field.eval(start));
				nEval += 1;
			}
			V3 err;
			sde_trystep(field, x, kx, t, out, err);
			nEval += 5;
			if (trial) {
				const double rr = norm(err) / P.tol;
				lastT = propTime;
				propTime *= 0.5;
				counter += 1;
				const bool more = rr > 1 && __builtin_fabs(propTime) >= minTime;
				if (more && counter > P.sdeMaxHalvings) {
					// another trial would need more than 2^sdeMaxHalvings sub-steps: stop the particle
					// without moving it and report it (CTR_SDE_LIMIT)
					nLimit++;
					alive = false;
					moved = false;
					done = true;
				} else if (!more) {
					{
						stepNumber = (double)(1u << (counter - 1)); // pow(2, counter - 1)
						allowed = div_rn(div_rn(TStep * sh, kCLight, kInvCLight), stepNumber, 1 / stepNumber);
						nRej += counter - 1;
						if (__double_as_longlong(allowed) == __double_as_longlong(lastT)) {
							start = out; // the first sub-step is the last trial
							j = 1;
						} else {
							start = pos;
							j = 0;
						}
						phase = SDE_SUB;
						if (j == stepNumber)
							done = true;
					}
				}
			} else {
				start = out;
				j += 1;
				if (j == stepNumber)
					done = true;
			}
			if (done && moved) {
				// DiffusionSDE.cpp:121-209: the tangent, the no-field case, the normal and binormal
				// displacements, the new direction
				const V3 T = sde_unit(v3(out.x - pos.x, out.y - pos.y, out.z - pos.z));
				const double tTest = norm(T);
				if (P.sdeAdvection)
					sde_drift(P, pos, h, lin);
				if (tTest != tTest) {
					if (P.sdeAdvection)
						nPos = v3(pos.x + lin[0], pos.y + lin[1], pos.z + lin[2]);
					else
						nPos = v3(pos.x + (dir.x * h) * kCLight, pos.y + (dir.y * h) * kCLight,
						          pos.z + (dir.z * h) * kCLight);
					next = dclip(5 * h * kCLight, P.minStep, P.maxStep);
				} else {
					crp_sde_rng r;
					crp_sde_rng_init(&r, P.sdeSeed, stream, rstep);
					r.next = 3; // after the three blocks of the eta draws
					V3 N = v3(0, 0, 0);
					int tries = 0;
					while (norm(N) == 0. && tries < 64) {
						N = sde_cross(T, sde_randvector(r));
						tries++;
					}
					N = sde_unit(N);
					const V3 Bv = sde_unit(sde_cross(T, N));
					const V3 PO = v3((out.x + lin[0]) + (N.x * NStep + Bv.x * BStep) * sh,
					                 (out.y + lin[1]) + (N.y * NStep + Bv.y * BStep) * sh,
					                 (out.z + lin[2]) + (N.z * NStep + Bv.z * BStep) * sh);
					const double por = norm(PO);
					if (por != por) {
						alive = false;
						moved = false;
					} else {
						const V3 D2 = sde_randcone(r, T, cosHalfPi);
						const double dr = norm(D2), idr = 1 / dr;
						nPos = PO;
						nDir = v3(div_rn(D2.x, dr, idr), div_rn(D2.y, dr, idr), div_rn(D2.z, dr, idr));
						// nextStep: h * pow(stepNumber, -2) * c_light; the power of two is exact
						next = stepNumber > 1 ? (h * __builtin_ldexp(1.0, -2 * (counter - 1))) * kCLight
						                      : 4 * h * kCLight;
					}
				}
			}
		}
		if (!done)
			continue;

		// ---- commit: position, direction, setCurrentStep(h c), then MaximumTrajectoryLength
		phase = SDE_BEGIN;
		double len = *cLen;
		if (moved) {
			pos = nPos;
			dir = nDir;
			const double hc = h * kCLight;
			*cCur = hc;
			len = len + hc;
			*cLen = len;
			*cTim = *cTim + div_rn(hc, kCLight, kInvCLight);
		}
		rstep++;
		nAcc++;
		steps++;
		// the modules after DiffusionSDE in the list, as k_propagate applies them (they run even when the
		// step deactivated the particle, as ModuleList::process does)
		if (MODS && P.box) {
			// PeriodicBox::process (src/module/Boundary.cpp:16-27)
			const double nbx = __builtin_floor(div_rn(pos.x - P.boxO[0], P.boxS[0], P.boxInv[0]));
			const double nby = __builtin_floor(div_rn(pos.y - P.boxO[1], P.boxS[1], P.boxInv[1]));
			const double nbz = __builtin_floor(div_rn(pos.z - P.boxO[2], P.boxS[2], P.boxInv[2]));
			if (!((nbx == 0) && (nby == 0) && (nbz == 0))) {
				const double shx = nbx * P.boxS[0], shy = nby * P.boxS[1], shz = nbz * P.boxS[2];
				pos = v3(pos.x - shx, pos.y - shy, pos.z - shz);
				*cPPx = *cPPx - shx; *cPPy = *cPPy - shy; *cPPz = *cPPz - shz;
				S.sx[idx] = S.sx[idx] - shx; S.sy[idx] = S.sy[idx] - shy; S.sz[idx] = S.sz[idx] - shz;
				S.cx[idx] = S.cx[idx] - shx; S.cy[idx] = S.cy[idx] - shy; S.cz[idx] = S.cz[idx] - shz;
			}
		}
		double nextObs = next; // nextStep as the observer sees it
		if (MODS && P.sph) {
			// SphericalBoundary::process (src/module/Boundary.cpp:216-223)
			const double d = norm(v3(pos.x - P.sphC[0], pos.y - P.sphC[1], pos.z - P.sphC[2]));
			if (d >= P.sphR)
				alive = false;
			if (P.sphLimit) {
				const double l = (P.sphR - d) + P.sphMargin;
				next = dmin(next, l);
				if (!(P.postObs & 1))
					nextObs = dmin(nextObs, l);
			}
		}
		if (P.maxLen > 0) { // MaximumTrajectoryLength::process
			if (len >= P.maxLen) {
				alive = false;
			} else {
				const double l = P.maxLen - len;
				next = dmin(next, l);
				if (!(P.postObs & 2))
					nextObs = dmin(nextObs, l);
			}
		}
		if (MODS && P.nObs > 0) {
			// Observer with ObserverTimeEvolution (src/module/Observer.cpp:26-57, 279-306)
			const int k = (int)*cDet;
			if (k < P.nObs) {
				const double distance = len - P.obsT[k];
				if (distance < 0.) {
					next = dmin(next, -distance);
				} else {
					if (k < P.nObs - 2) {
						const double l = P.obsT[k + 1] - len;
						next = dmin(next, l);
						nextObs = dmin(nextObs, l);
					}
					*cDet = (double)(k + 1);
					double *r = S.snap + (size_t)k * kSnapDoubles * S.cap + idx;
					r[0] = pos.x; r[S.cap] = pos.y; r[2 * S.cap] = pos.z;
					r[3 * S.cap] = dir.x; r[4 * S.cap] = dir.y; r[5 * S.cap] = dir.z;
					r[6 * S.cap] = len; r[7 * S.cap] = *cTim; r[8 * S.cap] = nextObs;
					if (P.obsDeactivate)
						alive = false;
				}
			}
		}
		nFin += alive ? 0 : 1;
		if (!alive || (P.maxSteps >= 0 && steps >= P.maxSteps)) {
			S.px[idx] = pos.x; S.py[idx] = pos.y; S.pz[idx] = pos.z;
			S.dx[idx] = dir.x; S.dy[idx] = dir.y; S.dz[idx] = dir.z;
			S.cur[idx] = *cCur; S.next[idx] = next; S.len[idx] = len; S.time[idx] = *cTim;
			S.active[idx] = alive ? 1 : 0;
			S.ppx[idx] = *cPPx; S.ppy[idx] = *cPPy; S.ppz[idx] = *cPPz;
			S.pdx[idx] = *cPDx; S.pdy[idx] = *cPDy; S.pdz[idx] = *cPDz;
			S.sdeStep[idx] = rstep;
			if (MODS && P.nObs > 0)
				S.det[idx] = (int32_t)*cDet;
			nPause += alive ? 1 : 0;
			have = false;
		}
	}
	const unsigned long long tAcc = wave_sum(nAcc), tRej = wave_sum(nRej), tEval = wave_sum(nEval),
	                         tFin = wave_sum(nFin), tPause = wave_sum(nPause), tLim = wave_sum(nLimit);
	if (lane == 0) {
		atomicAdd(&ctr[CTR_ACCEPTED], tAcc);
		atomicAdd(&ctr[CTR_REJECTED], tRej);
		atomicAdd(&ctr[CTR_EVALS], tEval);
		atomicAdd(&ctr[CTR_FINISHED], tFin);
		atomicAdd(&ctr[CTR_PAUSED], tPause);
		if (tLim)
			atomicAdd(&ctr[CTR_SDE_LIMIT], tLim);
	}
}

} // namespace crp
