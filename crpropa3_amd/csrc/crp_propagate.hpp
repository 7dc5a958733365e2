// The persistent propagation kernel (included by crp_kernels.hip).
//
// One particle per lane.  Each loop iteration performs one trial of
// PropagationBP/CK::process() for every busy lane, so the 1 / 3 / 6 field
// evaluations of a trial stay converged across the wave; an accepted trial
// commits the step (setPosition, setDirection, setCurrentStep, setNextStep),
// the fused MaximumTrajectoryLength condition and the step counter.  Lanes
// whose particle finished (inactive, or its step budget spent) store it and
// take the next queue entry: one atomicAdd per wave reserves popcount(ballot)
// entries, each lane takes its prefix-count slot (wave64 ballot/mbcnt).
//
// MIX = false: the reference's FP64 arithmetic bit for bit
//   (src/module/PropagationBP.cpp:62-124, src/module/PropagationCK.cpp:77-136,
//    src/module/BreakCondition.cpp:42-64, src/Candidate.cpp:92-104).
// MIX = true: FP32 field evaluation, rotation and direction stages; FP64
//   positions, CK error norm, step acceptance and step-size update.
#pragma once

#include "crp_composite.hpp"
#include "crp_device.hpp"

namespace crp {

// PK_SDE: DiffusionSDE (k_sde, crp_sde.hpp)
enum PropKind { PK_BP_FIXED = 0, PK_BP_ADAPTIVE = 1, PK_CK_FIXED = 2, PK_CK_ADAPTIVE = 3, PK_SDE = 4 };
// CTR_PAUSED: particles stored still active because this launch's step budget ran out
// CTR_SDE_LIMIT: DiffusionSDE particles stopped because a step needed too many sub-steps
enum { CTR_QUEUE = 0, CTR_ACCEPTED, CTR_REJECTED, CTR_EVALS, CTR_FINISHED, CTR_PAUSED, CTR_SDE_LIMIT, CTR_COUNT };
constexpr int kColdDoubles = 10; // dynamic LDS per thread of k_propagate, in doubles

struct Soa {
	double *px, *py, *pz, *dx, *dy, *dz, *E, *q, *cur, *next, *len, *time;
	double *ppx, *ppy, *ppz, *pdx, *pdy, *pdz;
	double *sx, *sy, *sz, *cx, *cy, *cz; // source and created positions (moved by PeriodicBox)
	int32_t *active;
	const uint32_t *order;
	int32_t *det;             // observer detections so far per particle ("DetectionIndex")
	double *snap;             // state at detection k of particle i: snap[(k * kSnapDoubles + c) * cap + i]
	unsigned long long cap;   // snapshot stride
	uint32_t *sdeStep;        // DiffusionSDE: steps taken (the random-number counter)
	const uint64_t *sdeStream; // DiffusionSDE: the particle's random stream
	const double *z;          // the candidate's redshift (read by fields that use it)
};

constexpr int kSnapDoubles = 9; // position x,y,z, direction x,y,z, trajectory length, time, nextStep

// One advection field of DiffusionSDE (src/advectionField/AdvectionField.cpp); `kind` is a
// crp_sde_advection_kind, `par` its parameters in the order the C ABI documents.
constexpr int kSdeAdvMax = 8;
struct SdeAdvTerm {
	int kind, pad;
	double par[9];
};

struct Params {
	double tol, minStep, maxStep, maxLen;
	long long maxSteps; // < 0: unlimited
	// Observer with one ObserverTimeEvolution after the length limit (nObs = 0: none)
	const double *obsT;
	int nObs, obsDeactivate;
	// PeriodicBox right after the propagator (box = 0: none); boxInv = RN(1/size)
	int box;
	double boxO[3], boxS[3], boxInv[3];
	// SphericalBoundary after the box (sph = 0: none): reject at |pos - center| >= radius;
	// sphLimit: limitNextStep(radius - d + margin)
	int sph, sphLimit;
	double sphC[3], sphR, sphMargin;
	// conditions that run after the observer in the module list (bit 0: spherical boundary, bit 1: length
	// limit); the observer sees, and records, nextStep limited by the modules before it only
	int postObs;
	// DiffusionSDE (k_sde): diffusion coefficient scale * 6.1e24 * (|E/q| / 4e9)^alpha, perpendicular
	// fraction epsilon, Philox seed; the advection field as sdeAdvection terms (device memory) whose
	// velocities are summed in order (0: no advection field)
	double sdeScale, sdeAlpha, sdeEpsilon;
	const SdeAdvTerm *sdeAdvTerms;
	int sdeAdvection;
	uint32_t sdeSeed;
	int sdeMaxHalvings; // trials beyond this many halvings stop the particle (CTR_SDE_LIMIT)
};

__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; } // std::max
__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; } // std::min
__device__ __forceinline__ double dclip(double x, double lo, double hi) { return dmax(lo, dmin(x, hi)); }

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
	for (int o = 32; o > 0; o >>= 1)
		v += __shfl_xor(v, o);
	return v;
}

// r^-0.2 of the step-size update (used as (step * 0.95) * r^-0.2, the
// reference's order): correctly rounded on the FP64 path (the reference calls
// glibc pow), the device libm on the mixed path
template <bool MIX>
__device__ __forceinline__ double pow_m02(double r) {
	if (MIX)
		return ::pow(r, -0.2);
	return crp_pow_cr(r, -0.2);
}

// Register budget: the FP64 path through the cell table is asked for 3 waves per SIMD (<= 168
// VGPRs; it needs 169 unconstrained and then runs at 2), everything else is left to the compiler.
template <class F>
struct WavesHint {
	static constexpr int fp64 = 1;
};
template <>
struct WavesHint<FieldGridCells> {
	static constexpr int fp64 = 3;
};

template <class F, int KIND, bool MIX>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MIX ? 1 : WavesHint<F>::fp64, 8)))
k_propagate(F field0, Params P, Soa S, unsigned long long n,
                                                   unsigned long long *__restrict__ ctr) {
	constexpr bool kCK = (KIND == PK_CK_FIXED || KIND == PK_CK_ADAPTIVE);
	constexpr bool kAdaptive = (KIND == PK_BP_ADAPTIVE || KIND == PK_CK_ADAPTIVE);
	constexpr unsigned long long kEvalsPerTrial = kCK ? 6 : (kAdaptive ? 3 : 1);
	const int lane = __lane_id();
	F field = field0; // the lane's copy: a field that reads the candidate's redshift binds it at refill

	bool have = false, exhausted = false, inProc = false;
	CkB0Store<FloatValuedField<F>::value> b0; // the field at the trial sequence's start (adaptive CK), for retries
	CkB0f b0f;
	b0.have = false;
	b0f.have = false;
	unsigned idx = 0;
	long long steps = 0;
	V3 pos = v3(0, 0, 0), dir = v3(0, 0, 0);
	float3 dirf = make_float3(0.f, 0.f, 0.f);
	double E = 0, q = 0, next = 0;
	double step = 0, newStep = 0;
	// per-lane counters, flushed to ctr[] before they could wrap
	unsigned nAcc = 0, nRej = 0, nEval = 0, nFin = 0, nPause = 0;
	// cold per-particle state (previous position and direction, currentStep, trajectoryLength,
	// time: read at refill, commit or store only) lives in LDS, kColdDoubles * blockDim.x doubles
	// passed as dynamic shared memory, which keeps it out of the register file
	// Layout: per wave, kColdDoubles rows of 64 lanes (constant row stride for every block size).
	extern __shared__ double crp_cold[];
	double *const cw = crp_cold + (threadIdx.x >> 6) * (kColdDoubles * 64) + (threadIdx.x & 63);
	double *const cPPx = cw, *const cPPy = cw + 64, *const cPPz = cw + 128;
	double *const cPDx = cw + 192, *const cPDy = cw + 256, *const cPDz = cw + 320;
	double *const cCur = cw + 384, *const cLen = cw + 448, *const cTim = cw + 512;
	double *const cDet = cw + 576; // observer detection count (small integer held as a double)

	for (;;) {
		if (nEval > 0x80000000u) { // flush this lane's counters long before they wrap
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
						if (MIX)
							dirf = make_float3((float)dir.x, (float)dir.y, (float)dir.z);
						E = S.E[idx];
						q = S.q[idx];
						if (UsesRedshift<F>::value)
							bind_redshift(field, S.z[idx]);
						*cCur = S.cur[idx];
						next = S.next[idx];
						*cLen = S.len[idx];
						*cTim = S.time[idx];
						if (P.nObs > 0)
							*cDet = (double)S.det[idx];
						steps = 0;
						inProc = false;
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

		// ---- one trial of process() for this lane
		bool commit = false;
		bool neutral = false;
		Y yOut;          // FP64 path result
		V3 oPos;         // mixed path result
		float3 oDir;
		if (!inProc) {
			*cPPx = pos.x; *cPPy = pos.y; *cPPz = pos.z; // candidate->previous = current
			if (MIX) {
				*cPDx = dirf.x; *cPDy = dirf.y; *cPDz = dirf.z;
			} else {
				*cPDx = dir.x; *cPDy = dir.y; *cPDz = dir.z;
			}
			if (q == 0) {
				// rectilinear propagation for neutral particles (PropagationBP.cpp:74-80)
				neutral = true;
				step = dclip(next, P.minStep, P.maxStep);
				if (MIX)
					oPos = v3(pos.x + (double)dirf.x * step, pos.y + (double)dirf.y * step, pos.z + (double)dirf.z * step);
				else
					yOut.x = v3(pos.x + dir.x * step, pos.y + dir.y * step, pos.z + dir.z * step);
				newStep = P.maxStep;
				commit = true;
			} else if (!kAdaptive) {
				step = P.maxStep;
				newStep = step;
				if (MIX) {
					if (kCK) {
						double eU;
						YM o;
						ck_trystep_mixed(field, YM{pos, dirf}, step / kCLight, (float)((q * kCLight * kCLight) / E), o, eU);
						oPos = o.x;
						oDir = o.u;
					} else {
						oPos = pos;
						oDir = dirf;
						bp_dy_mixed(field, oPos, oDir, step, (float)((q * kCLight * kCLight) / (2 * E) * step / kCLight));
					}
				} else {
					if (kCK) {
						double eU;
						ck_trystep(field, Y{pos, dir}, step / kCLight, (q * kCLight) / E, yOut, eU);
					} else {
						const double m_ = E / (kCLight * kCLight);
						yOut = bp_dy(field, pos, dir, step, q, m_, 1 / m_);
					}
				}
				nEval += kEvalsPerTrial;
				commit = true;
			} else {
				step = dclip(next, P.minStep, P.maxStep);
				newStep = step;
				inProc = true;
				b0.have = false;
				b0f.have = false;
			}
		}
		if (kAdaptive && inProc) {
			double r;
			if (MIX) {
				if (kCK) {
					double eU;
					YM o;
					ck_trystep_mixed(field, YM{pos, dirf}, step / kCLight, (float)((q * kCLight * kCLight) / E), o, eU,
					                 &b0f);
					oPos = o.x;
					oDir = o.u;
					r = eU / P.tol;
				} else {
					const double k0 = (q * kCLight * kCLight) / (2 * E) / kCLight;
					oPos = pos;
					oDir = dirf;
					bp_dy_mixed(field, oPos, oDir, step, (float)(k0 * step));
					V3 hPos = pos;
					float3 hDir = dirf;
					bp_dy_mixed(field, hPos, hDir, step / 2, (float)(k0 * (step / 2)));
					bp_dy_mixed(field, hPos, hDir, step / 2, (float)(k0 * (step / 2)));
					V3 d = v3(oPos.x - hPos.x, oPos.y - hPos.y, oPos.z - hPos.z);
					double Sv = norm(d) / (step * 0.75);
					r = __builtin_sqrt(3.0 * Sv * Sv) / P.tol;
				}
			} else {
				if (kCK) {
					double eU;
					ck_trystep(field, Y{pos, dir}, step / kCLight, (q * kCLight) / E, yOut, eU, &b0);
					r = eU / P.tol;
				} else {
					const double m_ = E / (kCLight * kCLight);
					const double im_ = 1 / m_;
					yOut = bp_dy(field, pos, dir, step, q, m_, im_);
					Y half = bp_dy(field, pos, dir, step / 2, q, m_, im_);
					Y cmp = bp_dy(field, half.x, half.u, step / 2, q, m_, im_);
					V3 d = v3(yOut.x.x - cmp.x.x, yOut.x.y - cmp.x.y, yOut.x.z - cmp.x.z);
					double Sv = norm(d) / (step * (1 - 1 / 4.));
					r = __builtin_sqrt((Sv * Sv + Sv * Sv) + Sv * Sv) / P.tol;
				}
			}
			nEval += kEvalsPerTrial; // the reference's getField calls, the reused first stage included
			if (r > 1) {
				if (step == P.minStep) {
					commit = true;
				} else {
					nRej++;
					newStep = step * 0.95 * pow_m02<MIX>(r);
					newStep = dmax(newStep, 0.1 * step);
					newStep = dmax(newStep, P.minStep);
					step = newStep;
				}
			} else {
				if (step != P.maxStep) {
					newStep = step * 0.95 * pow_m02<MIX>(r);
					newStep = dmin(newStep, 5 * step);
					newStep = dmin(newStep, P.maxStep);
				}
				commit = true;
			}
		}
		if (!commit)
			continue;

		// ---- commit: setPosition, setDirection(getUnitVector()), setCurrentStep, setNextStep
		inProc = false;
		if (MIX) {
			pos = oPos;
			if (!neutral)
				dirf = unit3(oDir);
		} else {
			pos = yOut.x;
			if (!neutral) {
				const double n1 = norm(yOut.u), i1 = 1 / n1;
				V3 u1 = v3(div_rn(yOut.u.x, n1, i1), div_rn(yOut.u.y, n1, i1), div_rn(yOut.u.z, n1, i1));
				const double n2 = norm(u1), i2 = 1 / n2;
				dir = v3(div_rn(u1.x, n2, i2), div_rn(u1.y, n2, i2), div_rn(u1.z, n2, i2));
			}
		}
		if (P.box) {
			// PeriodicBox::process (src/module/Boundary.cpp:16-27): n = floor((pos - origin) / size);
			// current, previous, source and created positions move by -n * size
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
		*cCur = step;
		const double len = *cLen + step;
		*cLen = len;
		*cTim = *cTim + step / kCLight;
		next = newStep;
		nAcc++;
		steps++;
		bool alive = true;
		double nextObs = next; // nextStep as the observer sees it
		if (P.sph) {
			// SphericalBoundary::process (src/module/Boundary.cpp:216-223), d = Vector3::getR of pos - center
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
		if (P.nObs > 0) {
			// Observer::process with ObserverTimeEvolution::checkDetection (src/module/Observer.cpp:26-57,
			// 279-306): at most one detection per step; the next step is limited towards the next time
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
					const V3 d = MIX ? v3(dirf.x, dirf.y, dirf.z) : dir;
					r[0] = pos.x; r[S.cap] = pos.y; r[2 * S.cap] = pos.z;
					r[3 * S.cap] = d.x; r[4 * S.cap] = d.y; r[5 * S.cap] = d.z;
					r[6 * S.cap] = len; r[7 * S.cap] = *cTim; r[8 * S.cap] = nextObs;
					if (P.obsDeactivate)
						alive = false;
				}
			}
		}
		nFin += alive ? 0 : 1;
		if (!alive || (P.maxSteps >= 0 && steps >= P.maxSteps)) {
			if (MIX)
				dir = v3(dirf.x, dirf.y, dirf.z);
			S.px[idx] = pos.x; S.py[idx] = pos.y; S.pz[idx] = pos.z;
			S.dx[idx] = dir.x; S.dy[idx] = dir.y; S.dz[idx] = dir.z;
			S.cur[idx] = *cCur; S.next[idx] = next; S.len[idx] = len; S.time[idx] = *cTim;
			S.active[idx] = alive ? 1 : 0;
			S.ppx[idx] = *cPPx; S.ppy[idx] = *cPPy; S.ppz[idx] = *cPPz;
			S.pdx[idx] = *cPDx; S.pdy[idx] = *cPDy; S.pdz[idx] = *cPDz;
			if (P.nObs > 0)
				S.det[idx] = (int32_t)*cDet;
			nPause += alive ? 1 : 0;
			have = false;
		}
	}
	const unsigned long long tAcc = wave_sum(nAcc), tRej = wave_sum(nRej), tEval = wave_sum(nEval),
	                         tFin = wave_sum(nFin), tPause = wave_sum(nPause);
	if (lane == 0) {
		atomicAdd(&ctr[CTR_ACCEPTED], tAcc);
		atomicAdd(&ctr[CTR_REJECTED], tRej);
		atomicAdd(&ctr[CTR_EVALS], tEval);
		atomicAdd(&ctr[CTR_FINISHED], tFin);
		atomicAdd(&ctr[CTR_PAUSED], tPause);
	}
}

} // namespace crp
