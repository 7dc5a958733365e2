// HIP kernels of the batched propagator and the C ABI declared in
// include/crpropa3_amd.h.
//
// Execution model (one particle per lane, persistent waves):
//   * particle state lives in registers for the whole run; HBM holds it as a
//     structure of arrays (one array per component, coalesced per wave);
//   * each loop iteration performs one trial of PropagationBP/CK::process()
//     for every busy lane (the 1 / 3 / 6 field evaluations stay converged
//     across the wave); accepted trials commit the step, the fused
//     MaximumTrajectoryLength condition, and the step counter;
//   * lanes whose particle finished (inactive, or its step budget spent)
//     store it and are refilled from a device queue: one atomicAdd per wave
//     reserves popcount(ballot) indices and each lane takes its mbcnt slot.
// Reference semantics: src/module/PropagationBP.cpp:62-124,
// src/module/PropagationCK.cpp:77-136, src/module/BreakCondition.cpp:42-64,
// src/Candidate.cpp:92-104, src/ParticleState.cpp:30-32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "crp_device.hpp"
#include "crp_propagate.hpp"
#include "crp_composite.hpp"
#include "crp_sde.hpp"
#include "crp_jf12.hpp"
#include "crp_adjoint.hpp"
#include "crp_turbulence.hpp"
#include "crp_probe.hpp"
#include "crpropa3_amd.h"

#include <dlfcn.h>
#include <hipfft/hipfft.h>

using namespace crp;

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;

static int fail(int code, const std::string &msg) {
	g_err = msg;
	return code;
}

#define CRP_HIP_TRY(call)                                                                                      \
	do {                                                                                                       \
		hipError_t e_ = (call);                                                                                \
		if (e_ != hipSuccess)                                                                                  \
			return fail(CRP_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));                        \
	} while (0)

// ------------------------------------------------------------------ objects
enum FieldKind { FK_NONE = 0, FK_UNIFORM = 1, FK_GRID = 2, FK_PLANEWAVE = 3, FK_MODULATED = 4, FK_JF12 = 5, FK_COMPOSITE = 6 };
// FK_COMPOSITE decorators (crp_field.deco)
enum DecoKind { DECO_PERIODIC = 1, DECO_LIST = 2, DECO_EVOLUTION = 3 };
constexpr int kDecoMaxChildren = 8;

struct crp_field {
	int kind;
	int device;
	FieldUniform uniform;
	GridDesc grid;
	GridDesc mod; // FK_MODULATED: the Grid1f modulation, one float per point
	FieldJF12 jf12; // FK_JF12: parameters and grids
	PlaneWaveDesc pw;
	float *cells; // cell-major corner table (periodic grids), or nullptr
	unsigned char *cells16; // FP16 corner table with brick scales (mixed path), or nullptr
	void *owned[8];
	// FK_COMPOSITE: a decorator over borrowed fields (the caller keeps them alive) and the device
	// program compiled from the tree (rebuilt before each use: a child's layout may have changed)
	int deco;
	const crp_field *child[kDecoMaxChildren];
	int nchild;
	double extends[3], porigin[3];
	int preflective;
	double evoM;
	CompositeDesc *dcomp; // device copy of the program
	CompositeDesc *hcomp; // host copy of what dcomp holds
};

struct crp_propagator {
	int type; // 0 BP, 1 CK, 2 DiffusionSDE
	const crp_field *field;
	double tolerance, minStep, maxStep;
	double maxLength;
	int precision;
	// launch configuration (crp_launch_config) and the geometry of the last call
	int block, blocksPerCU;
	long long stepsPerLaunch;
	int useGraph;
	mutable unsigned lastGrid, lastBlock;
	mutable double lastPPL;
	// fused Observer(ObserverTimeEvolution): detection lengths on the device
	double *obsT;
	int nObs, obsDeactivate;
	// fused PeriodicBox
	int box;
	double boxO[3], boxS[3];
	// fused SphericalBoundary
	int sph, sphLimit;
	double sphC[3], sphR, sphMargin;
	// Params::postObs of the module order
	int postObs;
	// DiffusionSDE (type 2)
	double sdeScale, sdeAlpha, sdeEpsilon;
	SdeAdvTerm sdeAdv[kSdeAdvMax];  // the advection field's terms (crp_propagator_add_sde_advection)
	int sdeAdvection;               // their number
	mutable SdeAdvTerm *dSdeAdv;    // device copy, made or refreshed by crp_propagate
	mutable int sdeAdvDirty;        // the terms changed since the last copy
	uint32_t sdeSeed;
};


struct crp_batch {
	size_t cap, n;
	int device;
	hipStream_t stream;
	double *base;
	double *px, *py, *pz, *dx, *dy, *dz, *E, *q, *cur, *next, *len, *time;
	double *ppx, *ppy, *ppz, *pdx, *pdy, *pdz;
	double *sx, *sy, *sz, *cx, *cy, *cz; // source and created positions
	double *z;                           // redshift (read by a top-level MagneticFieldEvolution)
	int32_t *active;
	uint32_t *order; // queue order: particle indices, longest expected work first
	unsigned long long *ctr;
	hipEvent_t ev0, ev1;
	int32_t *det;     // observer detections per particle (inside the main allocation)
	uint32_t *sdeStep; // DiffusionSDE steps taken per particle: the Philox counter (main allocation)
	uint64_t *sdeStream; // DiffusionSDE random stream per particle (main allocation, 8-byte aligned)
	double *snap;     // detection snapshots [snapK][kSnapDoubles][cap], allocated on first use
	int snapK;
	// captured launch sequence of the last graph-enabled call, keyed by the exact bytes of every
	// kernel argument and launch dimension it holds
	hipGraphExec_t gexec;
	size_t gkeyLen;
	unsigned char gkey[1024];
};


static Soa soa_of(const crp_batch *b) {
	Soa s;
	s.px = b->px; s.py = b->py; s.pz = b->pz;
	s.dx = b->dx; s.dy = b->dy; s.dz = b->dz;
	s.E = b->E; s.q = b->q; s.cur = b->cur; s.next = b->next; s.len = b->len; s.time = b->time;
	s.ppx = b->ppx; s.ppy = b->ppy; s.ppz = b->ppz;
	s.pdx = b->pdx; s.pdy = b->pdy; s.pdz = b->pdz;
	s.sx = b->sx; s.sy = b->sy; s.sz = b->sz;
	s.cx = b->cx; s.cy = b->cy; s.cz = b->cz;
	s.active = b->active;
	s.order = b->order;
	s.det = b->det;
	s.snap = b->snap;
	s.cap = b->cap;
	s.sdeStep = b->sdeStep;
	s.sdeStream = b->sdeStream;
	s.z = b->z;
	return s;
}

// ------------------------------------------------------------------ kernels
__global__ void k_zero_u64(unsigned long long *p, int n) {
	int i = threadIdx.x;
	if (i < n)
		p[i] = 0ull;
}

// Plain streaming copy (batch reset); also the known-size reference kernel
// against which HBM traffic counters are calibrated.
__global__ void __launch_bounds__(256) k_copy_f64(const double *__restrict__ src, double *__restrict__ dst, size_t n) {
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	const size_t stride = (size_t)gridDim.x * blockDim.x;
	for (; i < n; i += stride)
		dst[i] = src[i];
}

// Cell-major corner table of a periodic grid: cell (x0,y0,z0) stores the values
// of its 8 corners, in the order of Grid.h:572-579, as 24 floats + 8 zeros.
__global__ void __launch_bounds__(256) k_build_cells(const float *__restrict__ g, float *__restrict__ cells, int nx,
                                                     int ny, int nz) {
	const size_t total = (size_t)nx * ny * nz;
	const size_t stride = (size_t)gridDim.x * blockDim.x;
	for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < total; c += stride) {
		const int z0 = (int)(c % nz);
		const size_t t = c / nz;
		const int y0 = (int)(t % ny);
		const int x0 = (int)(t / ny);
		const int x1 = (x0 + 1) % nx, y1 = (y0 + 1) % ny, z1 = (z0 + 1) % nz;
		const int X[8] = {x0, x1, x0, x0, x1, x0, x1, x1};
		const int Yc[8] = {y0, y0, y1, y0, y0, y1, y1, y1};
		const int Z[8] = {z0, z0, z0, z1, z1, z1, z0, z1};
		float v[32];
#pragma unroll
		for (int k = 0; k < 8; k++) {
			const float *s = g + 3 * (((size_t)X[k] * ny + Yc[k]) * nz + Z[k]);
			v[3 * k] = s[0];
			v[3 * k + 1] = s[1];
			v[3 * k + 2] = s[2];
		}
#pragma unroll
		for (int k = 24; k < 32; k++)
			v[k] = 0.0f;
		float4 *o = reinterpret_cast<float4 *>(cells + (size_t)kCellFloats * c);
#pragma unroll
		for (int k = 0; k < 8; k++)
			o[k] = make_float4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
	}
}

// Per 8^3-point brick: the largest |component| over the brick's points and its
// +1 faces (periodic), i.e. over every corner of every cell in the brick.
__global__ void __launch_bounds__(256) k_brick_scales(const float *__restrict__ g, float *__restrict__ scales, int nx,
                                                      int ny, int nz, int bx, int by, int bz) {
	const size_t nb = (size_t)bx * by * bz;
	const size_t b = (size_t)blockIdx.x;
	if (b >= nb)
		return;
	const int kz = (int)(b % bz), ky = (int)((b / bz) % by), kx = (int)(b / ((size_t)bz * by));
	float m = 0.f;
	for (int t = threadIdx.x; t < 9 * 9 * 9; t += blockDim.x) {
		const int dz = t % 9, dy = (t / 9) % 9, dx = t / 81;
		const int x = (kx * 8 + dx) % nx, y = (ky * 8 + dy) % ny, z = (kz * 8 + dz) % nz;
		const float *s = g + 3 * (((size_t)x * ny + y) * nz + z);
		m = fmaxf(m, fmaxf(fabsf(s[0]), fmaxf(fabsf(s[1]), fabsf(s[2]))));
	}
	__shared__ float red[256];
	red[threadIdx.x] = m;
	__syncthreads();
	for (int o = blockDim.x / 2; o > 0; o >>= 1) {
		if (threadIdx.x < o)
			red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
		__syncthreads();
	}
	if (threadIdx.x == 0)
		scales[b] = red[0] > 0.f ? red[0] : 1.f;
}

// FP16 cell table: corners of cell (x0,y0,z0) divided by its brick's scale,
// rounded to FP16, then the scale (see FieldGridCells16).
__global__ void __launch_bounds__(256) k_build_cells16(const float *__restrict__ g, const float *__restrict__ scales,
                                                       unsigned char *__restrict__ cells, int nx, int ny, int nz,
                                                       int by, int bz) {
	const size_t total = (size_t)nx * ny * nz;
	const size_t stride = (size_t)gridDim.x * blockDim.x;
	for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < total; c += stride) {
		const int z0 = (int)(c % nz);
		const size_t t = c / nz;
		const int y0 = (int)(t % ny);
		const int x0 = (int)(t / ny);
		const int x1 = (x0 + 1) % nx, y1 = (y0 + 1) % ny, z1 = (z0 + 1) % nz;
		const int X[8] = {x0, x1, x0, x0, x1, x0, x1, x1};
		const int Yc[8] = {y0, y0, y1, y0, y0, y1, y1, y1};
		const int Z[8] = {z0, z0, z0, z1, z1, z1, z0, z1};
		const float scale = scales[((size_t)(x0 >> 3) * by + (y0 >> 3)) * bz + (z0 >> 3)];
		const float inv = 1.f / scale;
		unsigned short h[24];
#pragma unroll
		for (int k = 0; k < 8; k++) {
			const float *s = g + 3 * (((size_t)X[k] * ny + Yc[k]) * nz + Z[k]);
#pragma unroll
			for (int a = 0; a < 3; a++)
				h[3 * k + a] = __half_as_ushort(__float2half_rn(s[a] * inv));
		}
		unsigned int wv[16];
#pragma unroll
		for (int k = 0; k < 12; k++)
			wv[k] = (unsigned int)h[2 * k] | ((unsigned int)h[2 * k + 1] << 16);
		wv[12] = __float_as_uint(scale);
		wv[13] = wv[14] = wv[15] = 0u;
		uint4 *o = reinterpret_cast<uint4 *>(cells + (size_t)kCell16Bytes * c);
#pragma unroll
		for (int k = 0; k < 4; k++)
			o[k] = make_uint4(wv[4 * k], wv[4 * k + 1], wv[4 * k + 2], wv[4 * k + 3]);
	}
}

template <class F>
__global__ void k_field_evalf(F field, const double *__restrict__ pos, double *__restrict__ B, size_t n) {
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	V3 p = v3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
	float3 b = field.evalf(p);
	B[3 * i] = b.x;
	B[3 * i + 1] = b.y;
	B[3 * i + 2] = b.z;
}

template <class F>
__global__ void k_field_eval(F field, const double *__restrict__ pos, double *__restrict__ B, size_t n) {
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	V3 p = v3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
	V3 b = field.eval(p);
	B[3 * i] = b.x;
	B[3 * i + 1] = b.y;
	B[3 * i + 2] = b.z;
}

// The reference's public per-step functions on n phase points (PropagationBP::dY / tryStep,
// PropagationCK::dYdt / tryStep), with the same device code as the propagation kernels.
enum StepOp { SO_BP_DY = 0, SO_BP_TRY = 1, SO_CK_DYDT = 2, SO_CK_TRY = 3 };
struct StepArgs {
	const double *x, *u, *h, *z, *q, *m; // x, u AoS [3n]; m: BP mass, CK energy
	double *ox, *ou, *ex, *eu;            // AoS [3n]
	size_t n;
};

template <class F, int OP>
__global__ void k_step_api(F field0, StepArgs A) {
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= A.n)
		return;
	F field = field0;
	bind_redshift(field, A.z[i]);
	const V3 x = v3(A.x[3 * i], A.x[3 * i + 1], A.x[3 * i + 2]), u = v3(A.u[3 * i], A.u[3 * i + 1], A.u[3 * i + 2]);
	const double q = A.q[i], m = A.m[i];
	Y o, e;
	if (OP == SO_BP_DY || OP == SO_BP_TRY) {
		const double h = A.h[i], invm = 1 / m;
		o = bp_dy(field, x, u, h, q, m, invm);
		if (OP == SO_BP_TRY) { // error = Y(errorEstimation(out.x, dY(dY(h/2), h/2).x, h))
			const Y half = bp_dy(field, x, u, h / 2, q, m, invm);
			const Y cmp = bp_dy(field, half.x, half.u, h / 2, q, m, invm);
			const double S = norm(v3(o.x.x - cmp.x.x, o.x.y - cmp.x.y, o.x.z - cmp.x.z)) / (h * (1 - 1 / 4.));
			e.x = v3(S, S, S);
			e.u = e.x;
		}
	} else if (OP == SO_CK_DYDT) {
		o = ck_dydt(field, Y{x, u}, (q * kCLight) / m);
	} else {
		ck_trystep_err(field, Y{x, u}, A.h[i], (q * kCLight) / m, o, e);
	}
	const V3 w[4] = {o.x, o.u, e.x, e.u};
	double *dst[4] = {A.ox, A.ou, A.ex, A.eu};
	const int nout = (OP == SO_BP_TRY || OP == SO_CK_TRY) ? 4 : 2;
	for (int k = 0; k < nout; k++) {
		dst[k][3 * i] = w[k].x;
		dst[k][3 * i + 1] = w[k].y;
		dst[k][3 * i + 2] = w[k].z;
	}
}

// getField(position, z) per position: the field bound to each position's redshift
template <class F>
__global__ void k_field_eval_z(F field0, const double *__restrict__ pos, const double *__restrict__ z,
                               double *__restrict__ B, size_t n) {
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	F field = field0;
	bind_redshift(field, z[i]);
	V3 b = field.eval(v3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]));
	B[3 * i] = b.x;
	B[3 * i + 1] = b.y;
	B[3 * i + 2] = b.z;
}



// ------------------------------------------------------------------ launch helpers
static int num_cus() {
	static int numCU = -1;
	if (numCU < 0) {
		int dev, v = 0;
		CRP_HIP_TRY(hipGetDevice(&dev));
		CRP_HIP_TRY(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev));
		numCU = v;
	}
	return numCU;
}

// Append bytes to a graph key; false when the key buffer is full (then no graph is used).
struct KeyBuf {
	unsigned char *p;
	size_t len, cap;
	bool ok;
	template <class T> void put(const T &v) {
		if (!ok || len + sizeof(T) > cap) {
			ok = false;
			return;
		}
		std::memcpy(p + len, &v, sizeof(T));
		len += sizeof(T);
	}
};

// One crp_propagate call: [zero counters] then launches of k_propagate.  With
// steps_per_launch S and a step budget M >= 0 the call is ceil(M/S) launches of
// at most S steps (the queue counter is reset between them); with M < 0 it is
// repeated S-step launches until a launch leaves no particle paused.
template <class F, int KIND, bool MIX>
static auto propagate_kernel() {
	if constexpr (KIND == PK_SDE)
		return k_sde<F, MIX>; // for DiffusionSDE (FP64 only) the flag selects the compiled-in module chain
	else
		return k_propagate<F, KIND, MIX>;
}

template <class F, int KIND, bool MIX>
static int launch_propagate(const F &f, const crp_propagator *p, const Params &P0, crp_batch *b, int *launches) {
	auto kern = propagate_kernel<F, KIND, MIX>();
	const int ncu = num_cus();
	if (ncu < 0)
		return CRP_E_HIP;
	const int block = (p->block == 64 || p->block == 128) ? p->block : 256;
	int bpcMax = 0;
	const size_t lds = (size_t)kColdDoubles * block * sizeof(double);
	CRP_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpcMax, kern, block, lds));
	if (bpcMax < 1)
		bpcMax = 1;
	const int bpc = p->blocksPerCU > 0 ? std::min(p->blocksPerCU, bpcMax) : bpcMax;
	const size_t want = (b->n + block - 1) / block;
	const unsigned grid = (unsigned)std::max<size_t>(1, std::min<size_t>(want, (size_t)ncu * bpc));
	p->lastGrid = grid;
	p->lastBlock = (unsigned)block;
	p->lastPPL = (double)b->n / ((double)grid * block);
	const Soa soa = soa_of(b);
	const unsigned long long n = b->n;
	const long long M = P0.maxSteps, S = p->stepsPerLaunch;
	const bool chunked = S > 0 && (M < 0 || M > S);
	const bool open = chunked && M < 0; // to completion, S steps per launch
	const long long nchunks = !chunked ? 1 : (open ? 1 : (M + S - 1) / S);

	// the launch sequence, enqueued on stream st
	auto enqueue = [&](hipStream_t st, bool zeroAll) -> int {
		if (zeroAll)
			hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, b->ctr, (int)CTR_COUNT);
		for (long long j = 0; j < nchunks; j++) {
			Params P = P0;
			if (chunked)
				P.maxSteps = open ? S : std::min(S, M - j * S);
			if (j > 0 || !zeroAll)
				hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, b->ctr, 1); // queue counter only
			hipLaunchKernelGGL(kern, dim3(grid), dim3(block), lds, st, f, P, soa, n, b->ctr);
		}
		CRP_HIP_TRY(hipGetLastError());
		return CRP_OK;
	};

	// graph key: everything the captured sequence holds
	unsigned char key[sizeof(b->gkey)];
	KeyBuf kb{key, 0, sizeof(key), p->useGraph != 0};
	kb.put((const void *)kern);
	kb.put(f);
	kb.put(P0);
	kb.put(soa);
	kb.put(n);
	kb.put((const void *)b->ctr);
	kb.put(grid);
	kb.put(block);
	kb.put(S);
	kb.put(open);
	const bool useGraph = kb.ok;

	auto run_sequence = [&](bool zeroAll) -> int {
		if (!useGraph)
			return enqueue(b->stream, zeroAll);
		KeyBuf k2 = kb;
		k2.put(zeroAll);
		if (!k2.ok)
			return enqueue(b->stream, zeroAll);
		if (!(b->gexec && b->gkeyLen == k2.len && std::memcmp(b->gkey, key, k2.len) == 0)) {
			if (b->gexec) {
				CRP_HIP_TRY(hipGraphExecDestroy(b->gexec));
				b->gexec = nullptr;
			}
			hipGraph_t g = nullptr;
			CRP_HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
			int rc = enqueue(b->stream, zeroAll);
			hipError_t e = hipStreamEndCapture(b->stream, &g);
			if (rc != CRP_OK)
				return rc;
			if (e != hipSuccess)
				return fail(CRP_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
			e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
			(void)hipGraphDestroy(g);
			if (e != hipSuccess) {
				b->gexec = nullptr;
				return fail(CRP_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
			}
			std::memcpy(b->gkey, key, k2.len);
			b->gkeyLen = k2.len;
		}
		CRP_HIP_TRY(hipGraphLaunch(b->gexec, b->stream));
		return CRP_OK;
	};

	if (!open) {
		int rc = run_sequence(true);
		if (rc != CRP_OK)
			return rc;
		*launches += (int)nchunks;
		return CRP_OK;
	}
	// open-ended chunked run: counters zeroed once, then one S-step launch at a time
	hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, b->stream, b->ctr, (int)CTR_COUNT);
	CRP_HIP_TRY(hipGetLastError());
	unsigned long long pausedBefore = 0;
	for (;;) {
		int rc = run_sequence(false);
		if (rc != CRP_OK)
			return rc;
		(*launches)++;
		unsigned long long paused = 0;
		CRP_HIP_TRY(hipMemcpyAsync(&paused, b->ctr + CTR_PAUSED, sizeof(paused), hipMemcpyDeviceToHost, b->stream));
		CRP_HIP_TRY(hipStreamSynchronize(b->stream));
		if (paused == pausedBefore)
			break;
		pausedBefore = paused;
	}
	return CRP_OK;
}

template <class F, bool MIX>
static int dispatch_kind_p(const F &f, const crp_propagator *p, const Params &P, crp_batch *b, int *launches) {
	const bool adaptive = !(p->minStep == p->maxStep);
	if (p->type == 0)
		return adaptive ? launch_propagate<F, PK_BP_ADAPTIVE, MIX>(f, p, P, b, launches)
		                : launch_propagate<F, PK_BP_FIXED, MIX>(f, p, P, b, launches);
	return adaptive ? launch_propagate<F, PK_CK_ADAPTIVE, MIX>(f, p, P, b, launches)
	                : launch_propagate<F, PK_CK_FIXED, MIX>(f, p, P, b, launches);
}

template <class F>
static int dispatch_kind(const F &f, const crp_propagator *p, const Params &P, crp_batch *b, int *launches) {
	if (p->type == 2) {
		if (p->box || p->sph || p->nObs > 0)
			return launch_propagate<F, PK_SDE, true>(f, p, P, b, launches);
		return launch_propagate<F, PK_SDE, false>(f, p, P, b, launches);
	}
	if (p->precision == CRP_PRECISION_MIXED)
		return dispatch_kind_p<F, true>(f, p, P, b, launches);
	return dispatch_kind_p<F, false>(f, p, P, b, launches);
}

// ------------------------------------------------------------------ C ABI
extern "C" {

const char *crp_last_error(void) { return g_err.c_str(); }

int crp_abi_version(void) { return CRP_ABI_VERSION; }

int crp_set_device(int device) {
	CRP_HIP_TRY(hipSetDevice(device));
	return CRP_OK;
}

int crp_get_device_count(int *count) {
	if (!count)
		return fail(CRP_E_INVALID, "crp_get_device_count: count is NULL");
	CRP_HIP_TRY(hipGetDeviceCount(count));
	return CRP_OK;
}

int crp_device_synchronize(void) {
	CRP_HIP_TRY(hipDeviceSynchronize());
	return CRP_OK;
}

static crp_field *new_field(int kind) {
	crp_field *f = new (std::nothrow) crp_field;
	if (!f)
		return nullptr;
	std::memset(f, 0, sizeof(*f));
	f->kind = kind;
	(void)hipGetDevice(&f->device);
	return f;
}

int crp_field_create_uniform(const double B[3], crp_field_t **out) {
	if (!B || !out)
		return fail(CRP_E_INVALID, "crp_field_create_uniform: NULL argument");
	crp_field *f = new_field(FK_UNIFORM);
	if (!f)
		return fail(CRP_E_NOMEM, "crp_field_create_uniform: out of host memory");
	for (int a = 0; a < 3; a++)
		f->uniform.b[a] = B[a];
	*out = f;
	return CRP_OK;
}

static int fill_grid_desc(GridDesc &g, const float *d_values, int nx, int ny, int nz, const double origin[3],
                          const double spacing[3], int reflective, int clip) {
	if (nx < 1 || ny < 1 || nz < 1)
		return fail(CRP_E_INVALID, "crp_field_create_grid3f: grid size must be >= 1 in every dimension");
	if (!origin || !spacing)
		return fail(CRP_E_INVALID, "crp_field_create_grid3f: NULL origin or spacing");
	for (int a = 0; a < 3; a++)
		if (!(spacing[a] > 0))
			return fail(CRP_E_INVALID, "crp_field_create_grid3f: spacing must be > 0");
	if ((double)nx * ny * nz * 3.0 > 9.0e18)
		return fail(CRP_E_INVALID, "crp_field_create_grid3f: grid too large");
	g.values = d_values;
	g.nx = nx;
	g.ny = ny;
	g.nz = nz;
	g.interp = kInterpTrilinear;
	g.reflective = reflective ? 1 : 0;
	g.clip = clip ? 1 : 0;
	const int N[3] = {nx, ny, nz};
	for (int a = 0; a < 3; a++) {
		g.spacing[a] = spacing[a];
		g.origin[a] = origin[a];
		g.gridOrigin[a] = origin[a] + spacing[a] / 2;      // Grid::setOrigin
		g.edge[a] = origin[a] + (double)N[a] * spacing[a]; // Grid::interpolate clip box
		g.invSpacing[a] = 1.0 / spacing[a];
		g.invN[a] = 1.0 / (double)N[a];
	}
	g.fastDiv = 1;
	for (int a = 0; a < 3; a++)
		g.fastDiv &= (spacing[a] >= 0x1p-500 && spacing[a] <= 0x1p500) ? 1 : 0;
	return CRP_OK;
}

static int build_cells(crp_field *f) {
	const GridDesc &g = f->grid;
	const size_t ncell = (size_t)g.nx * g.ny * g.nz;
	float *cells = nullptr;
	hipError_t e = hipMalloc((void **)&cells, ncell * kCellFloats * sizeof(float));
	if (e != hipSuccess)
		return fail(CRP_E_NOMEM, std::string("crp_field: cell table hipMalloc: ") + hipGetErrorString(e));
	const unsigned blocks = (unsigned)std::min<size_t>((ncell + 255) / 256, 65536);
	hipLaunchKernelGGL(k_build_cells, dim3(blocks), dim3(256), 0, 0, g.values, cells, g.nx, g.ny, g.nz);
	e = hipGetLastError();
	if (e == hipSuccess)
		e = hipDeviceSynchronize();
	if (e != hipSuccess) {
		(void)hipFree(cells);
		return fail(CRP_E_HIP, std::string("crp_field: building the cell table: ") + hipGetErrorString(e));
	}
	if (f->cells)
		(void)hipFree(f->cells);
	f->cells = cells;
	return CRP_OK;
}

/* Periodic grids get the cell table by default; if it does not fit in device
 * memory the point layout is kept (same results, more traffic). */
static int build_cells_default(crp_field *f) {
	if (f->grid.reflective)
		return CRP_OK;
	if (build_cells(f) != CRP_OK)
		(void)hipGetLastError();
	return CRP_OK;
}

// a grid or modulated field with a Grid3f or Grid1f that is not read trilinear: FieldGridAny
static bool any_interp(const crp_field *f) {
	if (!f)
		return false;
	if (f->kind == FK_GRID)
		return f->grid.interp != kInterpTrilinear;
	if (f->kind == FK_MODULATED)
		return f->grid.interp != kInterpTrilinear || f->mod.interp != kInterpTrilinear;
	return false;
}

// a JF12 field whose turbulent grid is not read trilinear: FieldJF12Any
static bool jf12_any(const crp_field *f) {
	return f && f->kind == FK_JF12 && f->jf12.turbulent.interp != kInterpTrilinear;
}

static FieldGridAny grid_any(const crp_field *f) {
	FieldGridAny a;
	std::memset(&a, 0, sizeof(a));
	a.g = f->grid;
	if (f->kind == FK_MODULATED) {
		a.mod = f->mod;
		a.modulated = 1;
	}
	return a;
}

int crp_field_create_grid3f(const float *values, int nx, int ny, int nz, const double origin[3],
                            const double spacing[3], int reflective, int clip_volume, crp_field_t **out) {
	if (!values || !out)
		return fail(CRP_E_INVALID, "crp_field_create_grid3f: NULL argument");
	crp_field *f = new_field(FK_GRID);
	if (!f)
		return fail(CRP_E_NOMEM, "crp_field_create_grid3f: out of host memory");
	int rc = fill_grid_desc(f->grid, nullptr, nx, ny, nz, origin, spacing, reflective, clip_volume);
	if (rc != CRP_OK) {
		delete f;
		return rc;
	}
	size_t bytes = (size_t)nx * ny * nz * 3 * sizeof(float);
	void *d = nullptr;
	hipError_t e = hipMalloc(&d, bytes);
	if (e != hipSuccess) {
		delete f;
		return fail(CRP_E_NOMEM, std::string("crp_field_create_grid3f: hipMalloc: ") + hipGetErrorString(e));
	}
	e = hipMemcpy(d, values, bytes, hipMemcpyHostToDevice);
	if (e != hipSuccess) {
		(void)hipFree(d);
		delete f;
		return fail(CRP_E_HIP, std::string("crp_field_create_grid3f: hipMemcpy: ") + hipGetErrorString(e));
	}
	f->grid.values = (const float *)d;
	f->owned[0] = d;
	*out = f;
	return build_cells_default(f);
}

int crp_field_create_grid3f_device(const float *d_values, int nx, int ny, int nz, const double origin[3],
                                   const double spacing[3], int reflective, int clip_volume, crp_field_t **out) {
	if (!d_values || !out)
		return fail(CRP_E_INVALID, "crp_field_create_grid3f_device: NULL argument");
	crp_field *f = new_field(FK_GRID);
	if (!f)
		return fail(CRP_E_NOMEM, "crp_field_create_grid3f_device: out of host memory");
	int rc = fill_grid_desc(f->grid, d_values, nx, ny, nz, origin, spacing, reflective, clip_volume);
	if (rc != CRP_OK) {
		delete f;
		return rc;
	}
	*out = f;
	return build_cells_default(f);
}

int crp_field_create_modulated_grid3f(const float *values, int nx, int ny, int nz, const double origin[3],
                                      const double spacing[3], int reflective, int clip_volume, const float *modulation,
                                      int mnx, int mny, int mnz, const double morigin[3], const double mspacing[3],
                                      int mreflective, int mclip_volume, crp_field_t **out) {
	if (!modulation || !out)
		return fail(CRP_E_INVALID, "crp_field_create_modulated_grid3f: NULL argument");
	crp_field *f = nullptr;
	int rc = crp_field_create_grid3f(values, nx, ny, nz, origin, spacing, reflective, clip_volume, &f);
	if (rc != CRP_OK)
		return rc;
	f->kind = FK_MODULATED;
	rc = fill_grid_desc(f->mod, nullptr, mnx, mny, mnz, morigin, mspacing, mreflective, mclip_volume);
	if (rc != CRP_OK) {
		(void)crp_field_destroy(f);
		return rc;
	}
	const size_t bytes = (size_t)mnx * mny * mnz * sizeof(float);
	void *d = nullptr;
	hipError_t e = hipMalloc(&d, bytes);
	if (e == hipSuccess)
		e = hipMemcpy(d, modulation, bytes, hipMemcpyHostToDevice);
	if (e != hipSuccess) {
		if (d)
			(void)hipFree(d);
		(void)crp_field_destroy(f);
		return fail(CRP_E_HIP, std::string("crp_field_create_modulated_grid3f: ") + hipGetErrorString(e));
	}
	f->mod.values = (const float *)d;
	f->owned[1] = d;
	*out = f;
	return CRP_OK;
}

int crp_jf12_default_params(int planck, crp_jf12_params *out) {
	// JF12Field::JF12Field (src/magneticField/JF12Field.cpp:8-79) and PlanckJF12bField (:348-363),
	// with the units of include/crpropa/Units.h
	if (!out)
		return fail(CRP_E_INVALID, "crp_jf12_default_params: NULL argument");
	const double tesla = 1.0, gauss = 1e-4 * tesla, muG = 1e-6 * gauss;
	const double kpc = 1e3 * (648000 / M_PI * 149597870700.0);
	crp_jf12_params &p = *out;
	std::memset(&p, 0, sizeof(p));
	p.use_regular = 1;
	p.use_disk = 1;
	p.use_toroidal_halo = 1;
	p.use_x = 1;
	volatile double pitchv = 11.5 * M_PI / 180; // computed at run time with the host libm
	const double pitch = pitchv;
	p.sin_pitch = std::sin(pitch);
	p.cos_pitch = std::cos(pitch);
	volatile double hp = M_PI / 2 - pitch;
	p.tan90_minus_pitch = std::tan((double)hp);
	const double arms[8] = {5.1, 6.3, 7.1, 8.3, 9.8, 11.4, 12.7, 15.5};
	const double disk[8] = {0.1, 3.0, -0.9, -0.8, -2.0, -4.2, 0.0, 2.7};
	const double turb[8] = {10.81, 6.96, 9.59, 6.96, 1.96, 16.34, 37.29, 10.35};
	for (int i = 0; i < 8; i++) {
		p.r_arms[i] = arms[i] * kpc;
		p.b_disk[i] = disk[i] * muG;
		p.b_disk_turb[i] = turb[i] * muG;
	}
	p.b_ring = 0.1 * muG;
	p.h_disk = 0.40 * kpc;
	p.w_disk = 0.27 * kpc;
	p.b_north = 1.4 * muG;
	p.b_south = -1.1 * muG;
	p.r_north = 9.22 * kpc;
	p.r_south = 17 * kpc;
	p.w_halo = 0.20 * kpc;
	p.z0 = 5.3 * kpc;
	p.b_x = 4.6 * muG;
	volatile double thv = 49.0 * M_PI / 180;
	const double thetaX0 = thv;
	p.sin_theta_x0 = std::sin(thetaX0);
	p.cos_theta_x0 = std::cos(thetaX0);
	p.tan_theta_x0 = std::tan(thetaX0);
	p.r_xc = 4.8 * kpc;
	p.r_x = 2.9 * kpc;
	volatile double beta = 1.36;
	p.sqrt_beta = std::sqrt((double)beta);
	p.b_disk_turb5 = 7.63 * muG;
	p.z_disk_turb = 0.61 * kpc;
	p.b_halo_turb = 4.68 * muG;
	p.r_halo_turb = 10.97 * kpc;
	p.z_halo_turb = 2.84 * kpc;
	if (planck) {
		const double pt[8] = {3.12, 6.24, 3.12, 6.24, 3.12, 6.24, 3.12, 6.24};
		p.b_disk[5] = -3.5 * muG;
		p.b_x = 1.8 * muG;
		for (int i = 0; i < 8; i++)
			p.b_disk_turb[i] = pt[i] * muG;
		p.b_disk_turb5 = 3.90 * muG;
		p.b_halo_turb = 7.332 * muG;
	}
	return CRP_OK;
}

int crp_field_create_jf12(const crp_jf12_params *p, const float *striated, const int striated_n[3],
                          const double striated_origin[3], const double striated_spacing[3], int striated_reflective,
                          const float *turbulent, const int turbulent_n[3], const double turbulent_origin[3],
                          const double turbulent_spacing[3], int turbulent_reflective, int turbulent_clip,
                          crp_field_t **out) {
	if (!p || !out)
		return fail(CRP_E_INVALID, "crp_field_create_jf12: NULL argument");
	if (p->use_striated && (!striated || !striated_n))
		return fail(CRP_E_INVALID, "crp_field_create_jf12: use_striated needs a striated grid");
	if (p->use_turbulent && (!turbulent || !turbulent_n))
		return fail(CRP_E_INVALID, "crp_field_create_jf12: use_turbulent needs a turbulent grid");
	crp_field *f = new_field(FK_JF12);
	if (!f)
		return fail(CRP_E_NOMEM, "crp_field_create_jf12: out of host memory");
	std::memset(&f->jf12, 0, sizeof(f->jf12));
	f->jf12.p = *p;
	auto upload = [&](GridDesc &g, const float *v, const int n[3], const double *o, const double *sp, int refl, int clip,
	                  int comps, int slot) -> int {
		int rc = fill_grid_desc(g, nullptr, n[0], n[1], n[2], o, sp, refl, clip);
		if (rc != CRP_OK)
			return rc;
		const size_t bytes = (size_t)n[0] * n[1] * n[2] * comps * sizeof(float);
		void *d = nullptr;
		hipError_t e = hipMalloc(&d, bytes);
		if (e == hipSuccess)
			e = hipMemcpy(d, v, bytes, hipMemcpyHostToDevice);
		if (e != hipSuccess) {
			if (d)
				(void)hipFree(d);
			return fail(CRP_E_HIP, std::string("crp_field_create_jf12: ") + hipGetErrorString(e));
		}
		g.values = (const float *)d;
		f->owned[slot] = d;
		return CRP_OK;
	};
	int rc = CRP_OK;
	if (striated && striated_n)
		rc = upload(f->jf12.striated, striated, striated_n, striated_origin, striated_spacing, striated_reflective, 0, 1, 0);
	if (rc == CRP_OK && turbulent && turbulent_n)
		rc = upload(f->jf12.turbulent, turbulent, turbulent_n, turbulent_origin, turbulent_spacing, turbulent_reflective,
		            turbulent_clip, 3, 1);
	if (rc != CRP_OK) {
		(void)crp_field_destroy(f);
		return rc;
	}
	*out = f;
	return CRP_OK;
}

int crp_field_create_jf12_view(const crp_field_t *base, const crp_jf12_params *p, crp_field_t **out) {
	if (!base || !p || !out)
		return fail(CRP_E_INVALID, "crp_field_create_jf12_view: NULL argument");
	if (base->kind != FK_JF12)
		return fail(CRP_E_INVALID, "crp_field_create_jf12_view: base is not a JF12 field");
	if (p->use_striated && !base->jf12.striated.values)
		return fail(CRP_E_INVALID, "crp_field_create_jf12_view: use_striated, but base holds no striated grid");
	if (p->use_turbulent && !base->jf12.turbulent.values)
		return fail(CRP_E_INVALID, "crp_field_create_jf12_view: use_turbulent, but base holds no turbulent grid");
	crp_field *f = new_field(FK_JF12);
	if (!f)
		return fail(CRP_E_NOMEM, "crp_field_create_jf12_view: out of host memory");
	f->jf12 = base->jf12; // the grid descriptors point at base's memory; nothing is owned
	f->jf12.p = *p;
	*out = f;
	return CRP_OK;
}

int crp_field_set_grid_layout(crp_field_t *f, int layout) {
	if (!f || (f->kind != FK_GRID && f->kind != FK_MODULATED))
		return fail(CRP_E_INVALID, "crp_field_set_grid_layout: not a grid field");
	if (layout == CRP_GRID_LAYOUT_POINTS) {
		if (f->cells) {
			CRP_HIP_TRY(hipFree(f->cells));
			f->cells = nullptr;
		}
		return CRP_OK;
	}
	if (layout != CRP_GRID_LAYOUT_CELLS)
		return fail(CRP_E_INVALID, "crp_field_set_grid_layout: unknown layout");
	if (f->grid.reflective)
		return fail(CRP_E_INVALID, "crp_field_set_grid_layout: the cell table needs a periodic grid");
	if (f->grid.interp != kInterpTrilinear)
		return fail(CRP_E_INVALID, "crp_field_set_grid_layout: the cell table needs trilinear interpolation");
	return f->cells ? CRP_OK : build_cells(f);
}

int crp_field_set_grid_storage(crp_field_t *f, int storage) {
	if (!f || f->kind != FK_GRID)
		return fail(CRP_E_INVALID, "crp_field_set_grid_storage: not a grid field");
	if (storage == CRP_GRID_STORAGE_FP32) {
		if (f->cells16) {
			CRP_HIP_TRY(hipFree(f->cells16));
			f->cells16 = nullptr;
		}
		return CRP_OK;
	}
	if (storage != CRP_GRID_STORAGE_FP16_BRICK)
		return fail(CRP_E_INVALID, "crp_field_set_grid_storage: unknown storage");
	if (f->grid.reflective)
		return fail(CRP_E_INVALID, "crp_field_set_grid_storage: FP16 storage needs a periodic grid");
	if (f->grid.interp != kInterpTrilinear)
		return fail(CRP_E_INVALID, "crp_field_set_grid_storage: FP16 storage needs trilinear interpolation");
	if (f->cells16)
		return CRP_OK;
	const GridDesc &g = f->grid;
	const int bx = (g.nx + 7) / 8, by = (g.ny + 7) / 8, bz = (g.nz + 7) / 8;
	const size_t nb = (size_t)bx * by * bz, ncell = (size_t)g.nx * g.ny * g.nz;
	float *scales = nullptr;
	unsigned char *cells = nullptr;
	hipError_t e = hipMalloc((void **)&scales, nb * sizeof(float));
	if (e == hipSuccess)
		e = hipMalloc((void **)&cells, ncell * kCell16Bytes);
	if (e != hipSuccess) {
		if (scales)
			(void)hipFree(scales);
		return fail(CRP_E_NOMEM, std::string("crp_field_set_grid_storage: hipMalloc: ") + hipGetErrorString(e));
	}
	hipLaunchKernelGGL(k_brick_scales, dim3((unsigned)nb), dim3(256), 0, 0, g.values, scales, g.nx, g.ny, g.nz, bx, by, bz);
	const unsigned blocks = (unsigned)std::min<size_t>((ncell + 255) / 256, 65536);
	hipLaunchKernelGGL(k_build_cells16, dim3(blocks), dim3(256), 0, 0, g.values, scales, cells, g.nx, g.ny, g.nz, by, bz);
	e = hipGetLastError();
	if (e == hipSuccess)
		e = hipDeviceSynchronize();
	(void)hipFree(scales);
	if (e != hipSuccess) {
		(void)hipFree(cells);
		return fail(CRP_E_HIP, std::string("crp_field_set_grid_storage: ") + hipGetErrorString(e));
	}
	f->cells16 = cells;
	return CRP_OK;
}

static_assert((int)kInterpTrilinear == (int)CRP_GRID_TRILINEAR && (int)kInterpTricubic == (int)CRP_GRID_TRICUBIC &&
                  (int)kInterpNearest == (int)CRP_GRID_NEAREST_NEIGHBOUR,
              "device interpolation numbers follow crp_grid_interpolation");

// the grid of f that `grid` (crp_grid_select) names, or nullptr
static GridDesc *interp_grid(crp_field *f, int grid) {
	if (!f)
		return nullptr;
	if (grid == CRP_GRID_VALUES && (f->kind == FK_GRID || f->kind == FK_MODULATED))
		return &f->grid;
	if (grid == CRP_GRID_VALUES && f->kind == FK_JF12 && f->jf12.turbulent.values)
		return &f->jf12.turbulent;
	if (grid == CRP_GRID_MODULATION && f->kind == FK_MODULATED)
		return &f->mod;
	return nullptr;
}

int crp_field_set_grid_interpolation(crp_field_t *f, int grid, int type) {
	GridDesc *g = interp_grid(f, grid);
	if (!g)
		return fail(CRP_E_INVALID, "crp_field_set_grid_interpolation: the field holds no such grid");
	if (type != CRP_GRID_TRILINEAR && type != CRP_GRID_TRICUBIC && type != CRP_GRID_NEAREST_NEIGHBOUR)
		return fail(CRP_E_INVALID, "InterpolationType: unknown interpolation type");
	if (g->interp == type)
		return CRP_OK;
	g->interp = type;
	if (g != &f->grid)
		return CRP_OK;
	if (type == CRP_GRID_TRILINEAR)
		return build_cells_default(f); // periodic grids get their cell table back
	// the tables hold trilinear corners: read the points instead
	if (f->cells) {
		CRP_HIP_TRY(hipFree(f->cells));
		f->cells = nullptr;
	}
	if (f->cells16) {
		CRP_HIP_TRY(hipFree(f->cells16));
		f->cells16 = nullptr;
	}
	return CRP_OK;
}

int crp_field_get_grid_interpolation(const crp_field_t *f, int grid, int *type) {
	const GridDesc *g = interp_grid(const_cast<crp_field *>(f), grid);
	if (!g || !type)
		return fail(CRP_E_INVALID, "crp_field_get_grid_interpolation: the field holds no such grid");
	*type = g->interp;
	return CRP_OK;
}

// ---------------------------------------------------------------- field decorators (host side)
// Compiles a tree of PeriodicMagneticField / MagneticFieldList / MagneticFieldEvolution over
// device fields into the program FieldComposite runs (crp_composite.hpp).
struct CompBuilder {
	CompositeDesc D;
	int nl = 0, np = 0;
	std::string err;

	bool emit(int op, int a, int b, int c) {
		if (D.nops >= kCompMaxOps) {
			err = "more than " + std::to_string(kCompMaxOps) + " operations";
			return false;
		}
		D.ops[D.nops++] = CompOp{op, a, b, c};
		return true;
	}
	int leaf(const crp_field *f) {
		if (nl >= kCompMaxLeaves) {
			err = "more than " + std::to_string(kCompMaxLeaves) + " fields below the decorators";
			return -1;
		}
		CompLeaf &L = D.leaves[nl];
		std::memset(&L, 0, sizeof(L));
		switch (f ? f->kind : FK_NONE) {
		case FK_UNIFORM: L.kind = CL_UNIFORM; L.uniform = f->uniform; break;
		case FK_GRID:
			L.kind = any_interp(f) ? CL_GRID_ANY : CL_GRID;
			L.grid = f->grid;
			L.cells = f->cells;
			L.cells16 = f->cells16;
			break;
		case FK_MODULATED:
			L.kind = any_interp(f) ? CL_GRID_ANY : CL_MODULATED;
			L.grid = f->grid;
			L.cells = f->cells;
			L.mod = f->mod;
			break;
		case FK_PLANEWAVE: L.kind = CL_PLANEWAVE; L.pw = f->pw; break;
		case FK_JF12: L.kind = CL_JF12; L.jf12 = f->jf12; break;
		default: L.kind = CL_NONE; break;
		}
		return nl++;
	}
	// code for getField(pos[ps]) of f (the one-argument call), stored into acc[as] or added to it
	bool gen(const crp_field *f, int ps, int as, bool add) {
		if (!f || f->kind != FK_COMPOSITE) {
			const int l = leaf(f);
			return l >= 0 && emit(add ? CO_LEAF_ADD : CO_LEAF, l, ps, as);
		}
		switch (f->deco) {
		case DECO_PERIODIC: {
			if (ps + 1 >= kCompSlots || np >= kCompMaxPeriodic) {
				err = "periodic fields nested more than " + std::to_string(kCompSlots - 1) + " deep";
				return false;
			}
			CompPeriodic &c = D.per[np];
			for (int a = 0; a < 3; a++) {
				c.origin[a] = f->porigin[a];
				c.extends[a] = f->extends[a];
			}
			c.reflective = f->preflective;
			return emit(CO_PERIODIC, np++, ps, ps + 1) && gen(f->child[0], ps + 1, as, add);
		}
		case DECO_LIST: {
			// b = (0,0,0); b += field_i->getField(position): a list inside a sum is summed on its own
			const int t = add ? as + 1 : as;
			if (t >= kCompSlots) {
				err = "field lists nested more than " + std::to_string(kCompSlots - 1) + " deep";
				return false;
			}
			if (!emit(CO_ZERO, t, 0, 0))
				return false;
			for (int i = 0; i < f->nchild; i++)
				if (!gen(f->child[i], ps, t, true))
					return false;
			return add ? emit(CO_ADD, t, as, 0) : true;
		}
		default:
			// a MagneticFieldEvolution below the top: its one-argument getField is the base class's
			// zero; adding +0 to a sum that starts at +0 changes nothing
			return add ? true : emit(CO_ZERO, as, 0, 0);
		}
	}
	bool build(const crp_field *top) {
		std::memset(&D, 0, sizeof(D));
		if (top->deco == DECO_EVOLUTION) {
			D.evolve = 1;
			D.evoM = top->evoM;
			return gen(top->child[0], 0, 0, false) && emit(CO_EVOLVE, 0, 0, 0);
		}
		return gen(top, 0, 0, false);
	}
};

// Compiles f's tree and uploads the program when it changed (a child's device layout may have).
static int composite_refresh(const crp_field *cf) {
	crp_field *f = const_cast<crp_field *>(cf);
	CompBuilder B;
	if (!B.build(f))
		return fail(CRP_E_INVALID, "field decorator: " + B.err + " (not supported on the device)");
	if (!f->dcomp || !f->hcomp)
		return fail(CRP_E_STATE, "field decorator: no device program");
	if (std::memcmp(f->hcomp, &B.D, sizeof(CompositeDesc)) != 0) {
		CRP_HIP_TRY(hipMemcpy(f->dcomp, &B.D, sizeof(CompositeDesc), hipMemcpyHostToDevice));
		std::memcpy(f->hcomp, &B.D, sizeof(CompositeDesc));
	}
	return CRP_OK;
}

// Before a field is used: a decorator's program is brought up to date (no allocation).
static int field_prepare(const crp_field *f) {
	return (f && f->kind == FK_COMPOSITE) ? composite_refresh(f) : CRP_OK;
}

// getField(position) of a decorator tree whose top is a MagneticFieldEvolution is zero: the class
// overrides only getField(position, z)
static bool composite_one_arg_zero(const crp_field *f) {
	return f && f->kind == FK_COMPOSITE && f->deco == DECO_EVOLUTION;
}

// Launches the field evaluation of n device positions dpos into dB (both AoS xyz) on the null stream:
// the FP64 evaluation (k_field_eval) or the mixed path's FP32 evaluation (k_field_evalf, FP16 storage
// if set).  Shared by crp_field_eval, crp_field_eval_ex and crp_field_eval_noalloc.
static hipError_t launch_field_eval(const crp_field_t *f, const double *dpos, double *dB, size_t n, int precision) {
	const unsigned blocks = (unsigned)((n + 255) / 256);
	const int kind = f ? f->kind : FK_NONE;
	if (composite_one_arg_zero(f))
		return hipMemsetAsync(dB, 0, 3 * n * sizeof(double), 0);
	if (any_interp(f)) {
		if (precision == CRP_PRECISION_FP64)
			hipLaunchKernelGGL(k_field_eval<FieldGridAny>, blocks, 256, 0, 0, grid_any(f), dpos, dB, n);
		else
			hipLaunchKernelGGL(k_field_evalf<FieldGridAny>, blocks, 256, 0, 0, grid_any(f), dpos, dB, n);
		return hipGetLastError();
	}
	if (precision == CRP_PRECISION_FP64) {
		switch (kind) {
		case FK_UNIFORM: hipLaunchKernelGGL(k_field_eval<FieldUniform>, blocks, 256, 0, 0, f->uniform, dpos, dB, n); break;
		case FK_GRID:
			if (f->cells)
				hipLaunchKernelGGL(k_field_eval<FieldGridCells>, blocks, 256, 0, 0, FieldGridCells{f->grid, f->cells},
				                   dpos, dB, n);
			else
				hipLaunchKernelGGL(k_field_eval<FieldGrid>, blocks, 256, 0, 0, FieldGrid{f->grid}, dpos, dB, n);
			break;
		case FK_PLANEWAVE: hipLaunchKernelGGL(k_field_eval<FieldPlaneWave>, blocks, 256, 0, 0, FieldPlaneWave{f->pw}, dpos, dB, n); break;
		case FK_MODULATED:
			if (f->cells)
				hipLaunchKernelGGL(k_field_eval<FieldModulated<FieldGridCells>>, blocks, 256, 0, 0,
				                   FieldModulated<FieldGridCells>{FieldGridCells{f->grid, f->cells}, f->mod}, dpos, dB, n);
			else
				hipLaunchKernelGGL(k_field_eval<FieldModulated<FieldGrid>>, blocks, 256, 0, 0,
				                   FieldModulated<FieldGrid>{FieldGrid{f->grid}, f->mod}, dpos, dB, n);
			break;
		case FK_JF12:
			if (jf12_any(f))
				hipLaunchKernelGGL(k_field_eval<FieldJF12Any>, blocks, 256, 0, 0, FieldJF12Any{f->jf12}, dpos, dB, n);
			else
				hipLaunchKernelGGL(k_field_eval<FieldJF12>, blocks, 256, 0, 0, f->jf12, dpos, dB, n);
			break;
		case FK_COMPOSITE:
			hipLaunchKernelGGL(k_field_eval<FieldComposite>, blocks, 256, 0, 0, FieldComposite{f->dcomp, 1.0}, dpos, dB, n);
			break;
		default: hipLaunchKernelGGL(k_field_eval<FieldNone>, blocks, 256, 0, 0, FieldNone{}, dpos, dB, n); break;
		}
	} else {
		switch (kind) {
		case FK_UNIFORM: hipLaunchKernelGGL(k_field_evalf<FieldUniform>, blocks, 256, 0, 0, f->uniform, dpos, dB, n); break;
		case FK_GRID:
			if (f->cells16)
				hipLaunchKernelGGL(k_field_evalf<FieldGridCells16>, blocks, 256, 0, 0,
				                   FieldGridCells16{f->grid, f->cells16}, dpos, dB, n);
			else if (f->cells)
				hipLaunchKernelGGL(k_field_evalf<FieldGridCells>, blocks, 256, 0, 0, FieldGridCells{f->grid, f->cells},
				                   dpos, dB, n);
			else
				hipLaunchKernelGGL(k_field_evalf<FieldGrid>, blocks, 256, 0, 0, FieldGrid{f->grid}, dpos, dB, n);
			break;
		case FK_PLANEWAVE: hipLaunchKernelGGL(k_field_evalf<FieldPlaneWave>, blocks, 256, 0, 0, FieldPlaneWave{f->pw}, dpos, dB, n); break;
		case FK_MODULATED:
			if (f->cells)
				hipLaunchKernelGGL(k_field_evalf<FieldModulated<FieldGridCells>>, blocks, 256, 0, 0,
				                   FieldModulated<FieldGridCells>{FieldGridCells{f->grid, f->cells}, f->mod}, dpos, dB, n);
			else
				hipLaunchKernelGGL(k_field_evalf<FieldModulated<FieldGrid>>, blocks, 256, 0, 0,
				                   FieldModulated<FieldGrid>{FieldGrid{f->grid}, f->mod}, dpos, dB, n);
			break;
		case FK_JF12:
			if (jf12_any(f))
				hipLaunchKernelGGL(k_field_evalf<FieldJF12Any>, blocks, 256, 0, 0, FieldJF12Any{f->jf12}, dpos, dB, n);
			else
				hipLaunchKernelGGL(k_field_evalf<FieldJF12>, blocks, 256, 0, 0, f->jf12, dpos, dB, n);
			break;
		case FK_COMPOSITE:
			hipLaunchKernelGGL(k_field_evalf<FieldComposite>, blocks, 256, 0, 0, FieldComposite{f->dcomp, 1.0}, dpos, dB, n);
			break;
		default: hipLaunchKernelGGL(k_field_evalf<FieldNone>, blocks, 256, 0, 0, FieldNone{}, dpos, dB, n); break;
		}
	}
	return hipGetLastError();
}

int crp_field_eval_ex(const crp_field_t *f, const double *pos, size_t n, double *B, int precision) {
	if (precision == CRP_PRECISION_FP64)
		return crp_field_eval(f, pos, n, B);
	if (precision != CRP_PRECISION_MIXED)
		return fail(CRP_E_INVALID, "crp_field_eval_ex: unknown precision");
	if (!pos || !B)
		return fail(CRP_E_INVALID, "crp_field_eval_ex: NULL argument");
	if (n == 0)
		return CRP_OK;
	if (int rc = field_prepare(f))
		return rc;
	double *dpos = nullptr, *dB = nullptr;
	CRP_HIP_TRY(hipMalloc(&dpos, 3 * n * sizeof(double)));
	hipError_t e = hipMalloc(&dB, 3 * n * sizeof(double));
	if (e == hipSuccess)
		e = hipMemcpy(dpos, pos, 3 * n * sizeof(double), hipMemcpyHostToDevice);
	if (e == hipSuccess)
		e = launch_field_eval(f, dpos, dB, n, CRP_PRECISION_MIXED);
	if (e == hipSuccess)
		e = hipMemcpy(B, dB, 3 * n * sizeof(double), hipMemcpyDeviceToHost);
	if (dpos)
		(void)hipFree(dpos);
	if (dB)
		(void)hipFree(dB);
	if (e != hipSuccess)
		return fail(CRP_E_HIP, std::string("crp_field_eval_ex: ") + hipGetErrorString(e));
	return CRP_OK;
}

// ---------------------------------------------------------------- field evaluation without allocation
// Staging of crp_field_eval_noalloc: CRP_EVAL_NOALLOC_MAX positions followed by as many field values
// (AoS xyz doubles), 6 MiB per device for the life of the process.  Host code never uses this name as
// an address: the device address comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) double scratch[2 * 3 * CRP_EVAL_NOALLOC_MAX];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[kMaxDevices] = {};
} // namespace

// The device address of `scratch` on the current device (looked up on the first call per device).
static int noalloc_scratch(double **out) {
	int dev = 0;
	hipError_t e = hipGetDevice(&dev);
	if (e != hipSuccess)
		return fail(CRP_E_HIP, std::string("crp_field_eval_noalloc: hipGetDevice: ") + hipGetErrorString(e));
	if (dev < 0 || dev >= kMaxDevices)
		return fail(CRP_E_INVALID, "crp_field_eval_noalloc: device index beyond the supported 64");
	std::lock_guard<std::mutex> lock(g_scratch_mutex);
	if (!g_scratch_ptr[dev]) {
		void *p = nullptr;
		e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !p)
			return fail(CRP_E_HIP, std::string("crp_field_eval_noalloc: hipGetSymbolAddress: ") +
			                           (e != hipSuccess ? hipGetErrorString(e) : "null address"));
		g_scratch_ptr[dev] = (double *)p;
	}
	*out = g_scratch_ptr[dev];
	return CRP_OK;
}

int crp_field_eval_noalloc(const crp_field_t *f, const double *pos, size_t n, double *B, int precision) {
	if (precision != CRP_PRECISION_FP64 && precision != CRP_PRECISION_MIXED)
		return fail(CRP_E_INVALID, "crp_field_eval_noalloc: unknown precision");
	if (!pos || !B)
		return fail(CRP_E_INVALID, "crp_field_eval_noalloc: NULL argument");
	if (n > CRP_EVAL_NOALLOC_MAX)
		return fail(CRP_E_INVALID, "crp_field_eval_noalloc: " + std::to_string(n) +
		                               " positions exceed the staging capacity of " +
		                               std::to_string((unsigned long long)CRP_EVAL_NOALLOC_MAX) +
		                               " per call (split the call, or use crp_field_eval)");
	if (n == 0)
		return CRP_OK;
	if (int rc = field_prepare(f))
		return rc;
	double *st = nullptr;
	if (int rc = noalloc_scratch(&st))
		return rc;
	double *dpos = st, *dB = st + 3 * (size_t)CRP_EVAL_NOALLOC_MAX;
	hipError_t e = hipMemcpy(dpos, pos, 3 * n * sizeof(double), hipMemcpyHostToDevice);
	if (e == hipSuccess)
		e = launch_field_eval(f, dpos, dB, n, precision);
	if (e == hipSuccess)
		e = hipMemcpy(B, dB, 3 * n * sizeof(double), hipMemcpyDeviceToHost);
	if (e != hipSuccess)
		return fail(CRP_E_HIP, std::string("crp_field_eval_noalloc: ") + hipGetErrorString(e));
	return CRP_OK;
}

int crp_field_get_grid_layout(const crp_field_t *f, int *layout) {
	if (!f || !layout || f->kind != FK_GRID)
		return fail(CRP_E_INVALID, "crp_field_get_grid_layout: not a grid field");
	*layout = f->cells ? CRP_GRID_LAYOUT_CELLS : CRP_GRID_LAYOUT_POINTS;
	return CRP_OK;
}

int crp_field_create_planewave(int nm, const double *k, const double *Ak, const double *xi, const double *kappa,
                               const double *beta, crp_field_t **out) {
	if (!k || !Ak || !xi || !kappa || !beta || !out)
		return fail(CRP_E_INVALID, "crp_field_create_planewave: NULL argument");
	if (nm <= 1) // PlaneWaveTurbulence.cpp:93-97
		return fail(CRP_E_INVALID, "crp_field_create_planewave: Nm <= 1. Specify at least two wavemodes");
	crp_field *f = new_field(FK_PLANEWAVE);
	if (!f)
		return fail(CRP_E_NOMEM, "crp_field_create_planewave: out of host memory");
	// xi*Ak is the first product of B += xi[i] * Ak[i] * cos(...): formed once here, same IEEE op.
	// Mixed-precision records: phase in turns = p . (k kappa / 2pi) + beta / 2pi; amplitudes in FP32.
	const double inv2pi = 1.0 / (2.0 * M_PI);
	std::vector<PwMode64> m64(nm);
	std::vector<PwModeMixed> mm(nm);
	for (int i = 0; i < nm; i++) {
		PwMode64 &a = m64[i];
		a.k = k[i];
		a.beta = beta[i];
		a.kx = kappa[3 * i];
		a.ky = kappa[3 * i + 1];
		a.kz = kappa[3 * i + 2];
		a.ax = xi[3 * i] * Ak[i];
		a.ay = xi[3 * i + 1] * Ak[i];
		a.az = xi[3 * i + 2] * Ak[i];
		PwModeMixed &b = mm[i];
		b.ktx = k[i] * kappa[3 * i] * inv2pi;
		b.kty = k[i] * kappa[3 * i + 1] * inv2pi;
		b.ktz = k[i] * kappa[3 * i + 2] * inv2pi;
		b.bt = beta[i] * inv2pi;
		b.ax = (float)a.ax;
		b.ay = (float)a.ay;
		b.az = (float)a.az;
		b.pad = 0.f;
	}
	const void *src[2] = {m64.data(), mm.data()};
	const size_t sz[2] = {(size_t)nm * sizeof(PwMode64), (size_t)nm * sizeof(PwModeMixed)};
	for (int j = 0; j < 2; j++) {
		void *d = nullptr;
		hipError_t e = hipMalloc(&d, sz[j]);
		if (e == hipSuccess)
			e = hipMemcpy(d, src[j], sz[j], hipMemcpyHostToDevice);
		if (e != hipSuccess) {
			if (d)
				(void)hipFree(d);
			for (int i = 0; i < j; i++)
				(void)hipFree(f->owned[i]);
			delete f;
			return fail(CRP_E_HIP, std::string("crp_field_create_planewave: ") + hipGetErrorString(e));
		}
		f->owned[j] = d;
	}
	f->pw.nm = nm;
	f->pw.m64 = (const PwMode64 *)f->owned[0];
	f->pw.mm = (const PwModeMixed *)f->owned[1];
	*out = f;
	return CRP_OK;
}

int crp_field_destroy(crp_field_t *f) {
	if (!f)
		return CRP_OK;
	for (int i = 0; i < 8; i++)
		if (f->owned[i])
			(void)hipFree(f->owned[i]);
	if (f->cells)
		(void)hipFree(f->cells);
	if (f->cells16)
		(void)hipFree(f->cells16);
	if (f->dcomp)
		(void)hipFree(f->dcomp);
	delete f->hcomp;
	delete f;
	return CRP_OK;
}

// ---------------------------------------------------------------- decorators (C ABI)
static int new_composite(int deco, crp_field_t **out, crp_field **res) {
	crp_field *f = new_field(FK_COMPOSITE);
	if (!f)
		return fail(CRP_E_NOMEM, "field decorator: out of host memory");
	f->deco = deco;
	hipError_t e = hipMalloc((void **)&f->dcomp, sizeof(CompositeDesc));
	f->hcomp = new (std::nothrow) CompositeDesc;
	if (e != hipSuccess || !f->hcomp) {
		if (e == hipSuccess)
			(void)hipFree(f->dcomp);
		delete f->hcomp;
		delete f;
		return fail(CRP_E_NOMEM, "field decorator: device or host memory");
	}
	std::memset(f->hcomp, 0xff, sizeof(CompositeDesc)); // differs from any program: the first use uploads
	*res = f;
	*out = f;
	return CRP_OK;
}

// The tree below f must compile (depth and size limits of the device program).
static int composite_check(crp_field *f) {
	CompBuilder B;
	if (!B.build(f))
		return fail(CRP_E_INVALID, "field decorator: " + B.err + " (not supported on the device)");
	return CRP_OK;
}

int crp_field_create_periodic(const crp_field_t *field, const double extends[3], const double origin[3], int reflective,
                              crp_field_t **out) {
	if (!extends || !out)
		return fail(CRP_E_INVALID, "crp_field_create_periodic: NULL argument");
	crp_field *f = nullptr;
	if (int rc = new_composite(DECO_PERIODIC, out, &f))
		return rc;
	f->child[0] = field;
	f->nchild = 1;
	for (int a = 0; a < 3; a++) {
		f->extends[a] = extends[a];
		f->porigin[a] = origin ? origin[a] : 0.0;
	}
	f->preflective = reflective ? 1 : 0;
	if (int rc = composite_check(f)) {
		crp_field_destroy(f);
		*out = nullptr;
		return rc;
	}
	return CRP_OK;
}

int crp_field_create_list(crp_field_t **out) {
	if (!out)
		return fail(CRP_E_INVALID, "crp_field_create_list: NULL argument");
	crp_field *f = nullptr;
	return new_composite(DECO_LIST, out, &f);
}

int crp_field_list_add(crp_field_t *list, const crp_field_t *field) {
	if (!list || list->kind != FK_COMPOSITE || list->deco != DECO_LIST)
		return fail(CRP_E_INVALID, "crp_field_list_add: not a field list");
	if (list->nchild >= kDecoMaxChildren)
		return fail(CRP_E_INVALID, "crp_field_list_add: at most " + std::to_string(kDecoMaxChildren) +
		                               " fields per list on the device");
	list->child[list->nchild++] = field;
	if (int rc = composite_check(list)) {
		list->nchild--;
		return rc;
	}
	return CRP_OK;
}

int crp_field_create_evolution(const crp_field_t *field, double m, crp_field_t **out) {
	if (!out)
		return fail(CRP_E_INVALID, "crp_field_create_evolution: NULL argument");
	crp_field *f = nullptr;
	if (int rc = new_composite(DECO_EVOLUTION, out, &f))
		return rc;
	f->child[0] = field;
	f->nchild = 1;
	f->evoM = m;
	if (int rc = composite_check(f)) {
		crp_field_destroy(f);
		*out = nullptr;
		return rc;
	}
	return CRP_OK;
}

int crp_field_eval_z(const crp_field_t *f, const double *pos, const double *z, size_t n, double *B) {
	if (!pos || !z || !B)
		return fail(CRP_E_INVALID, "crp_field_eval_z: NULL argument");
	if (!(f && f->kind == FK_COMPOSITE && f->deco == DECO_EVOLUTION))
		return crp_field_eval(f, pos, n, B); // only a top-level evolution reads z
	if (n == 0)
		return CRP_OK;
	if (int rc = field_prepare(f))
		return rc;
	double *dpos = nullptr, *dB = nullptr, *dz = nullptr;
	hipError_t e = hipMalloc(&dpos, 3 * n * sizeof(double));
	if (e == hipSuccess)
		e = hipMalloc(&dB, 3 * n * sizeof(double));
	if (e == hipSuccess)
		e = hipMalloc(&dz, n * sizeof(double));
	if (e == hipSuccess)
		e = hipMemcpy(dpos, pos, 3 * n * sizeof(double), hipMemcpyHostToDevice);
	if (e == hipSuccess)
		e = hipMemcpy(dz, z, n * sizeof(double), hipMemcpyHostToDevice);
	if (e == hipSuccess) {
		hipLaunchKernelGGL(k_field_eval_z<FieldComposite>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0,
		                   FieldComposite{f->dcomp, 1.0}, dpos, dz, dB, n);
		e = hipGetLastError();
	}
	if (e == hipSuccess)
		e = hipMemcpy(B, dB, 3 * n * sizeof(double), hipMemcpyDeviceToHost);
	if (dpos)
		(void)hipFree(dpos);
	if (dB)
		(void)hipFree(dB);
	if (dz)
		(void)hipFree(dz);
	if (e != hipSuccess)
		return fail(CRP_E_HIP, std::string("crp_field_eval_z: ") + hipGetErrorString(e));
	return CRP_OK;
}

// ---------------------------------------------------------------- per-step functions (C ABI)
} // extern "C"

// Calls fn(field) with the FP64 device field of f, as crp_propagate dispatches it.
template <class Fn> static int with_fp64_field(const crp_field *f, Fn &&fn) {
	if (any_interp(f))
		return fn(grid_any(f));
	switch (f ? f->kind : FK_NONE) {
	case FK_UNIFORM: return fn(f->uniform);
	case FK_GRID: return f->cells ? fn(FieldGridCells{f->grid, f->cells}) : fn(FieldGrid{f->grid});
	case FK_PLANEWAVE: return fn(FieldPlaneWave{f->pw});
	case FK_MODULATED:
		return f->cells ? fn(FieldModulated<FieldGridCells>{FieldGridCells{f->grid, f->cells}, f->mod})
		                : fn(FieldModulated<FieldGrid>{FieldGrid{f->grid}, f->mod});
	case FK_JF12: return jf12_any(f) ? fn(FieldJF12Any{f->jf12}) : fn(f->jf12);
	case FK_COMPOSITE: return fn(FieldComposite{f->dcomp, 1.0});
	default: return fn(FieldNone{});
	}
}

static int step_api(const crp_propagator *p, int op, size_t n, const double *x, const double *u, const double *h,
                    const double *z, const double *q, const double *m, double *ox, double *ou, double *ex, double *eu,
                    const char *who) {
	const bool tryStep = op == SO_BP_TRY || op == SO_CK_TRY;
	if (!p || !x || !u || (op != SO_CK_DYDT && !h) || !q || !m || !ox || !ou || (tryStep && (!ex || !eu)))
		return fail(CRP_E_INVALID, std::string(who) + ": NULL argument");
	if ((op == SO_BP_DY || op == SO_BP_TRY) != (p->type == 0) || p->type == 2)
		return fail(CRP_E_INVALID, std::string(who) + ": wrong propagator type");
	if (n == 0)
		return CRP_OK;
	if (int rc = field_prepare(p->field))
		return rc;
	double *d = nullptr;
	CRP_HIP_TRY(hipMalloc(&d, 22 * n * sizeof(double)));
	StepArgs A;
	double *dx = d, *du = d + 3 * n, *dh = d + 6 * n, *dz = d + 7 * n, *dq = d + 8 * n, *dm = d + 9 * n;
	A.x = dx; A.u = du; A.h = dh; A.z = dz; A.q = dq; A.m = dm;
	A.ox = d + 10 * n; A.ou = d + 13 * n; A.ex = d + 16 * n; A.eu = d + 19 * n;
	A.n = n;
	std::vector<double> zeros(n, 0.0);
	hipError_t e = hipMemcpy(dx, x, 3 * n * sizeof(double), hipMemcpyHostToDevice);
	if (e == hipSuccess)
		e = hipMemcpy(du, u, 3 * n * sizeof(double), hipMemcpyHostToDevice);
	if (e == hipSuccess)
		e = hipMemcpy(dh, h ? h : zeros.data(), n * sizeof(double), hipMemcpyHostToDevice);
	if (e == hipSuccess)
		e = hipMemcpy(dz, z ? z : zeros.data(), n * sizeof(double), hipMemcpyHostToDevice);
	if (e == hipSuccess)
		e = hipMemcpy(dq, q, n * sizeof(double), hipMemcpyHostToDevice);
	if (e == hipSuccess)
		e = hipMemcpy(dm, m, n * sizeof(double), hipMemcpyHostToDevice);
	if (e == hipSuccess) {
		const unsigned blocks = (unsigned)((n + 255) / 256);
		with_fp64_field(p->field, [&](auto fld) {
			typedef decltype(fld) F;
			switch (op) {
			case SO_BP_DY: hipLaunchKernelGGL((k_step_api<F, SO_BP_DY>), blocks, 256, 0, 0, fld, A); break;
			case SO_BP_TRY: hipLaunchKernelGGL((k_step_api<F, SO_BP_TRY>), blocks, 256, 0, 0, fld, A); break;
			case SO_CK_DYDT: hipLaunchKernelGGL((k_step_api<F, SO_CK_DYDT>), blocks, 256, 0, 0, fld, A); break;
			default: hipLaunchKernelGGL((k_step_api<F, SO_CK_TRY>), blocks, 256, 0, 0, fld, A); break;
			}
			return 0;
		});
		e = hipGetLastError();
	}
	if (e == hipSuccess)
		e = hipMemcpy(ox, A.ox, 3 * n * sizeof(double), hipMemcpyDeviceToHost);
	if (e == hipSuccess)
		e = hipMemcpy(ou, A.ou, 3 * n * sizeof(double), hipMemcpyDeviceToHost);
	if (e == hipSuccess && tryStep)
		e = hipMemcpy(ex, A.ex, 3 * n * sizeof(double), hipMemcpyDeviceToHost);
	if (e == hipSuccess && tryStep)
		e = hipMemcpy(eu, A.eu, 3 * n * sizeof(double), hipMemcpyDeviceToHost);
	(void)hipFree(d);
	if (e != hipSuccess)
		return fail(CRP_E_HIP, std::string(who) + ": " + hipGetErrorString(e));
	return CRP_OK;
}

extern "C" {

int crp_bp_dy(const crp_propagator_t *p, size_t n, const double *pos, const double *dir, const double *step,
              const double *z, const double *q, const double *m, double *out_pos, double *out_dir) {
	return step_api(p, SO_BP_DY, n, pos, dir, step, z, q, m, out_pos, out_dir, nullptr, nullptr, "crp_bp_dy");
}

int crp_bp_try_step(const crp_propagator_t *p, size_t n, const double *pos, const double *dir, const double *h,
                    const double *z, const double *q, const double *m, double *out_pos, double *out_dir,
                    double *err_pos, double *err_dir) {
	return step_api(p, SO_BP_TRY, n, pos, dir, h, z, q, m, out_pos, out_dir, err_pos, err_dir, "crp_bp_try_step");
}

int crp_ck_dydt(const crp_propagator_t *p, size_t n, const double *x, const double *u, const double *z,
                const double *q, const double *E, double *dxdt, double *dudt) {
	return step_api(p, SO_CK_DYDT, n, x, u, nullptr, z, q, E, dxdt, dudt, nullptr, nullptr, "crp_ck_dydt");
}

int crp_ck_try_step(const crp_propagator_t *p, size_t n, const double *x, const double *u, const double *t,
                    const double *z, const double *q, const double *E, double *out_x, double *out_u, double *err_x,
                    double *err_u) {
	return step_api(p, SO_CK_TRY, n, x, u, t, z, q, E, out_x, out_u, err_x, err_u, "crp_ck_try_step");
}

int crp_field_eval(const crp_field_t *f, const double *pos, size_t n, double *B) {
	if (!pos || !B)
		return fail(CRP_E_INVALID, "crp_field_eval: NULL argument");
	if (n == 0)
		return CRP_OK;
	if (int rc = field_prepare(f))
		return rc;
	double *dpos = nullptr, *dB = nullptr;
	CRP_HIP_TRY(hipMalloc(&dpos, 3 * n * sizeof(double)));
	hipError_t e = hipMalloc(&dB, 3 * n * sizeof(double));
	if (e != hipSuccess) {
		(void)hipFree(dpos);
		return fail(CRP_E_NOMEM, "crp_field_eval: hipMalloc failed");
	}
	e = hipMemcpy(dpos, pos, 3 * n * sizeof(double), hipMemcpyHostToDevice);
	if (e == hipSuccess)
		e = launch_field_eval(f, dpos, dB, n, CRP_PRECISION_FP64);
	if (e == hipSuccess)
		e = hipMemcpy(B, dB, 3 * n * sizeof(double), hipMemcpyDeviceToHost);
	(void)hipFree(dpos);
	(void)hipFree(dB);
	if (e != hipSuccess)
		return fail(CRP_E_HIP, std::string("crp_field_eval: ") + hipGetErrorString(e));
	return CRP_OK;
}

static int validate_prop(const char *who, double tol, double minStep, double maxStep) {
	// setTolerance / setMaximumStep / setMinimumStep (PropagationBP.cpp:162-183, PropagationCK.cpp:160-179)
	if ((tol > 1) || (tol < 0))
		return fail(CRP_E_INVALID, std::string(who) + ": target error not in range 0-1");
	if (minStep < 0)
		return fail(CRP_E_INVALID, std::string(who) + ": minStep < 0 ");
	if (minStep > maxStep)
		return fail(CRP_E_INVALID, std::string(who) + ": minStep > maxStep");
	return CRP_OK;
}

static int create_prop(int type, const char *who, const crp_field_t *field, double tol, double minStep,
                       double maxStep, crp_propagator_t **out) {
	if (!out)
		return fail(CRP_E_INVALID, std::string(who) + ": out is NULL");
	int rc = validate_prop(who, tol, minStep, maxStep);
	if (rc != CRP_OK)
		return rc;
	crp_propagator *p = new (std::nothrow) crp_propagator;
	if (!p)
		return fail(CRP_E_NOMEM, std::string(who) + ": out of host memory");
	p->type = type;
	p->field = field;
	p->tolerance = tol;
	p->minStep = minStep;
	p->maxStep = maxStep;
	p->maxLength = 0;
	p->precision = CRP_PRECISION_FP64;
	p->block = 256;
	p->blocksPerCU = 0;
	p->stepsPerLaunch = 0;
	p->useGraph = 0;
	p->lastGrid = p->lastBlock = 0;
	p->lastPPL = 0;
	p->obsT = nullptr;
	p->nObs = 0;
	p->obsDeactivate = 0;
	p->box = 0;
	for (int a = 0; a < 3; a++)
		p->boxO[a] = p->boxS[a] = p->sphC[a] = 0;
	p->sph = p->sphLimit = p->postObs = 0;
	p->sphR = p->sphMargin = 0;
	p->sdeScale = 1;
	p->sdeAlpha = 1. / 3.;
	p->sdeEpsilon = 0.1;
	std::memset(p->sdeAdv, 0, sizeof(p->sdeAdv));
	p->sdeAdvection = 0;
	p->dSdeAdv = nullptr;
	p->sdeAdvDirty = 0;
	p->sdeSeed = 0;
	*out = p;
	return CRP_OK;
}

int crp_propagator_create_bp(const crp_field_t *field, double tolerance, double min_step, double max_step,
                             crp_propagator_t **out) {
	return create_prop(0, "PropagationBP", field, tolerance, min_step, max_step, out);
}

int crp_propagator_create_ck(const crp_field_t *field, double tolerance, double min_step, double max_step,
                             crp_propagator_t **out) {
	return create_prop(1, "PropagationCK", field, tolerance, min_step, max_step, out);
}

int crp_propagator_create_sde(const crp_field_t *field, double tolerance, double min_step, double max_step,
                              double epsilon, crp_propagator_t **out) {
	// DiffusionSDE's setters (src/module/DiffusionSDE.cpp:272-300)
	if (!out)
		return fail(CRP_E_INVALID, "DiffusionSDE: out is NULL");
	if (min_step < 0)
		return fail(CRP_E_INVALID, "DiffusionSDE: minStep < 0 ");
	if (min_step > max_step)
		return fail(CRP_E_INVALID, "DiffusionSDE: minStep > maxStep");
	if ((tolerance > 1) || (tolerance < 0))
		return fail(CRP_E_INVALID, "DiffusionSDE: tolerance error not in range 0-1");
	if ((epsilon > 1) || (epsilon < 0))
		return fail(CRP_E_INVALID, "DiffusionSDE: epsilon not in range 0-1");
	int rc = create_prop(2, "DiffusionSDE", field, tolerance, min_step, max_step, out);
	if (rc == CRP_OK)
		(*out)->sdeEpsilon = epsilon;
	return rc;
}

int crp_propagator_set_sde(crp_propagator_t *p, double alpha, double scale, double epsilon, uint32_t seed) {
	if (!p || p->type != 2)
		return fail(CRP_E_INVALID, "crp_propagator_set_sde: not a DiffusionSDE propagator");
	if ((alpha > 2.) || (alpha < 0))
		return fail(CRP_E_INVALID, "DiffusionSDE: alpha not in range 0-2");
	if (scale < 0)
		return fail(CRP_E_INVALID, "DiffusionSDE: Scale error: Scale < 0");
	if ((epsilon > 1) || (epsilon < 0))
		return fail(CRP_E_INVALID, "DiffusionSDE: epsilon not in range 0-1");
	p->sdeAlpha = alpha;
	p->sdeScale = scale;
	p->sdeEpsilon = epsilon;
	p->sdeSeed = seed;
	return CRP_OK;
}

// the number of parameters of each crp_sde_advection_kind (index = kind)
static const int kAdvParams[8] = {0, 3, 4, 7, 3, 5, 4, 8};

// Copies changed terms to the device before a propagation reads them (the setters stay host-only).
// Rare (the terms change between runs at most), so one lock for all propagators serves.
static std::mutex g_adv_mutex;

static int upload_sde_advection(const crp_propagator *p) {
	std::lock_guard<std::mutex> lock(g_adv_mutex);
	if (!p->sdeAdvDirty || p->sdeAdvection == 0)
		return CRP_OK;
	if (!p->dSdeAdv)
		CRP_HIP_TRY(hipMalloc((void **)&p->dSdeAdv, sizeof(p->sdeAdv)));
	CRP_HIP_TRY(hipDeviceSynchronize()); // no kernel may read the terms while they change
	CRP_HIP_TRY(hipMemcpy(p->dSdeAdv, p->sdeAdv, sizeof(p->sdeAdv), hipMemcpyHostToDevice));
	p->sdeAdvDirty = 0;
	return CRP_OK;
}

int crp_propagator_clear_sde_advection(crp_propagator_t *p) {
	if (!p || p->type != 2)
		return fail(CRP_E_INVALID, "crp_propagator_clear_sde_advection: not a DiffusionSDE propagator");
	std::memset(p->sdeAdv, 0, sizeof(p->sdeAdv));
	p->sdeAdvection = 0;
	p->sdeAdvDirty = 1;
	return CRP_OK;
}

int crp_propagator_add_sde_advection(crp_propagator_t *p, int kind, const double *par, int npar) {
	if (!p || p->type != 2)
		return fail(CRP_E_INVALID, "crp_propagator_add_sde_advection: not a DiffusionSDE propagator");
	if (kind < CRP_ADV_UNIFORM || kind > CRP_ADV_SPHERICAL_SHOCK)
		return fail(CRP_E_INVALID, "crp_propagator_add_sde_advection: unknown advection field kind");
	if (npar != kAdvParams[kind] || !par)
		return fail(CRP_E_INVALID, "crp_propagator_add_sde_advection: kind " + std::to_string(kind) + " takes " +
		                               std::to_string(kAdvParams[kind]) + " parameters");
	if (p->sdeAdvection >= kSdeAdvMax)
		return fail(CRP_E_INVALID, "crp_propagator_add_sde_advection: more than " + std::to_string(kSdeAdvMax) +
		                               " advection fields");
	SdeAdvTerm &T = p->sdeAdv[p->sdeAdvection];
	std::memset(&T, 0, sizeof(T));
	T.kind = kind;
	for (int i = 0; i < npar; i++)
		T.par[i] = par[i];
	p->sdeAdvection++;
	p->sdeAdvDirty = 1;
	return CRP_OK;
}

static_assert((int)CRP_SDE_ADVECTION_MAX == kSdeAdvMax, "advection term count");

int crp_propagator_set_sde_advection(crp_propagator_t *p, const double v[3]) {
	if (!p || p->type != 2)
		return fail(CRP_E_INVALID, "crp_propagator_set_sde_advection: not a DiffusionSDE propagator");
	if (int rc = crp_propagator_clear_sde_advection(p))
		return rc;
	return v ? crp_propagator_add_sde_advection(p, CRP_ADV_UNIFORM, v, 3) : CRP_OK;
}

int crp_propagator_set_sde_constant_spherical_advection(crp_propagator_t *p, const double origin[3], double v_wind) {
	if (!p || p->type != 2 || !origin)
		return fail(CRP_E_INVALID, "crp_propagator_set_sde_constant_spherical_advection: bad arguments");
	if (int rc = crp_propagator_clear_sde_advection(p))
		return rc;
	const double par[4] = {origin[0], origin[1], origin[2], v_wind};
	return crp_propagator_add_sde_advection(p, CRP_ADV_CONSTANT_SPHERICAL, par, 4);
}

int crp_propagator_set_sde_spherical_advection(crp_propagator_t *p, const double origin[3], double radius,
                                               double v_max, double tau, double alpha) {
	if (!p || p->type != 2 || !origin)
		return fail(CRP_E_INVALID, "crp_propagator_set_sde_spherical_advection: bad arguments");
	if (int rc = crp_propagator_clear_sde_advection(p))
		return rc;
	const double par[7] = {origin[0], origin[1], origin[2], radius, v_max, tau, alpha};
	return crp_propagator_add_sde_advection(p, CRP_ADV_SPHERICAL, par, 7);
}

int crp_propagator_set_max_trajectory_length(crp_propagator_t *p, double max_length) {
	if (!p)
		return fail(CRP_E_INVALID, "crp_propagator_set_max_trajectory_length: NULL propagator");
	p->maxLength = max_length > 0 ? max_length : 0;
	return CRP_OK;
}

int crp_propagator_set_precision(crp_propagator_t *p, int precision) {
	if (!p)
		return fail(CRP_E_INVALID, "crp_propagator_set_precision: NULL propagator");
	if (precision != CRP_PRECISION_FP64 && precision != CRP_PRECISION_MIXED)
		return fail(CRP_E_INVALID, "crp_propagator_set_precision: unknown precision");
	if (p->type == 2 && precision != CRP_PRECISION_FP64)
		return fail(CRP_E_INVALID, "crp_propagator_set_precision: DiffusionSDE runs in FP64 only");
	p->precision = precision;
	return CRP_OK;
}

int crp_propagator_set_observer(crp_propagator_t *p, const double *times, int ntimes, int deactivate) {
	if (!p)
		return fail(CRP_E_INVALID, "crp_propagator_set_observer: NULL propagator");
	if (ntimes < 0 || (ntimes > 0 && !times))
		return fail(CRP_E_INVALID, "crp_propagator_set_observer: bad time list");
	if (p->obsT) {
		CRP_HIP_TRY(hipDeviceSynchronize());
		CRP_HIP_TRY(hipFree(p->obsT));
		p->obsT = nullptr;
	}
	p->nObs = 0;
	p->obsDeactivate = deactivate ? 1 : 0;
	if (ntimes == 0)
		return CRP_OK;
	CRP_HIP_TRY(hipMalloc((void **)&p->obsT, (size_t)ntimes * sizeof(double)));
	CRP_HIP_TRY(hipMemcpy(p->obsT, times, (size_t)ntimes * sizeof(double), hipMemcpyHostToDevice));
	p->nObs = ntimes;
	return CRP_OK;
}

int crp_propagator_set_periodic_box(crp_propagator_t *p, const double origin[3], const double size[3]) {
	if (!p || !origin || !size)
		return fail(CRP_E_INVALID, "crp_propagator_set_periodic_box: NULL argument");
	const bool off = size[0] == 0 && size[1] == 0 && size[2] == 0;
	if (!off && !(size[0] > 0 && size[1] > 0 && size[2] > 0))
		return fail(CRP_E_INVALID, "crp_propagator_set_periodic_box: size must be > 0 (or all 0 to remove the box)");
	p->box = off ? 0 : 1;
	for (int a = 0; a < 3; a++) {
		p->boxO[a] = origin[a];
		p->boxS[a] = size[a];
	}
	return CRP_OK;
}

int crp_propagator_set_spherical_boundary(crp_propagator_t *p, const double center[3], double radius,
                                          double margin, int limit_step) {
	if (!p)
		return fail(CRP_E_INVALID, "crp_propagator_set_spherical_boundary: NULL propagator");
	if (!center) {
		p->sph = 0;
		return CRP_OK;
	}
	if (!(radius == radius) || !(margin == margin))
		return fail(CRP_E_INVALID, "crp_propagator_set_spherical_boundary: radius and margin must be numbers");
	p->sph = 1;
	p->sphLimit = limit_step ? 1 : 0;
	for (int a = 0; a < 3; a++)
		p->sphC[a] = center[a];
	p->sphR = radius;
	p->sphMargin = margin;
	return CRP_OK;
}

int crp_propagator_set_module_order(crp_propagator_t *p, const int *modules, int n) {
	if (!p || n < 0 || (n > 0 && !modules))
		return fail(CRP_E_INVALID, "crp_propagator_set_module_order: bad arguments");
	int seen = 0, obsAt = -1, sphAt = -1, lenAt = -1;
	for (int m = 0; m < n; m++) {
		const int id = modules[m];
		if (id < CRP_MODULE_PERIODIC_BOX || id > CRP_MODULE_OBSERVER || (seen & (1 << id)))
			return fail(CRP_E_INVALID, "crp_propagator_set_module_order: unknown or repeated module");
		if (id == CRP_MODULE_PERIODIC_BOX && m != 0)
			return fail(CRP_E_INVALID, "crp_propagator_set_module_order: the periodic box must come first");
		seen |= 1 << id;
		if (id == CRP_MODULE_OBSERVER) obsAt = m;
		if (id == CRP_MODULE_SPHERICAL_BOUNDARY) sphAt = m;
		if (id == CRP_MODULE_MAX_TRAJECTORY_LENGTH) lenAt = m;
	}
	p->postObs = (obsAt >= 0 && sphAt > obsAt ? 1 : 0) | (obsAt >= 0 && lenAt > obsAt ? 2 : 0);
	return CRP_OK;
}

int crp_propagator_destroy(crp_propagator_t *p) {
	if (p && (p->obsT || p->dSdeAdv)) {
		(void)hipDeviceSynchronize();
		if (p->obsT)
			(void)hipFree(p->obsT);
		if (p->dSdeAdv)
			(void)hipFree(p->dSdeAdv);
	}
	delete p;
	return CRP_OK;
}

int crp_batch_create(size_t capacity, crp_batch_t **out) {
	if (!out || capacity == 0)
		return fail(CRP_E_INVALID, "crp_batch_create: capacity must be > 0");
	crp_batch *b = new (std::nothrow) crp_batch;
	if (!b)
		return fail(CRP_E_NOMEM, "crp_batch_create: out of host memory");
	std::memset(b, 0, sizeof(*b));
	b->cap = capacity;
	(void)hipGetDevice(&b->device);
	// one allocation: 25 double arrays, the active flags, the queue order, then a
	// 256-byte aligned counter block
	const size_t nd = 25;
	size_t off = nd * capacity * sizeof(double) + 2 * capacity * sizeof(int32_t) + 2 * capacity * sizeof(uint32_t) +
	             capacity * sizeof(uint64_t);
	off = (off + 255) & ~(size_t)255;
	const size_t bytes = off + 64 * sizeof(unsigned long long);
	hipError_t e = hipMalloc((void **)&b->base, bytes);
	if (e != hipSuccess) {
		delete b;
		return fail(CRP_E_NOMEM, std::string("crp_batch_create: hipMalloc: ") + hipGetErrorString(e));
	}
	double **arr[25] = {&b->px, &b->py, &b->pz, &b->dx, &b->dy, &b->dz, &b->E, &b->q, &b->cur,
	                    &b->next, &b->len, &b->time, &b->ppx, &b->ppy, &b->ppz, &b->pdx, &b->pdy, &b->pdz,
	                    &b->sx, &b->sy, &b->sz, &b->cx, &b->cy, &b->cz, &b->z};
	for (size_t i = 0; i < nd; i++)
		*arr[i] = b->base + i * capacity;
	b->active = (int32_t *)(b->base + nd * capacity);
	b->order = (uint32_t *)(b->active + capacity);
	b->det = (int32_t *)(b->order + capacity);
	b->sdeStep = (uint32_t *)(b->det + capacity);
	b->sdeStream = (uint64_t *)(b->sdeStep + capacity); // 16 bytes per particle before it: 8-byte aligned
	b->ctr = (unsigned long long *)((char *)b->base + off);
	e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
	if (e == hipSuccess)
		e = hipEventCreate(&b->ev0);
	if (e == hipSuccess)
		e = hipEventCreate(&b->ev1);
	if (e != hipSuccess) {
		(void)hipFree(b->base);
		delete b;
		return fail(CRP_E_HIP, std::string("crp_batch_create: ") + hipGetErrorString(e));
	}
	*out = b;
	return CRP_OK;
}

int crp_batch_destroy(crp_batch_t *b) {
	if (!b)
		return CRP_OK;
	(void)hipStreamSynchronize(b->stream);
	if (b->gexec)
		(void)hipGraphExecDestroy(b->gexec);
	(void)hipEventDestroy(b->ev0);
	(void)hipEventDestroy(b->ev1);
	(void)hipStreamDestroy(b->stream);
	(void)hipFree(b->base);
	if (b->snap)
		(void)hipFree(b->snap);
	delete b;
	return CRP_OK;
}

int crp_batch_size(const crp_batch_t *b, size_t *n) {
	if (!b || !n)
		return fail(CRP_E_INVALID, "crp_batch_size: NULL argument");
	*n = b->n;
	return CRP_OK;
}

int crp_batch_upload(crp_batch_t *b, size_t n, const crp_host_state *s) {
	if (!b || !s)
		return fail(CRP_E_INVALID, "crp_batch_upload: NULL argument");
	if (n > b->cap)
		return fail(CRP_E_INVALID, "crp_batch_upload: n exceeds the batch capacity");
	if (!s->pos || !s->dir || !s->energy || !s->charge || !s->active)
		return fail(CRP_E_INVALID, "crp_batch_upload: pos, dir, energy, charge and active are required");
	b->n = n;
	if (n == 0)
		return CRP_OK;
	std::vector<double> tmp(n);
	double *dst3[6] = {b->px, b->py, b->pz, b->dx, b->dy, b->dz};
	for (int c = 0; c < 6; c++) {
		const double *src = c < 3 ? s->pos : s->dir;
		for (size_t i = 0; i < n; i++)
			tmp[i] = src[3 * i + (c % 3)];
		CRP_HIP_TRY(hipMemcpy(dst3[c], tmp.data(), n * sizeof(double), hipMemcpyHostToDevice));
	}
	const double *src1[7] = {s->energy, s->charge, s->current_step, s->next_step, s->trajectory_length, s->time,
	                         s->redshift};
	double *dst1[7] = {b->E, b->q, b->cur, b->next, b->len, b->time, b->z};
	for (int c = 0; c < 7; c++) {
		if (src1[c]) {
			CRP_HIP_TRY(hipMemcpy(dst1[c], src1[c], n * sizeof(double), hipMemcpyHostToDevice));
		} else {
			std::fill(tmp.begin(), tmp.end(), 0.0);
			CRP_HIP_TRY(hipMemcpy(dst1[c], tmp.data(), n * sizeof(double), hipMemcpyHostToDevice));
		}
	}
	CRP_HIP_TRY(hipMemcpy(b->active, s->active, n * sizeof(int32_t), hipMemcpyHostToDevice));
	{
		std::vector<int32_t> det(n, 0);
		if (s->detection_index)
			std::copy(s->detection_index, s->detection_index + n, det.begin());
		CRP_HIP_TRY(hipMemcpy(b->det, det.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
		std::vector<uint32_t> sst(n, 0);
		if (s->sde_step)
			std::copy(s->sde_step, s->sde_step + n, sst.begin());
		CRP_HIP_TRY(hipMemcpy(b->sdeStep, sst.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
		std::vector<uint64_t> sid(n);
		for (size_t i = 0; i < n; i++)
			sid[i] = s->sde_stream ? s->sde_stream[i] : (uint64_t)i;
		CRP_HIP_TRY(hipMemcpy(b->sdeStream, sid.data(), n * sizeof(uint64_t), hipMemcpyHostToDevice));
	}
	// Queue order: ascending magnetic rigidity E/|q| (neutral particles last).  Low-rigidity
	// particles take the most adaptive steps, so starting them first shortens the tail of a
	// run (longest-processing-time-first scheduling); results do not depend on the order.
	if (n > 0xffffffffull)
		return fail(CRP_E_INVALID, "crp_batch_upload: at most 2^32-1 particles per batch");
	{
		std::vector<uint32_t> ord(n);
		std::vector<double> key(n);
		for (size_t i = 0; i < n; i++) {
			ord[i] = (uint32_t)i;
			const double aq = std::fabs(s->charge[i]);
			key[i] = aq > 0 ? s->energy[i] / aq : HUGE_VAL;
		}
		std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t c) { return key[a] < key[c]; });
		CRP_HIP_TRY(hipMemcpy(b->order, ord.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
	}
	// previous := current until the first step
	CRP_HIP_TRY(hipMemcpy(b->ppx, b->px, 3 * b->cap * sizeof(double), hipMemcpyDeviceToDevice));
	CRP_HIP_TRY(hipMemcpy(b->pdx, b->dx, 3 * b->cap * sizeof(double), hipMemcpyDeviceToDevice));
	// source and created positions: given, or the current position
	const double *srcp[2] = {s->source_pos, s->created_pos};
	double *dsts[2][3] = {{b->sx, b->sy, b->sz}, {b->cx, b->cy, b->cz}};
	for (int v = 0; v < 2; v++)
		for (int c = 0; c < 3; c++) {
			if (srcp[v]) {
				for (size_t i = 0; i < n; i++)
					tmp[i] = srcp[v][3 * i + c];
				CRP_HIP_TRY(hipMemcpy(dsts[v][c], tmp.data(), n * sizeof(double), hipMemcpyHostToDevice));
			} else {
				CRP_HIP_TRY(hipMemcpy(dsts[v][c], dst3[c], n * sizeof(double), hipMemcpyDeviceToDevice));
			}
		}
	return CRP_OK;
}

int crp_batch_download(const crp_batch_t *b, crp_host_state *s) {
	if (!b || !s)
		return fail(CRP_E_INVALID, "crp_batch_download: NULL argument");
	const size_t n = b->n;
	if (n == 0)
		return CRP_OK;
	CRP_HIP_TRY(hipStreamSynchronize(b->stream));
	std::vector<double> tmp(n);
	const double *src3[12] = {b->px, b->py, b->pz, b->dx, b->dy, b->dz, b->ppx, b->ppy, b->ppz, b->pdx, b->pdy, b->pdz};
	double *dstv[4] = {s->pos, s->dir, s->prev_pos, s->prev_dir};
	for (int c = 0; c < 12; c++) {
		double *dst = dstv[c / 3];
		if (!dst)
			continue;
		CRP_HIP_TRY(hipMemcpy(tmp.data(), src3[c], n * sizeof(double), hipMemcpyDeviceToHost));
		for (size_t i = 0; i < n; i++)
			dst[3 * i + (c % 3)] = tmp[i];
	}
	const double *src1[6] = {b->E, b->q, b->cur, b->next, b->len, b->time};
	double *dst1[6] = {s->energy, s->charge, s->current_step, s->next_step, s->trajectory_length, s->time};
	for (int c = 0; c < 6; c++)
		if (dst1[c])
			CRP_HIP_TRY(hipMemcpy(dst1[c], src1[c], n * sizeof(double), hipMemcpyDeviceToHost));
	if (s->active)
		CRP_HIP_TRY(hipMemcpy(s->active, b->active, n * sizeof(int32_t), hipMemcpyDeviceToHost));
	if (s->detection_index)
		CRP_HIP_TRY(hipMemcpy(s->detection_index, b->det, n * sizeof(int32_t), hipMemcpyDeviceToHost));
	if (s->sde_step)
		CRP_HIP_TRY(hipMemcpy(s->sde_step, b->sdeStep, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
	const double *srcv[6] = {b->sx, b->sy, b->sz, b->cx, b->cy, b->cz};
	double *dstv2[2] = {s->source_pos, s->created_pos};
	for (int c = 0; c < 6; c++) {
		double *dst = dstv2[c / 3];
		if (!dst)
			continue;
		CRP_HIP_TRY(hipMemcpy(tmp.data(), srcv[c], n * sizeof(double), hipMemcpyDeviceToHost));
		for (size_t i = 0; i < n; i++)
			dst[3 * i + (c % 3)] = tmp[i];
	}
	return CRP_OK;
}

static int ensure_snapshots(crp_batch *b, int k) {
	if (b->snap && b->snapK >= k)
		return CRP_OK;
	double *ns = nullptr;
	CRP_HIP_TRY(hipMalloc((void **)&ns, (size_t)k * kSnapDoubles * b->cap * sizeof(double)));
	if (b->snap) {
		CRP_HIP_TRY(hipMemcpy(ns, b->snap, (size_t)b->snapK * kSnapDoubles * b->cap * sizeof(double),
		                      hipMemcpyDeviceToDevice));
		CRP_HIP_TRY(hipFree(b->snap));
	}
	b->snap = ns;
	b->snapK = k;
	return CRP_OK;
}

int crp_batch_detections(const crp_batch_t *b, int ntimes, double *out) {
	if (!b || (ntimes > 0 && !out) || ntimes < 0)
		return fail(CRP_E_INVALID, "crp_batch_detections: bad argument");
	const size_t n = b->n;
	if (n == 0 || ntimes == 0)
		return CRP_OK;
	CRP_HIP_TRY(hipStreamSynchronize(b->stream));
	std::vector<int32_t> det(n);
	CRP_HIP_TRY(hipMemcpy(det.data(), b->det, n * sizeof(int32_t), hipMemcpyDeviceToHost));
	std::fill(out, out + (size_t)ntimes * n * kSnapDoubles, NAN);
	const int kk = std::min(ntimes, b->snap ? b->snapK : 0);
	std::vector<double> col(n);
	for (int k = 0; k < kk; k++)
		for (int c = 0; c < kSnapDoubles; c++) {
			CRP_HIP_TRY(hipMemcpy(col.data(), b->snap + ((size_t)k * kSnapDoubles + c) * b->cap, n * sizeof(double),
			                      hipMemcpyDeviceToHost));
			for (size_t i = 0; i < n; i++)
				if (k < det[i])
					out[((size_t)k * n + i) * kSnapDoubles + c] = col[i];
		}
	return CRP_OK;
}

int crp_batch_copy(crp_batch_t *dst, const crp_batch_t *src) {
	if (!dst || !src)
		return fail(CRP_E_INVALID, "crp_batch_copy: NULL argument");
	if (src->n > dst->cap)
		return fail(CRP_E_INVALID, "crp_batch_copy: destination capacity too small");
	const size_t n = src->n;
	dst->n = n;
	if (n == 0)
		return CRP_OK;
	CRP_HIP_TRY(hipStreamSynchronize(src->stream));
	const double *s[25] = {src->px, src->py, src->pz, src->dx, src->dy, src->dz, src->E, src->q, src->cur,
	                       src->next, src->len, src->time, src->ppx, src->ppy, src->ppz, src->pdx, src->pdy, src->pdz,
	                       src->sx, src->sy, src->sz, src->cx, src->cy, src->cz, src->z};
	double *d[25] = {dst->px, dst->py, dst->pz, dst->dx, dst->dy, dst->dz, dst->E, dst->q, dst->cur,
	                 dst->next, dst->len, dst->time, dst->ppx, dst->ppy, dst->ppz, dst->pdx, dst->pdy, dst->pdz,
	                 dst->sx, dst->sy, dst->sz, dst->cx, dst->cy, dst->cz, dst->z};
	const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
	for (int c = 0; c < 25; c++)
		hipLaunchKernelGGL(k_copy_f64, dim3(blocks), dim3(256), 0, dst->stream, s[c], d[c], n);
	CRP_HIP_TRY(hipGetLastError());
	CRP_HIP_TRY(hipMemcpyAsync(dst->active, src->active, n * sizeof(int32_t), hipMemcpyDeviceToDevice, dst->stream));
	CRP_HIP_TRY(hipMemcpyAsync(dst->order, src->order, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, dst->stream));
	CRP_HIP_TRY(hipMemcpyAsync(dst->det, src->det, n * sizeof(int32_t), hipMemcpyDeviceToDevice, dst->stream));
	CRP_HIP_TRY(hipMemcpyAsync(dst->sdeStep, src->sdeStep, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, dst->stream));
	CRP_HIP_TRY(hipMemcpyAsync(dst->sdeStream, src->sdeStream, n * sizeof(uint64_t), hipMemcpyDeviceToDevice,
	                           dst->stream));
	if (src->snap) {
		int rc = ensure_snapshots(dst, src->snapK);
		if (rc != CRP_OK)
			return rc;
		for (int r = 0; r < src->snapK * kSnapDoubles; r++)
			CRP_HIP_TRY(hipMemcpyAsync(dst->snap + (size_t)r * dst->cap, src->snap + (size_t)r * src->cap,
			                           n * sizeof(double), hipMemcpyDeviceToDevice, dst->stream));
	}
	CRP_HIP_TRY(hipStreamSynchronize(dst->stream));
	return CRP_OK;
}

int crp_propagate(const crp_propagator_t *p, crp_batch_t *b, long long max_steps, crp_run_stats *stats) {
	if (!p || !b)
		return fail(CRP_E_INVALID, "crp_propagate: NULL argument");
	if (max_steps < 0 && !(p->maxLength > 0 && p->minStep > 0))
		return fail(CRP_E_UNBOUNDED, "crp_propagate: max_steps < 0 needs a maximum trajectory length and minStep > 0");
	if (p->type == 2)
		if (int rc = upload_sde_advection(p))
			return rc;
	crp_run_stats st;
	std::memset(&st, 0, sizeof(st));
	if (b->n == 0 || max_steps == 0) {
		if (stats)
			*stats = st;
		return CRP_OK;
	}
	Params P;
	std::memset(&P, 0, sizeof(P)); // padding too: the bytes key the captured graph
	P.tol = p->tolerance;
	P.minStep = p->minStep;
	P.maxStep = p->maxStep;
	P.maxLen = p->maxLength;
	P.maxSteps = max_steps;
	P.obsT = p->obsT;
	P.nObs = p->nObs;
	P.obsDeactivate = p->obsDeactivate;
	P.box = p->box;
	for (int a = 0; a < 3; a++) {
		P.boxO[a] = p->boxO[a];
		P.boxS[a] = p->boxS[a];
		P.boxInv[a] = p->box ? 1.0 / p->boxS[a] : 0.0;
		P.sphC[a] = p->sphC[a];
	}
	P.sph = p->sph;
	P.sphLimit = p->sphLimit;
	P.sphR = p->sphR;
	P.sphMargin = p->sphMargin;
	P.postObs = p->postObs;
	if (p->type == 2) {
		P.sdeScale = p->sdeScale;
		P.sdeAlpha = p->sdeAlpha;
		P.sdeEpsilon = p->sdeEpsilon;
		P.sdeAdvection = p->sdeAdvection;
		P.sdeAdvTerms = p->dSdeAdv;
		P.sdeSeed = p->sdeSeed;
		// CRP_SDE_MAX_HALVINGS (1..20) lowers the sub-step limit, so that tests can reach it on small inputs
		P.sdeMaxHalvings = kSdeMaxHalvings;
		if (const char *e = std::getenv("CRP_SDE_MAX_HALVINGS"))
			P.sdeMaxHalvings = std::max(1, std::min(kSdeMaxHalvings, std::atoi(e)));
	}
	if (p->nObs > 0) {
		int rc0 = ensure_snapshots(b, p->nObs);
		if (rc0 != CRP_OK)
			return rc0;
	}
	if (int rc0 = field_prepare(p->field)) // a decorator's program, before the timed region
		return rc0;
	CRP_HIP_TRY(hipEventRecord(b->ev0, b->stream));
	int rc;
	const crp_field *f = p->field;
	switch (any_interp(f) ? -1 : (f ? f->kind : FK_NONE)) {
	case -1: rc = dispatch_kind(grid_any(f), p, P, b, &st.launches); break;
	case FK_UNIFORM: rc = dispatch_kind(f->uniform, p, P, b, &st.launches); break;
	case FK_GRID:
		if (p->precision == CRP_PRECISION_MIXED && f->cells16)
			rc = dispatch_kind_p<FieldGridCells16, true>(FieldGridCells16{f->grid, f->cells16}, p, P, b, &st.launches);
		else
			rc = f->cells ? dispatch_kind(FieldGridCells{f->grid, f->cells}, p, P, b, &st.launches)
			              : dispatch_kind(FieldGrid{f->grid}, p, P, b, &st.launches);
		break;
	case FK_PLANEWAVE: rc = dispatch_kind(FieldPlaneWave{f->pw}, p, P, b, &st.launches); break;
	case FK_MODULATED:
		rc = f->cells ? dispatch_kind(FieldModulated<FieldGridCells>{FieldGridCells{f->grid, f->cells}, f->mod}, p, P, b,
		                              &st.launches)
		              : dispatch_kind(FieldModulated<FieldGrid>{FieldGrid{f->grid}, f->mod}, p, P, b, &st.launches);
		break;
	case FK_JF12:
		rc = jf12_any(f) ? dispatch_kind(FieldJF12Any{f->jf12}, p, P, b, &st.launches)
		                 : dispatch_kind(f->jf12, p, P, b, &st.launches);
		break;
	case FK_COMPOSITE: rc = dispatch_kind(FieldComposite{f->dcomp, 1.0}, p, P, b, &st.launches); break;
	default: rc = dispatch_kind(FieldNone{}, p, P, b, &st.launches); break;
	}
	if (rc != CRP_OK)
		return rc;
	CRP_HIP_TRY(hipEventRecord(b->ev1, b->stream));
	unsigned long long h[CTR_COUNT];
	CRP_HIP_TRY(hipMemcpyAsync(h, b->ctr, sizeof(h), hipMemcpyDeviceToHost, b->stream));
	CRP_HIP_TRY(hipStreamSynchronize(b->stream));
	float ms = 0;
	CRP_HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
	st.accepted_steps = h[CTR_ACCEPTED];
	st.rejected_trials = h[CTR_REJECTED];
	st.field_evals = h[CTR_EVALS];
	st.finished = h[CTR_FINISHED];
	st.kernel_ms = ms;
	if (stats)
		*stats = st;
	if (h[CTR_SDE_LIMIT] > 0)
		return fail(CRP_E_STATE, "DiffusionSDE: " + std::to_string(h[CTR_SDE_LIMIT]) +
		                             " particle(s) needed more than 2^" + std::to_string(P.sdeMaxHalvings) +
		                             " sub-steps in one step (minStep far below the field's scale) and were stopped");
	return CRP_OK;
}

int crp_propagator_set_launch(crp_propagator_t *p, const crp_launch_config *cfg) {
	if (!p || !cfg)
		return fail(CRP_E_INVALID, "crp_propagator_set_launch: NULL argument");
	if (cfg->block != 0 && cfg->block != 64 && cfg->block != 128 && cfg->block != 256)
		return fail(CRP_E_INVALID, "crp_propagator_set_launch: block must be 64, 128 or 256");
	if (cfg->blocks_per_cu < 0)
		return fail(CRP_E_INVALID, "crp_propagator_set_launch: blocks_per_cu must be >= 0");
	p->block = cfg->block ? cfg->block : 256;
	p->blocksPerCU = cfg->blocks_per_cu;
	p->stepsPerLaunch = cfg->steps_per_launch > 0 ? cfg->steps_per_launch : 0;
	p->useGraph = cfg->use_graph ? 1 : 0;
	return CRP_OK;
}

int crp_propagator_get_launch(const crp_propagator_t *p, crp_launch_config *cfg) {
	if (!p || !cfg)
		return fail(CRP_E_INVALID, "crp_propagator_get_launch: NULL argument");
	cfg->block = p->block;
	cfg->blocks_per_cu = p->blocksPerCU;
	cfg->steps_per_launch = p->stepsPerLaunch;
	cfg->use_graph = p->useGraph;
	cfg->particles_per_lane = p->lastPPL;
	return CRP_OK;
}

int crp_propagator_autotune(crp_propagator_t *p, const crp_batch_t *b, long long probe_steps,
                            crp_launch_config *chosen) {
	if (!p || !b)
		return fail(CRP_E_INVALID, "crp_propagator_autotune: NULL argument");
	if (probe_steps < 1)
		return fail(CRP_E_INVALID, "crp_propagator_autotune: probe_steps must be >= 1");
	crp_launch_config keep;
	crp_propagator_get_launch(p, &keep);
	if (b->n == 0) {
		if (chosen)
			*chosen = keep;
		return CRP_OK;
	}
	crp_batch *tmp = nullptr;
	int rc = crp_batch_create(b->n, &tmp);
	if (rc != CRP_OK)
		return rc;
	struct Cand {
		int block, bpc;
		long long spl;
		unsigned grid;
		double ms;
	};
	std::vector<Cand> tried;
	// time one configuration: best of two runs on fresh copies of b
	auto probe = [&](int block, int bpc, long long spl, Cand &out) -> int {
		p->block = block;
		p->blocksPerCU = bpc;
		p->stepsPerLaunch = spl;
		p->useGraph = 0;
		double best = 1e300;
		for (int r = 0; r < 2; r++) {
			int e = crp_batch_copy(tmp, b);
			if (e != CRP_OK)
				return e;
			crp_run_stats st;
			e = crp_propagate(p, tmp, probe_steps, &st);
			if (e != CRP_OK)
				return e;
			best = std::min(best, (double)st.kernel_ms);
		}
		out = Cand{block, bpc, spl, p->lastGrid, best};
		return CRP_OK;
	};
	const int bpcs[] = {1, 2, 3, 4, 6, 8, 12, 16};
	for (int block : {256, 128, 64}) {
		for (int bpc : bpcs) {
			Cand c;
			rc = probe(block, bpc, 0, c);
			if (rc != CRP_OK)
				break;
			bool dup = false;
			for (const Cand &t : tried)
				dup |= (t.block == c.block && t.grid == c.grid);
			if (dup)
				break; // the occupancy limit clamps blocks per CU: larger values launch the same grid
			tried.push_back(c);
			if ((size_t)c.grid * (size_t)block >= b->n)
				break; // every particle has a lane already
		}
		if (rc != CRP_OK)
			break;
	}
	if (rc == CRP_OK && !tried.empty()) {
		Cand best = tried[0];
		for (const Cand &t : tried)
			if (t.ms < best.ms)
				best = t;
		// steps per launch: split the probe into launches of a quarter and a half of it
		for (long long spl : {probe_steps / 4, probe_steps / 2}) {
			if (spl < 1)
				continue;
			Cand c;
			rc = probe(best.block, best.bpc, spl, c);
			if (rc != CRP_OK)
				break;
			tried.push_back(c);
			if (c.ms < best.ms)
				best = c;
		}
		if (rc == CRP_OK) {
			p->block = best.block;
			p->blocksPerCU = best.bpc;
			p->stepsPerLaunch = best.spl;
			p->useGraph = keep.use_graph;
			p->lastPPL = (double)b->n / ((double)best.grid * best.block);
			if (chosen)
				crp_propagator_get_launch(p, chosen);
		}
	}
	if (rc != CRP_OK) {
		p->block = keep.block;
		p->blocksPerCU = keep.blocks_per_cu;
		p->stepsPerLaunch = keep.steps_per_launch;
		p->useGraph = keep.use_graph;
	}
	crp_batch_destroy(tmp);
	return rc;
}

static int adj_args(const crp_adjoint_config *cfg, size_t n, AdjArgs &A) {
	if (!cfg || !cfg->grid)
		return fail(CRP_E_INVALID, "crp_adjoint: NULL config or grid");
	if (cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1 || cfg->nsteps < 0 || cfg->checkpoint_every < 1)
		return fail(CRP_E_INVALID, "crp_adjoint: bad grid size, step count or checkpoint interval");
	for (int a = 0; a < 3; a++)
		if (!(cfg->spacing[a] > 0))
			return fail(CRP_E_INVALID, "crp_adjoint: spacing must be > 0");
	std::memset(&A, 0, sizeof(A));
	A.G.g = cfg->grid;
	A.G.nx = cfg->nx;
	A.G.ny = cfg->ny;
	A.G.nz = cfg->nz;
	for (int a = 0; a < 3; a++) {
		A.G.gridOrigin[a] = cfg->origin[a] + cfg->spacing[a] / 2;
		A.G.invSpacing[a] = 1.0 / cfg->spacing[a];
		A.G.invN[a] = 1.0 / (double)(a == 0 ? cfg->nx : (a == 1 ? cfg->ny : cfg->nz));
	}
	A.G.scale = cfg->scale;
	A.h = cfg->step;
	A.nsteps = cfg->nsteps;
	A.K = cfg->checkpoint_every;
	A.n = n;
	return CRP_OK;
}

// Plan of the blocked backward (crp_adjoint.hpp): segments per block, deposit tiles and the scratch
// layout.  Every per-block buffer is doubled so that the reverse sweep of one block overlaps the
// recompute of the next and the accumulation of the previous one.  Deposit tiles are 8^3 cells, or 16^3
// when the grid has more than kAdjMaxTiles of those; grids with more than kAdjMaxTiles tiles of 16^3
// cells use the checkpointed backward with direct atomics.
constexpr int kAdjMaxTiles = 8192;
constexpr size_t kAdjBlockBudget = (size_t)4 << 30; // bytes of one buffer set (sub-checkpoints, records)
constexpr int kAdjMinBlocks = 8;

struct AdjPlan {
	bool blocked;
	int nseg, segPerBlock, nblocks, npg;
	int tileShift; // 3: 8^3-cell tiles, 4: 16^3
	int ntx, nty, ntz, ntiles;
	size_t maxRecords; // records of the largest block
	// byte offsets into the scratch
	size_t oCarry, oSub[2], oSRec[2], oCounts[2], oTileStart[2], oWgStart[2], total;
};

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

static AdjPlan adj_plan(const crp_adjoint_config *cfg, size_t n) {
	AdjPlan P;
	std::memset(&P, 0, sizeof(P));
	const int K = cfg->checkpoint_every;
	P.nseg = (cfg->nsteps + K - 1) / K;
	P.npg = (int)((n + kAdjGroup - 1) / kAdjGroup);
	// CRP_ADJOINT_TILE_SHIFT=4 asks for 16^3-cell tiles on grids where 8^3 would do (tests)
	const char *ts = std::getenv("CRP_ADJOINT_TILE_SHIFT");
	long long tiles = 0;
	for (int shift = (ts && std::atoi(ts) == 4) ? 4 : 3; shift <= 4; shift++) {
		const int c = 1 << shift;
		P.tileShift = shift;
		P.ntx = (cfg->nx + c - 1) / c;
		P.nty = (cfg->ny + c - 1) / c;
		P.ntz = (cfg->nz + c - 1) / c;
		tiles = (long long)P.ntx * P.nty * P.ntz;
		if (tiles <= kAdjMaxTiles)
			break;
	}
	// per segment and buffer set: the sub-checkpoints (48 B per kAdjSub steps) and the records (32 B a step)
	const double perSeg = (double)K * (double)n * (48.0 / kAdjSub + 32.0);
	// CRP_ADJOINT_BLOCK_BYTES overrides the buffer-set budget (0: per-step atomics), so that tests
	// can run the multi-block and the atomic paths on small inputs
	double budget = (double)kAdjBlockBudget;
	if (const char *e = std::getenv("CRP_ADJOINT_BLOCK_BYTES"))
		budget = std::strtod(e, nullptr);
	P.blocked = budget > 0 && cfg->nsteps > 0 && n > 0 && tiles <= kAdjMaxTiles && (double)K * n < 2e9;
	if (!P.blocked) {
		P.total = (size_t)K * 6 * n * sizeof(double);
		return P;
	}
	P.ntiles = (int)tiles;
	// at least kAdjMinBlocks blocks where there are that many segments, so that the three streams
	// overlap over most of the backward
	const double cap = std::ceil((double)P.nseg / kAdjMinBlocks);
	P.segPerBlock = (int)std::max(1.0, std::min(cap, std::floor(budget / perSeg)));
	while ((double)P.segPerBlock * K * n >= 2e9 && P.segPerBlock > 1)
		P.segPerBlock--;
	P.nblocks = (P.nseg + P.segPerBlock - 1) / P.segPerBlock;
	P.segPerBlock = (P.nseg + P.nblocks - 1) / P.nblocks; // the same number of segments in every block
	P.maxRecords = (size_t)P.segPerBlock * K * n;
	size_t o = 0;
	P.oCarry = o;      o += align256(6 * n * sizeof(double));
	for (int k = 0; k < 2; k++) {
		P.oSub[k] = o;        o += align256(((size_t)P.segPerBlock * K + kAdjSub - 1) / kAdjSub * 6 * n * sizeof(double));
		P.oSRec[k] = o;       o += align256(P.maxRecords * sizeof(DepRec));
		P.oCounts[k] = o;     o += align256((size_t)P.npg * P.ntiles * sizeof(uint32_t));
		P.oTileStart[k] = o;  o += align256(((size_t)P.ntiles + 1) * sizeof(uint32_t));
		P.oWgStart[k] = o;    o += align256(((size_t)P.ntiles + 1) * sizeof(uint32_t));
	}
	P.total = o;
	return P;
}

// the two side streams of the blocked backward, per device, default priority
static std::mutex g_adj_stream_mutex;
static hipStream_t g_adj_stream[64][2];

static hipError_t adj_side_streams(hipStream_t *b, hipStream_t *c) {
	int dev = 0;
	hipError_t e = hipGetDevice(&dev);
	if (e != hipSuccess)
		return e;
	if (dev < 0 || dev >= 64)
		return hipErrorInvalidDevice;
	std::lock_guard<std::mutex> lock(g_adj_stream_mutex);
	for (int k = 0; k < 2; k++) {
		if (!g_adj_stream[dev][k]) {
			e = hipStreamCreateWithFlags(&g_adj_stream[dev][k], hipStreamNonBlocking);
			if (e != hipSuccess)
				return e;
		}
	}
	*b = g_adj_stream[dev][0];
	*c = g_adj_stream[dev][1];
	return hipSuccess;
}

// The forward keeps a cell table of the grid (k_build_cells: the 8 corner values of every cell in one
// 128-byte line) after the checkpoints, for itself and the backward's recompute, when it takes at most
// kAdjCellsMax bytes; larger grids are gathered point by point.
constexpr size_t kAdjCellsMax = (size_t)8 << 30;

static size_t adj_cells_bytes(const crp_adjoint_config *cfg) {
	const size_t b = (size_t)cfg->nx * cfg->ny * cfg->nz * kCellFloats * sizeof(float);
	return b <= kAdjCellsMax ? b : 0;
}

static size_t adj_checkpoint_bytes(const crp_adjoint_config *cfg, size_t n) {
	const size_t nseg = ((size_t)cfg->nsteps + cfg->checkpoint_every - 1) / cfg->checkpoint_every;
	return align256((nseg > 0 ? nseg : 1) * 6 * n * sizeof(double));
}

int crp_adjoint_workspace(const crp_adjoint_config *cfg, size_t n, size_t *checkpoint_bytes, size_t *scratch_bytes) {
	AdjArgs A;
	int rc = adj_args(cfg, n, A);
	if (rc != CRP_OK)
		return rc;
	if (checkpoint_bytes)
		*checkpoint_bytes = adj_checkpoint_bytes(cfg, n) + adj_cells_bytes(cfg);
	if (scratch_bytes)
		*scratch_bytes = std::max<size_t>(adj_plan(cfg, n).total, 8);
	return CRP_OK;
}

int crp_adjoint_forward(const crp_adjoint_config *cfg, size_t n, const double *x0, const double *u0,
                        const double *kq, double *checkpoints, double *x_out, double *u_out, void *stream) {
	AdjArgs A;
	int rc = adj_args(cfg, n, A);
	if (rc != CRP_OK)
		return rc;
	if (n == 0)
		return CRP_OK;
	if (!x0 || !u0 || !kq || !checkpoints || !x_out || !u_out)
		return fail(CRP_E_INVALID, "crp_adjoint_forward: NULL pointer");
	A.x0 = x0;
	A.u0 = u0;
	A.kq = kq;
	A.ckpt = checkpoints;
	A.xout = x_out;
	A.uout = u_out;
	if (adj_cells_bytes(cfg)) {
		float *cells = reinterpret_cast<float *>(reinterpret_cast<char *>(checkpoints) + adj_checkpoint_bytes(cfg, n));
		const size_t ncell = (size_t)cfg->nx * cfg->ny * cfg->nz;
		const unsigned blocks = (unsigned)std::min<size_t>((ncell + 255) / 256, 65535);
		hipLaunchKernelGGL(k_build_cells, dim3(blocks), dim3(256), 0, (hipStream_t)stream, cfg->grid, cells, cfg->nx,
		                   cfg->ny, cfg->nz);
		A.G.cells = cells;
	}
	hipLaunchKernelGGL(k_adjoint_forward, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, A);
	CRP_HIP_TRY(hipGetLastError());
	return CRP_OK;
}

// events of one blocked backward: released (after the work they mark has completed) when it returns
struct AdjEvents {
	std::vector<hipEvent_t> ev;
	~AdjEvents() {
		for (hipEvent_t e : ev)
			(void)hipEventDestroy(e);
	}
	hipError_t make(hipEvent_t *out) {
		hipError_t e = hipEventCreateWithFlags(out, hipEventDisableTiming);
		if (e == hipSuccess)
			ev.push_back(*out);
		return e;
	}
};

int crp_adjoint_backward(const crp_adjoint_config *cfg, size_t n, const double *kq, const double *checkpoints,
                         const double *grad_x_out, const double *grad_u_out, double *segment_scratch, double *grad_grid,
                         double *grad_x0, double *grad_u0, double *grad_scale, void *stream) {
	AdjArgs A;
	int rc = adj_args(cfg, n, A);
	if (rc != CRP_OK)
		return rc;
	if (n == 0)
		return CRP_OK;
	if (!kq || !checkpoints || !grad_u_out || !segment_scratch || !grad_grid || !grad_x0 || !grad_u0 || !grad_scale)
		return fail(CRP_E_INVALID, "crp_adjoint_backward: NULL pointer");
	A.kq = kq;
	A.ckpt = const_cast<double *>(checkpoints);
	A.gxout = grad_x_out;
	A.guout = grad_u_out;
	A.scratch = segment_scratch;
	A.gradG = grad_grid;
	A.gradX0 = grad_x0;
	A.gradU0 = grad_u0;
	A.gradScale = grad_scale;
	if (adj_cells_bytes(cfg))
		A.G.cells = reinterpret_cast<const float *>(reinterpret_cast<const char *>(checkpoints) + adj_checkpoint_bytes(cfg, n));
	hipStream_t sa = (hipStream_t)stream;
	const AdjPlan P = adj_plan(cfg, n);
	if (!P.blocked) {
		hipLaunchKernelGGL(k_adjoint_backward, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, sa, A);
		CRP_HIP_TRY(hipGetLastError());
		return CRP_OK;
	}
	// stream A (the caller's) runs the reverse sweeps, block after block; stream B the recompute of
	// the block after (in backward order), stream C the accumulation of the block before: the three
	// are bound by FP64 issue at one wave per SIMD, the random-line rate of the cell gathers and LDS
	// atomics respectively
	hipStream_t sb, sc;
	CRP_HIP_TRY(adj_side_streams(&sb, &sc));
	AdjEvents E;
	const int NB = P.nblocks;
	std::vector<hipEvent_t> recomputed(NB), reversed(NB), sorted(NB);
	for (int b = 0; b < NB; b++) {
		CRP_HIP_TRY(E.make(&recomputed[b]));
		CRP_HIP_TRY(E.make(&reversed[b]));
		CRP_HIP_TRY(E.make(&sorted[b]));
	}
	hipEvent_t start, done;
	CRP_HIP_TRY(E.make(&start));
	CRP_HIP_TRY(E.make(&done));

	char *base = reinterpret_cast<char *>(segment_scratch);
	const int K = cfg->checkpoint_every;
	auto block = [&](int b) {
		AdjBlock Bk;
		std::memset(&Bk, 0, sizeof(Bk));
		const int buf = b & 1;
		Bk.carry = reinterpret_cast<double *>(base + P.oCarry);
		Bk.sub = reinterpret_cast<double *>(base + P.oSub[buf]);
		Bk.srec = reinterpret_cast<DepRec *>(base + P.oSRec[buf]);
		Bk.counts = reinterpret_cast<uint32_t *>(base + P.oCounts[buf]);
		Bk.tileStart = reinterpret_cast<uint32_t *>(base + P.oTileStart[buf]);
		Bk.wgStart = reinterpret_cast<uint32_t *>(base + P.oWgStart[buf]);
		Bk.npg = P.npg;
		Bk.ntx = P.ntx;
		Bk.nty = P.nty;
		Bk.ntz = P.ntz;
		Bk.ntiles = P.ntiles;
		Bk.seg0 = b * P.segPerBlock;
		Bk.nseg = std::min(P.segPerBlock, P.nseg - Bk.seg0);
		Bk.s0 = Bk.seg0 * K;
		Bk.s1 = std::min(cfg->nsteps, (Bk.seg0 + Bk.nseg) * K);
		Bk.first = b == NB - 1;
		Bk.last = b == 0;
		Bk.m = (size_t)(Bk.s1 - Bk.s0) * n;
		return Bk;
	};
	const size_t T = (size_t)P.ntiles;
	// recompute of block b: its sub-checkpoints, the records' counts per particle group and tile, and
	// the scans
	auto recompute = [&](int b, hipStream_t st) {
		const AdjBlock Bk = block(b);
		const size_t nc = (size_t)Bk.npg * T;
		hipLaunchKernelGGL(k_dep_zero, dim3((unsigned)((nc + kAdjGroup - 1) / kAdjGroup)), dim3(kAdjGroup), 0, st, Bk);
		const dim3 grid((unsigned)((size_t)Bk.nseg * Bk.npg));
		if (P.tileShift == 3)
			hipLaunchKernelGGL((k_adj_recompute<3, kAdjSub>), grid, dim3(kAdjGroup), T * sizeof(uint32_t), st, A, Bk);
		else
			hipLaunchKernelGGL((k_adj_recompute<4, kAdjSub>), grid, dim3(kAdjGroup), T * sizeof(uint32_t), st, A, Bk);
		hipLaunchKernelGGL(k_dep_scan_chunks, dim3((unsigned)((T + 255) / 256)), dim3(256), 0, st, Bk);
		hipLaunchKernelGGL(k_dep_scan_tiles, dim3(1), dim3(1024), 0, st, Bk);
	};
	auto reverse = [&](int b, hipStream_t st) {
		const AdjBlock Bk = block(b);
		const dim3 g((unsigned)Bk.npg), w(kAdjGroup);
		const size_t lb = 2 * T * sizeof(uint32_t);
		if (P.tileShift == 3)
			hipLaunchKernelGGL((k_adj_reverse<3, kAdjSub>), g, w, lb, st, A, Bk);
		else
			hipLaunchKernelGGL((k_adj_reverse<4, kAdjSub>), g, w, lb, st, A, Bk);
	};
	auto accumulate = [&](int b, hipStream_t st) {
		const AdjBlock Bk = block(b);
		const size_t groups = T + (Bk.m + kDepPerGroup - 1) / kDepPerGroup;
		if (P.tileShift == 3)
			hipLaunchKernelGGL(k_dep_tiles<3>, dim3((unsigned)groups), dim3(Tile<3>::threads), 0, st, A, Bk);
		else
			hipLaunchKernelGGL(k_dep_tiles<4>, dim3((unsigned)groups), dim3(Tile<4>::threads), 0, st, A, Bk);
	};

	// streams B and C start after the work already queued on stream A (e.g. the zeroed gradients)
	CRP_HIP_TRY(hipEventRecord(start, sa));
	CRP_HIP_TRY(hipStreamWaitEvent(sb, start, 0));
	CRP_HIP_TRY(hipStreamWaitEvent(sc, start, 0));
	recompute(NB - 1, sb);
	CRP_HIP_TRY(hipEventRecord(recomputed[NB - 1], sb));
	for (int b = NB - 1; b >= 0; b--) {
		// A: reverse sweep of block b once its sub-checkpoints and counts are there and the accumulation
		// of block b + 2 (the same record buffer) has read its records
		CRP_HIP_TRY(hipStreamWaitEvent(sa, recomputed[b], 0));
		if (b + 2 < NB)
			CRP_HIP_TRY(hipStreamWaitEvent(sa, sorted[b + 2], 0));
		reverse(b, sa);
		CRP_HIP_TRY(hipEventRecord(reversed[b], sa));
		// B: recompute of block b - 1 into the buffers that the reverse sweep of block b + 1 has
		// finished with (its counts were last read by the accumulation of block b + 1)
		if (b > 0) {
			if (b + 1 < NB) {
				CRP_HIP_TRY(hipStreamWaitEvent(sb, reversed[b + 1], 0));
				CRP_HIP_TRY(hipStreamWaitEvent(sb, sorted[b + 1], 0));
			}
			recompute(b - 1, sb);
			CRP_HIP_TRY(hipEventRecord(recomputed[b - 1], sb));
		}
		// C: accumulation of block b
		CRP_HIP_TRY(hipStreamWaitEvent(sc, reversed[b], 0));
		accumulate(b, sc);
		CRP_HIP_TRY(hipEventRecord(sorted[b], sc));
	}
	CRP_HIP_TRY(hipGetLastError());
	// stream A continues once the last accumulation is in the gradient
	CRP_HIP_TRY(hipEventRecord(done, sc));
	CRP_HIP_TRY(hipStreamWaitEvent(sa, done, 0));
	return CRP_OK;
}

// ------------------------------------------------------------------ turbulent grids
// hipFFT is loaded when first needed, so the propagator does not depend on it.
struct HipfftApi {
	decltype(&hipfftPlan3d) plan3d = nullptr;
	decltype(&hipfftExecC2R) execC2R = nullptr;
	decltype(&hipfftDestroy) destroy = nullptr;
	bool ok = false;
};

static const HipfftApi &hipfft_api() {
	static HipfftApi api = [] {
		HipfftApi a;
		void *h = dlopen("libhipfft.so.0", RTLD_NOW | RTLD_GLOBAL);
		if (!h)
			h = dlopen("libhipfft.so", RTLD_NOW | RTLD_GLOBAL);
		if (!h)
			return a;
		a.plan3d = reinterpret_cast<decltype(&hipfftPlan3d)>(dlsym(h, "hipfftPlan3d"));
		a.execC2R = reinterpret_cast<decltype(&hipfftExecC2R)>(dlsym(h, "hipfftExecC2R"));
		a.destroy = reinterpret_cast<decltype(&hipfftDestroy)>(dlsym(h, "hipfftDestroy"));
		a.ok = a.plan3d && a.execC2R && a.destroy;
		return a;
	}();
	return api;
}

// GridTurbulence::checkGridRequirements (GridTurbulence.cpp:126-143) for a cubic grid of equal spacing
static int turb_desc(const char *fn, const crp_turbulence_desc *d, TurbDesc &D) {
	if (!d)
		return fail(CRP_E_INVALID, std::string(fn) + ": NULL descriptor");
	if (d->n < 2 || d->n > 2048 || !(d->spacing > 0))
		return fail(CRP_E_INVALID, std::string(fn) + ": need 2 <= n <= 2048 and spacing > 0");
	if (d->lmin < 2 * d->spacing)
		return fail(CRP_E_INVALID, "turbulentField: lMin < 2 * spacing");
	if (d->lmax > d->n * d->spacing)
		return fail(CRP_E_INVALID, "turbulentField: lMax > size");
	if (d->lmax < d->lmin)
		return fail(CRP_E_INVALID, "lMax < lMin");
	if (d->kind != CRP_TURBULENCE_SIMPLE && d->kind != CRP_TURBULENCE_GENERAL &&
	    d->kind != CRP_TURBULENCE_GENERAL_POWERLAW)
		return fail(CRP_E_INVALID, std::string(fn) + ": unknown spectrum kind");
	D.n = d->n;
	D.n2 = d->n / 2 + 1;
	D.kind = d->kind;
	D.kMin = d->spacing / d->lmax;
	D.kMax = d->spacing / d->lmin;
	D.alpha = d->alpha;
	D.lambda = 1 / d->spacing * 2 * M_PI;
	D.lbo = d->lbendover;
	D.sidx = d->sindex;
	D.qidx = d->qindex;
	return CRP_OK;
}

// per-row inside-mode counts; rowOffset gets their exclusive prefix sums, total the sum
static int turb_rows(const TurbDesc &D, std::vector<uint64_t> &rowOffset, uint64_t &total) {
	const int rows = D.n * D.n;
	uint32_t *dCount = nullptr;
	CRP_HIP_TRY(hipMalloc((void **)&dCount, rows * sizeof(uint32_t)));
	hipLaunchKernelGGL(k_turb_row_count, dim3((rows + 255) / 256), dim3(256), 0, 0, D, dCount);
	std::vector<uint32_t> count(rows);
	hipError_t e = hipGetLastError();
	if (e == hipSuccess)
		e = hipMemcpy(count.data(), dCount, rows * sizeof(uint32_t), hipMemcpyDeviceToHost);
	(void)hipFree(dCount);
	if (e != hipSuccess)
		return fail(CRP_E_HIP, std::string("turbulence row count: ") + hipGetErrorString(e));
	rowOffset.resize(rows);
	total = 0;
	for (int r = 0; r < rows; r++) {
		rowOffset[r] = total;
		total += count[r];
	}
	return CRP_OK;
}

int crp_turbulence_mode_count(const crp_turbulence_desc *d, uint64_t *modes) {
	TurbDesc D;
	int rc = turb_desc("crp_turbulence_mode_count", d, D);
	if (rc != CRP_OK)
		return rc;
	if (!modes)
		return fail(CRP_E_INVALID, "crp_turbulence_mode_count: NULL output");
	std::vector<uint64_t> off;
	uint64_t total = 0;
	rc = turb_rows(D, off, total);
	if (rc == CRP_OK)
		*modes = total;
	return rc;
}

int crp_turbulence_generate(const crp_turbulence_desc *d, const uint32_t *draws, uint64_t ndraws, float *grid) {
	TurbDesc D;
	int rc = turb_desc("crp_turbulence_generate", d, D);
	if (rc != CRP_OK)
		return rc;
	if (!grid || (ndraws > 0 && !draws))
		return fail(CRP_E_INVALID, "crp_turbulence_generate: NULL argument");
	std::vector<uint64_t> off;
	uint64_t total = 0;
	rc = turb_rows(D, off, total);
	if (rc != CRP_OK)
		return rc;
	const uint64_t perMode = D.kind == TURB_SIMPLE ? 4 : 2;
	if (ndraws != perMode * total)
		return fail(CRP_E_INVALID, "crp_turbulence_generate: " + std::to_string(total) + " modes need " +
		                               std::to_string(perMode * total) + " draws, got " + std::to_string(ndraws));
	const HipfftApi &fft = hipfft_api();
	if (!fft.ok)
		return fail(CRP_E_STATE, "crp_turbulence_generate: hipFFT (libhipfft.so) could not be loaded");
	const int rows = D.n * D.n;
	const size_t plane = (size_t)rows * D.n2; // complex points per component
	const size_t npts = (size_t)rows * D.n;
	uint64_t *dOff = nullptr;
	uint32_t *dDraws = nullptr;
	float2 *B = nullptr;
	float *dGrid = nullptr;
	hipfftHandle plan = nullptr;
	hipError_t e = hipSuccess;
	auto cleanup = [&]() {
		if (plan)
			(void)fft.destroy(plan);
		(void)hipFree(dOff);
		(void)hipFree(dDraws);
		(void)hipFree(B);
		(void)hipFree(dGrid);
	};
#define TURB_TRY(call)                                                                                         \
	do {                                                                                                       \
		e = (call);                                                                                            \
		if (e != hipSuccess) {                                                                                 \
			cleanup();                                                                                         \
			return fail(e == hipErrorOutOfMemory ? CRP_E_NOMEM : CRP_E_HIP,                                    \
			            std::string("crp_turbulence_generate: ") + hipGetErrorString(e));                      \
		}                                                                                                      \
	} while (0)
	TURB_TRY(hipMalloc((void **)&dOff, rows * sizeof(uint64_t)));
	TURB_TRY(hipMemcpy(dOff, off.data(), rows * sizeof(uint64_t), hipMemcpyHostToDevice));
	TURB_TRY(hipMalloc((void **)&dDraws, std::max<uint64_t>(ndraws, 1) * sizeof(uint32_t)));
	if (ndraws)
		TURB_TRY(hipMemcpy(dDraws, draws, ndraws * sizeof(uint32_t), hipMemcpyHostToDevice));
	TURB_TRY(hipMalloc((void **)&B, 3 * plane * sizeof(float2)));
	hipLaunchKernelGGL(k_turb_coeffs, dim3((rows + 255) / 256), dim3(256), 0, 0, D, dOff, dDraws, B);
	TURB_TRY(hipGetLastError());
	hipLaunchKernelGGL(k_turb_hermitize, dim3((rows + 255) / 256), dim3(256), 0, 0, B, D.n, D.n2, 0);
	TURB_TRY(hipGetLastError());
	if (D.n % 2 == 0) {
		hipLaunchKernelGGL(k_turb_hermitize, dim3((rows + 255) / 256), dim3(256), 0, 0, B, D.n, D.n2, D.n / 2);
		TURB_TRY(hipGetLastError());
	}
	TURB_TRY(hipFree(dDraws));
	dDraws = nullptr;
	// in place, complex to real, unnormalised, per component (GridTurbulence.cpp:155-171)
	if (fft.plan3d(&plan, D.n, D.n, D.n, HIPFFT_C2R) != HIPFFT_SUCCESS) {
		plan = nullptr;
		cleanup();
		return fail(CRP_E_HIP, "crp_turbulence_generate: hipfftPlan3d failed");
	}
	for (int c = 0; c < 3; c++) {
		hipfftComplex *in = reinterpret_cast<hipfftComplex *>(B + c * plane);
		if (fft.execC2R(plan, in, reinterpret_cast<hipfftReal *>(in)) != HIPFFT_SUCCESS) {
			cleanup();
			return fail(CRP_E_HIP, "crp_turbulence_generate: hipfftExecC2R failed");
		}
	}
	TURB_TRY(hipDeviceSynchronize());
	TURB_TRY(hipMalloc((void **)&dGrid, 3 * npts * sizeof(float)));
	hipLaunchKernelGGL(k_turb_pack, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, 0,
	                   reinterpret_cast<const float *>(B), D.n, D.n2, dGrid);
	TURB_TRY(hipGetLastError());
	TURB_TRY(hipMemcpy(grid, dGrid, 3 * npts * sizeof(float), hipMemcpyDefault));
#undef TURB_TRY
	cleanup();
	return CRP_OK;
}

// ------------------------------------------------------------------ random-gather ceiling
int crp_measure_gather(uint64_t table_bytes, int bytes_per_line, double *lines_per_s) {
	if (!lines_per_s || (bytes_per_line != 64 && bytes_per_line != 96 && bytes_per_line != 128) ||
	    table_bytes < 128 * 1024 || table_bytes / 128 > 0xffffffffull)
		return fail(CRP_E_INVALID, "crp_measure_gather: bytes_per_line 64/96/128, table 128 KiB .. 512 GiB");
	int dev = 0, ncu = 0;
	CRP_HIP_TRY(hipGetDevice(&dev));
	CRP_HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
	const unsigned nlines = (unsigned)(table_bytes / 128);
	const unsigned blocks = (unsigned)ncu * 8, threads = blocks * 256;
	const int iters = 256;
	float4 *tab = nullptr;
	float *out = nullptr;
	hipEvent_t a = nullptr, b = nullptr;
	auto cleanup = [&]() {
		(void)hipFree(tab);
		(void)hipFree(out);
		if (a) (void)hipEventDestroy(a);
		if (b) (void)hipEventDestroy(b);
	};
	hipError_t e = hipMalloc((void **)&tab, (size_t)nlines * 128);
	if (e == hipSuccess) e = hipMalloc((void **)&out, threads * sizeof(float));
	if (e == hipSuccess) e = hipMemset(tab, 0, (size_t)nlines * 128);
	if (e == hipSuccess) e = hipEventCreate(&a);
	if (e == hipSuccess) e = hipEventCreate(&b);
	float ms = 0.f;
	for (int rep = 0; rep < 2 && e == hipSuccess; rep++) { // first pass warms up
		e = hipEventRecord(a, 0);
		for (int r = 0; r < 3 && e == hipSuccess; r++) {
			if (bytes_per_line == 64)
				hipLaunchKernelGGL(k_probe_gather<4>, blocks, 256, 0, 0, tab, nlines, iters, out);
			else if (bytes_per_line == 96)
				hipLaunchKernelGGL(k_probe_gather<6>, blocks, 256, 0, 0, tab, nlines, iters, out);
			else
				hipLaunchKernelGGL(k_probe_gather<8>, blocks, 256, 0, 0, tab, nlines, iters, out);
			e = hipGetLastError();
		}
		if (e == hipSuccess) e = hipEventRecord(b, 0);
		if (e == hipSuccess) e = hipEventSynchronize(b);
		if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
	}
	cleanup();
	if (e != hipSuccess)
		return fail(e == hipErrorOutOfMemory ? CRP_E_NOMEM : CRP_E_HIP,
		            std::string("crp_measure_gather: ") + hipGetErrorString(e));
	*lines_per_s = 3.0 * threads * (double)iters / (ms * 1e-3);
	return CRP_OK;
}

} // extern "C"
