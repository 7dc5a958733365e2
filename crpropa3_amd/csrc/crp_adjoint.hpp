// Adjoint (reverse-mode derivative) of the fixed-step Boris push through a
// periodic Grid3f (included by crp_kernels.hip).
//
// Differentiable model, per step of length h (PropagationBP::dY,
// src/module/PropagationBP.cpp:19-39, followed by the two normalisations of
// src/module/PropagationBP.cpp:121 and src/ParticleState.cpp:30-32):
//   x1 = x + u h/2
//   B  = s * T(x1)            T: trilinear interpolation of the grid values G,
//                              evaluated in FP64 (a smooth model of Grid.h:540-582)
//   t  = kq B                  kq = q c h / (2 E)  ( = q/(2m) h/c with m = E/c^2 )
//   s2 = 2 t / (1 + t.t)
//   v  = u + u x t
//   u' = u + v x s2
//   x' = x1 + u' h/2,  u_next = n(n(u')),  n(a) = a/|a|
// Gradients flow to the grid values G (scatter-add, FP64 atomics), to the
// initial positions and directions, and to the field scale s.
//
// The model's arithmetic is its own (compared with the oracle's FP64 gradient within a tolerance, not
// bit for bit), so its functions contract a*b + c into fused multiply-adds: `#pragma clang fp
// contract(on)` at the top of their bodies (the translation unit is built with -ffp-contract=off, which
// the propagation kernels need).  The contraction is decided per expression in the source, so the
// forward and every recompute produce the same bits.
// Forward stores the state every K steps (checkpoints); backward walks the
// segments from last to first, recomputes the segment's K states into a
// per-particle scratch area, and applies the reverse step K times.
#pragma once

#include "crp_device.hpp"

namespace crp {

struct AdjGrid {
	const float *g;     // (nx*ny + iy)*nz + iz, 3 floats per point
	const float *cells; // null, or per cell (ix*ny + iy)*nz + iz its 8 corner values in corner order, 32 floats
	int nx, ny, nz;
	double gridOrigin[3], invSpacing[3], invN[3];
	double scale;
};

// FP64 trilinear interpolation and the corner data the reverse pass needs:
// lower and upper corner index per axis (periodic) and the fractions.
struct TriPoint {
	int lo[3], hi[3];
	double fx, fy, fz;
};

__device__ __forceinline__ void adj_cell(const AdjGrid &G, const V3 &p, TriPoint &tp) {
	const double r[3] = {(p.x - G.gridOrigin[0]) * G.invSpacing[0], (p.y - G.gridOrigin[1]) * G.invSpacing[1],
	                     (p.z - G.gridOrigin[2]) * G.invSpacing[2]};
	const int n[3] = {G.nx, G.ny, G.nz};
	double f[3];
#pragma unroll
	for (int a = 0; a < 3; a++) {
		const double fl = __builtin_floor(r[a]);
		f[a] = r[a] - fl;
		// (int) saturates beyond +-2^31; wrap_index maps any int into [0, n)
		const int i = wrap_index((int)fl, n[a], G.invN[a]);
		tp.lo[a] = i;
		tp.hi[a] = (i + 1 == n[a]) ? 0 : i + 1;
	}
	tp.fx = f[0];
	tp.fy = f[1];
	tp.fz = f[2];
}

// flat point index of corner c (corner order of Grid.h:572-579)
__device__ __forceinline__ size_t adj_corner(const AdjGrid &G, const TriPoint &tp, int c) {
	const int cx[8] = {0, 1, 0, 0, 1, 0, 1, 1}, cy[8] = {0, 0, 1, 0, 0, 1, 1, 1}, cz[8] = {0, 0, 0, 1, 1, 1, 0, 1};
	const int X = cx[c] ? tp.hi[0] : tp.lo[0], Yv = cy[c] ? tp.hi[1] : tp.lo[1], Z = cz[c] ? tp.hi[2] : tp.lo[2];
	return ((size_t)X * G.ny + Yv) * G.nz + Z;
}

// corner weight and its gradient with respect to the fractions
__device__ __forceinline__ void adj_weight(const TriPoint &tp, int c, double &w, double &dwx, double &dwy, double &dwz) {
	const int cx[8] = {0, 1, 0, 0, 1, 0, 1, 1}, cy[8] = {0, 0, 1, 0, 0, 1, 1, 1}, cz[8] = {0, 0, 0, 1, 1, 1, 0, 1};
	const double wx = cx[c] ? tp.fx : 1 - tp.fx, wy = cy[c] ? tp.fy : 1 - tp.fy, wz = cz[c] ? tp.fz : 1 - tp.fz;
	const double sx = cx[c] ? 1.0 : -1.0, sy = cy[c] ? 1.0 : -1.0, sz = cz[c] ? 1.0 : -1.0;
	w = wx * wy * wz;
	dwx = sx * wy * wz;
	dwy = wx * sy * wz;
	dwz = wx * wy * sz;
}

// the 24 corner values of the cell, in corner order: one 128-byte line of the cell table when there is
// one, else 8 gathers of the point grid
__device__ __forceinline__ void adj_gather(const AdjGrid &G, const TriPoint &tp, float v[24]) {
	if (G.cells) {
		const size_t c = ((size_t)tp.lo[0] * G.ny + tp.lo[1]) * G.nz + tp.lo[2];
		const float4 *q = reinterpret_cast<const float4 *>(G.cells + 32 * c);
#pragma unroll
		for (int k = 0; k < 6; k++) {
			const float4 f = q[k];
			v[4 * k] = f.x; v[4 * k + 1] = f.y; v[4 * k + 2] = f.z; v[4 * k + 3] = f.w;
		}
	} else {
#pragma unroll
		for (int k = 0; k < 8; k++) {
			const float *gv = G.g + 3 * adj_corner(G, tp, k);
			v[3 * k] = gv[0]; v[3 * k + 1] = gv[1]; v[3 * k + 2] = gv[2];
		}
	}
}

__device__ __forceinline__ V3 adj_interp(const AdjGrid &G, const TriPoint &tp) {
#pragma clang fp contract(on)
	float v[24];
	adj_gather(G, tp, v);
	double b[3] = {0, 0, 0};
#pragma unroll
	for (int c = 0; c < 8; c++) {
		double w, a, bb, cc;
		adj_weight(tp, c, w, a, bb, cc);
		b[0] += w * (double)v[3 * c];
		b[1] += w * (double)v[3 * c + 1];
		b[2] += w * (double)v[3 * c + 2];
	}
	return v3(b[0], b[1], b[2]);
}

__device__ __forceinline__ V3 vadd(const V3 &a, const V3 &b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 vsub(const V3 &a, const V3 &b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 vmul(const V3 &a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ double vdot(const V3 &a, const V3 &b) {
#pragma clang fp contract(on)
	return a.x * b.x + a.y * b.y + a.z * b.z;
}
__device__ __forceinline__ V3 vcross(const V3 &a, const V3 &b) {
#pragma clang fp contract(on)
	return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// the step after the interpolation: T = trilinear(grid) at the half-step position x1
__device__ __forceinline__ void adj_push(const AdjGrid &G, double h, double kq, const V3 &T, const V3 &x1, V3 &x, V3 &u) {
#pragma clang fp contract(on)
	const V3 B = vmul(T, G.scale);
	const V3 t = vmul(B, kq);
	const double D = 1 + vdot(t, t);
	const V3 s2 = vmul(t, 2 / D);
	const V3 v = vadd(u, vcross(u, t));
	const V3 up = vadd(u, vcross(v, s2));
	x = vadd(x1, vmul(up, h / 2));
	const double r1 = __builtin_sqrt(vdot(up, up));
	const V3 n1 = vmul(up, 1 / r1);
	const double r2 = __builtin_sqrt(vdot(n1, n1));
	u = vmul(n1, 1 / r2);
}

// one forward step of the differentiable model
__device__ __forceinline__ void adj_step(const AdjGrid &G, double h, double kq, V3 &x, V3 &u) {
	const V3 x1 = vadd(x, vmul(u, h / 2));
	TriPoint tp;
	adj_cell(G, x1, tp);
	adj_push(G, h, kq, adj_interp(G, tp), x1, x, u);
}

// reverse of one step: (x, u) is the state before the step; (ax, au) the
// adjoint after it on entry and before it on return.
__device__ __forceinline__ void adj_step_reverse(const AdjGrid &G, double h, double kq, const V3 &x, const V3 &u, V3 &ax,
                                                 V3 &au, double *__restrict__ gradG, double &gradScale) {
#pragma clang fp contract(on)
	// recompute the forward intermediates
	const V3 x1 = vadd(x, vmul(u, h / 2));
	TriPoint tp;
	adj_cell(G, x1, tp);
	const V3 T = adj_interp(G, tp);
	const V3 B = vmul(T, G.scale);
	const V3 t = vmul(B, kq);
	const double D = 1 + vdot(t, t);
	const V3 s2 = vmul(t, 2 / D);
	const V3 v = vadd(u, vcross(u, t));
	const V3 up = vadd(u, vcross(v, s2));
	const double r1 = __builtin_sqrt(vdot(up, up));
	const V3 n1 = vmul(up, 1 / r1);
	const double r2 = __builtin_sqrt(vdot(n1, n1));
	const V3 un = vmul(n1, 1 / r2);
	// u_next = n(n1), n1 = n(u')
	V3 an1 = vmul(vsub(au, vmul(un, vdot(un, au))), 1 / r2);
	V3 aup = vmul(vsub(an1, vmul(n1, vdot(n1, an1))), 1 / r1);
	// x' = x1 + u' h/2
	aup = vadd(aup, vmul(ax, h / 2));
	V3 ax1 = ax;
	// u' = u + v x s2
	V3 aU = aup;
	const V3 av = vcross(s2, aup);
	const V3 as2 = vcross(aup, v);
	// v = u + u x t
	aU = vadd(aU, vadd(av, vcross(t, av)));
	V3 at = vcross(av, u);
	// s2 = 2 t / D
	at = vadd(at, vsub(vmul(as2, 2 / D), vmul(t, 4 / (D * D) * vdot(t, as2))));
	// t = kq B, B = s T
	const V3 aB = vmul(at, kq);
	gradScale += vdot(T, aB);
	const V3 aT = vmul(aB, G.scale);
	// T = sum_c w_c G_c
	double gx = 0, gy = 0, gz = 0;
#pragma unroll
	for (int c = 0; c < 8; c++) {
		double w, dwx, dwy, dwz;
		adj_weight(tp, c, w, dwx, dwy, dwz);
		const size_t ic = adj_corner(G, tp, c);
		const float *gv = G.g + 3 * ic;
		const double dotc = (double)gv[0] * aT.x + (double)gv[1] * aT.y + (double)gv[2] * aT.z;
		gx += dwx * dotc;
		gy += dwy * dotc;
		gz += dwz * dotc;
		// hardware FP64 atomic add in L2 (the gradient buffer is ordinary device memory, not
		// fine-grained host-coherent memory, so the compare-and-swap loop of atomicAdd is not needed)
		double *o = gradG + 3 * ic;
		unsafeAtomicAdd(o, w * aT.x);
		unsafeAtomicAdd(o + 1, w * aT.y);
		unsafeAtomicAdd(o + 2, w * aT.z);
	}
	ax1 = vadd(ax1, v3(gx * G.invSpacing[0], gy * G.invSpacing[1], gz * G.invSpacing[2]));
	// x1 = x + u h/2
	ax = ax1;
	au = vadd(aU, vmul(ax1, h / 2));
}

struct AdjArgs {
	AdjGrid G;
	double h;
	int nsteps, K;
	size_t n;
	const double *x0, *u0, *kq; // AoS (n,3), (n,3), (n)
	double *ckpt;               // [nseg+1][6][n]
	double *xout, *uout;        // AoS
	const double *gxout, *guout; // AoS adjoints of the outputs (gxout may be null)
	double *scratch;            // [K][6][n]
	double *gradG, *gradX0, *gradU0, *gradScale;
};

__global__ void __launch_bounds__(256) k_adjoint_forward(AdjArgs A) {
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= A.n)
		return;
	V3 x = v3(A.x0[3 * i], A.x0[3 * i + 1], A.x0[3 * i + 2]);
	V3 u = v3(A.u0[3 * i], A.u0[3 * i + 1], A.u0[3 * i + 2]);
	const double kq = A.kq[i];
	const size_t n = A.n;
	for (int s = 0; s < A.nsteps; s++) {
		if (s % A.K == 0) {
			double *c = A.ckpt + (size_t)(s / A.K) * 6 * n;
			c[i] = x.x; c[n + i] = x.y; c[2 * n + i] = x.z;
			c[3 * n + i] = u.x; c[4 * n + i] = u.y; c[5 * n + i] = u.z;
		}
		adj_step(A.G, A.h, kq, x, u);
	}
	A.xout[3 * i] = x.x; A.xout[3 * i + 1] = x.y; A.xout[3 * i + 2] = x.z;
	A.uout[3 * i] = u.x; A.uout[3 * i + 1] = u.y; A.uout[3 * i + 2] = u.z;
}

__global__ void __launch_bounds__(256) k_adjoint_backward(AdjArgs A) {
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	const bool live = i < A.n; // every lane takes part in the final wave reduction
	const size_t n = A.n;
	double gs = 0;
	const int nseg = live ? (A.nsteps + A.K - 1) / A.K : 0;
	const double kq = live ? A.kq[i] : 0.0;
	V3 ax = (live && A.gxout) ? v3(A.gxout[3 * i], A.gxout[3 * i + 1], A.gxout[3 * i + 2]) : v3(0, 0, 0);
	V3 au = live ? v3(A.guout[3 * i], A.guout[3 * i + 1], A.guout[3 * i + 2]) : v3(0, 0, 0);
	for (int seg = nseg - 1; seg >= 0; seg--) {
		const int s0 = seg * A.K;
		const int len = min(A.K, A.nsteps - s0);
		const double *c = A.ckpt + (size_t)seg * 6 * n;
		V3 x = v3(c[i], c[n + i], c[2 * n + i]);
		V3 u = v3(c[3 * n + i], c[4 * n + i], c[5 * n + i]);
		for (int j = 0; j < len; j++) {
			double *s = A.scratch + (size_t)j * 6 * n;
			s[i] = x.x; s[n + i] = x.y; s[2 * n + i] = x.z;
			s[3 * n + i] = u.x; s[4 * n + i] = u.y; s[5 * n + i] = u.z;
			adj_step(A.G, A.h, kq, x, u);
		}
		for (int j = len - 1; j >= 0; j--) {
			const double *s = A.scratch + (size_t)j * 6 * n;
			const V3 xs = v3(s[i], s[n + i], s[2 * n + i]);
			const V3 us = v3(s[3 * n + i], s[4 * n + i], s[5 * n + i]);
			adj_step_reverse(A.G, A.h, kq, xs, us, ax, au, A.gradG, gs);
		}
	}
	if (live) {
		A.gradX0[3 * i] = ax.x; A.gradX0[3 * i + 1] = ax.y; A.gradX0[3 * i + 2] = ax.z;
		A.gradU0[3 * i] = au.x; A.gradU0[3 * i + 1] = au.y; A.gradU0[3 * i + 2] = au.z;
	}
	// wave reduction of the scale gradient, one atomic per wave
#pragma unroll
	for (int o = 32; o > 0; o >>= 1)
		gs += __shfl_xor(gs, o);
	if ((threadIdx.x & 63) == 0)
		unsafeAtomicAdd(A.gradScale, gs);
}


// ---------------------------------------------------------------- backward in blocks of segments
// The checkpointed backward above walks every particle's segments one after the other: each
// segment's recompute and reverse sweep wait on one gather per step at one wave per SIMD, and
// every reverse step adds 24 FP64 atomics into the grid gradient.  The blocked backward instead
//   1. recomputes a block of segments of every particle at once (segments x n independent lanes),
//      stores the state (x, u) before every kAdjSub-th step of the block (sub-checkpoints, 6 doubles
//      per kAdjSub steps) and counts, per workgroup, the steps whose half-step position falls in each
//      tile of 2^S cells per axis;
//   2. scans the counts into every tile's start and every particle group's place within each tile;
//   3. runs the reverse steps of the block per particle (one lane per particle, the only sequential
//      part): for each run of kAdjSub steps, from the last to the first, it recomputes the run from
//      its sub-checkpoint, keeping in registers what the reverse steps read (the direction u before
//      the step, the interpolated field T and its derivatives with respect to the cell fractions,
//      the cell and the fractions), then applies the reverse steps and writes each step's 32-byte
//      record (with the adjoint aT of T: the gradient of grid value G_c is w_c aT) into its tile's range;
//   4. accumulates each tile's records into its 9^3 grid points in LDS (FP64), one workgroup per
//      kDepPerGroup records of a tile, and adds the points into the gradient.
// Per step the backward thus reads the cell line twice (recompute, reverse) and moves one record
// through memory; the stored values of a step never leave the registers.  The host runs the reverse
// sweeps on the caller's stream and the recompute of the next block and the accumulation of the
// previous one on two more streams beside it.  The arithmetic of a step is that of adj_step /
// adj_step_reverse, except that the position gradient is taken as aT . dT/df; the deposits are
// summed in another order.
// deposit tiles of 2^S cells per axis: 8^3 (S = 3, 9^3 points = 17.5 KB of FP64 LDS per accumulating
// workgroup of 256) or, for grids with too many of those, 16^3 (S = 4, 118 KB, workgroups of 1024)
template <int S>
struct Tile {
	static constexpr int shift = S, cells = 1 << S, pts = cells + 1, ptsAll = pts * pts * pts, cellBits = 3 * S;
	static constexpr int threads = S == 3 ? 256 : 1024;
};
constexpr int kAdjGroup = 256;       // particles per recompute / reverse workgroup
constexpr int kDepPerGroup = 65536;  // records per accumulating workgroup
constexpr int kDepBatch = 4;
constexpr int kAdjSub = 8;           // steps between the sub-checkpoints of a block
// what one reverse step reads: u, T, dT/dfx, dT/dfy, dT/dfz
constexpr int kAdjStored = 15;

struct __align__(16) DepRec {
	uint32_t key; // tile << 3S | cell in tile (x << 2S | y << S | z) for tiles of 2^S cells per axis
	float f[3];   // fractions of the half-step position in its cell
	float a[3];   // adjoint of the interpolated field
	uint32_t pad;
};

struct AdjBlock {
	int seg0, nseg;      // segments of this block
	int s0, s1;          // steps [s0, s1)
	int first, last;     // first block of the backward (reads the output adjoints); last block
	int npg;             // particle groups: ceil(n / kAdjGroup)
	double *sub;         // [q / kAdjSub][6][n]: x, u before step s0 + q, q a multiple of kAdjSub
	double *carry;       // [6][n]: adjoint (x, u) between blocks
	DepRec *srec;        // [m], grouped by tile
	uint32_t *counts;    // [npg][ntiles]: records per particle group and tile, then their offsets
	uint32_t *tileStart; // [ntiles + 1]
	uint32_t *wgStart;   // [ntiles + 1]: first accumulating workgroup of every tile
	int ntx, nty, ntz, ntiles;
	size_t m;            // records of the block: (s1 - s0) * n
};

__global__ void __launch_bounds__(kAdjGroup) k_dep_zero(AdjBlock Bk) {
	const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k < (size_t)Bk.npg * Bk.ntiles)
		Bk.counts[k] = 0u;
}

// the record key of a cell: tile << 3S | cell in tile (x << 2S | y << S | z)
template <int S>
__device__ __forceinline__ uint32_t adj_key(const AdjBlock &Bk, const TriPoint &tp) {
	const int cx = tp.lo[0], cy = tp.lo[1], cz = tp.lo[2];
	const uint32_t tile = (uint32_t)(((cx >> S) * Bk.nty + (cy >> S)) * Bk.ntz + (cz >> S));
	const uint32_t m = (1u << S) - 1;
	return (tile << (3 * S)) | (((uint32_t)cx & m) << (2 * S)) | (((uint32_t)cy & m) << S) | ((uint32_t)cz & m);
}

// One lane per (segment, particle); a workgroup holds kAdjGroup particles of one segment.  Every step
// is the forward's (adj_step); the lane stores the sub-checkpoints (every L-th step of the block) that
// fall in its segment and counts the step's tile.
template <int S, int L>
__global__ void __launch_bounds__(kAdjGroup) k_adj_recompute(AdjArgs A, AdjBlock Bk) {
	extern __shared__ uint32_t hist[];
	const int T = Bk.ntiles;
	for (int k = threadIdx.x; k < T; k += blockDim.x)
		hist[k] = 0u;
	__syncthreads();
	const int segl = blockIdx.x / Bk.npg, pg = blockIdx.x % Bk.npg;
	const size_t n = A.n, i = (size_t)pg * kAdjGroup + threadIdx.x;
	if (i < n) {
		const int seg = Bk.seg0 + segl;
		const int len = min(A.K, A.nsteps - seg * A.K);
		const double *c = A.ckpt + (size_t)seg * 6 * n;
		V3 x = v3(c[i], c[n + i], c[2 * n + i]);
		V3 u = v3(c[3 * n + i], c[4 * n + i], c[5 * n + i]);
		const double kq = A.kq[i];
		int q = seg * A.K - Bk.s0;
		for (int j = 0; j < len; j++, q++) {
			if (q % L == 0) {
				double *o = Bk.sub + (size_t)(q / L) * 6 * n + i;
				__builtin_nontemporal_store(x.x, o);
				__builtin_nontemporal_store(x.y, o + n);
				__builtin_nontemporal_store(x.z, o + 2 * n);
				__builtin_nontemporal_store(u.x, o + 3 * n);
				__builtin_nontemporal_store(u.y, o + 4 * n);
				__builtin_nontemporal_store(u.z, o + 5 * n);
			}
			const V3 x1 = vadd(x, vmul(u, A.h / 2));
			TriPoint tp;
			adj_cell(A.G, x1, tp);
			const V3 Tv = adj_interp(A.G, tp);
			atomicAdd(&hist[adj_key<S>(Bk, tp) >> (3 * S)], 1u);
			adj_push(A.G, A.h, kq, Tv, x1, x, u);
		}
	}
	__syncthreads();
	uint32_t *cnt = Bk.counts + (size_t)pg * T;
	for (int k = threadIdx.x; k < T; k += blockDim.x)
		if (hist[k])
			atomicAdd(cnt + k, hist[k]);
}

// per tile: every particle group's offset within the tile (exclusive scan over the groups, loaded
// kScanBatch groups at a time), and the tile's total in tileStart[t + 1]
constexpr int kScanBatch = 16;
__global__ void __launch_bounds__(256) k_dep_scan_chunks(AdjBlock Bk) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= Bk.ntiles)
		return;
	uint32_t acc = 0;
	for (int c0 = 0; c0 < Bk.npg; c0 += kScanBatch) {
		uint32_t v[kScanBatch];
#pragma unroll
		for (int k = 0; k < kScanBatch; k++)
			v[k] = (c0 + k < Bk.npg) ? Bk.counts[(size_t)(c0 + k) * Bk.ntiles + t] : 0u;
#pragma unroll
		for (int k = 0; k < kScanBatch; k++) {
			if (c0 + k < Bk.npg)
				Bk.counts[(size_t)(c0 + k) * Bk.ntiles + t] = acc;
			acc += v[k];
		}
	}
	Bk.tileStart[t + 1] = acc;
}

// exclusive scans of the tile totals into tile starts and of the tiles' workgroup counts into
// wgStart (one workgroup; at most kScanPer tiles per thread).  Every total is read into registers
// before the first barrier and every start written after it, since thread j's starts overwrite the
// totals thread j - 1 reads.
constexpr int kScanPer = 8;
__global__ void __launch_bounds__(1024) k_dep_scan_tiles(AdjBlock Bk) {
	__shared__ uint32_t part[1024], partw[1024];
	const int T = Bk.ntiles;
	const int per = (T + 1023) / 1024;
	const int b = threadIdx.x * per, e = min(b + per, T);
	uint32_t v[kScanPer];
	uint32_t sum = 0, sumw = 0;
#pragma unroll
	for (int k = 0; k < kScanPer; k++) {
		v[k] = (b + k < e) ? Bk.tileStart[b + k + 1] : 0u;
		sum += v[k];
		sumw += (v[k] + kDepPerGroup - 1) / kDepPerGroup;
	}
	part[threadIdx.x] = sum;
	partw[threadIdx.x] = sumw;
	__syncthreads();
	if (threadIdx.x == 0) {
		uint32_t acc = 0, accw = 0;
		for (int k = 0; k < 1024; k++) {
			const uint32_t x = part[k], y = partw[k];
			part[k] = acc;
			partw[k] = accw;
			acc += x;
			accw += y;
		}
	}
	__syncthreads();
	uint32_t acc = part[threadIdx.x], accw = partw[threadIdx.x];
#pragma unroll
	for (int k = 0; k < kScanPer; k++) {
		if (b + k < e) {
			Bk.tileStart[b + k] = acc;
			Bk.wgStart[b + k] = accw;
			acc += v[k];
			accw += (v[k] + kDepPerGroup - 1) / kDepPerGroup;
		}
	}
	if (e == T && b < e) {
		Bk.tileStart[T] = acc;
		Bk.wgStart[T] = accw;
	}
}

// the values one reverse step reads
struct AdjStored {
	double d[kAdjStored]; // u, T, dT/dfx, dT/dfy, dT/dfz
	uint32_t rec[4];      // key, fractions (plain words: a vector-type member sends copies through memory)
};

// the stored values of a step from its cell, corner values cv and direction u: T in the order of
// adj_interp, and its derivatives with respect to the fractions
template <int S>
__device__ __forceinline__ void adj_stored(const AdjBlock &Bk, const TriPoint &tp, const float cv[24], const V3 &u,
                                           AdjStored &d) {
#pragma clang fp contract(on)
	double b[3] = {0, 0, 0}, jx[3] = {0, 0, 0}, jy[3] = {0, 0, 0}, jz[3] = {0, 0, 0};
#pragma unroll
	for (int k = 0; k < 8; k++) {
		double w, dwx, dwy, dwz;
		adj_weight(tp, k, w, dwx, dwy, dwz);
		const double g0 = cv[3 * k], g1 = cv[3 * k + 1], g2 = cv[3 * k + 2];
		b[0] += w * g0; b[1] += w * g1; b[2] += w * g2;
		jx[0] += dwx * g0; jx[1] += dwx * g1; jx[2] += dwx * g2;
		jy[0] += dwy * g0; jy[1] += dwy * g1; jy[2] += dwy * g2;
		jz[0] += dwz * g0; jz[1] += dwz * g1; jz[2] += dwz * g2;
	}
	d.d[0] = u.x; d.d[1] = u.y; d.d[2] = u.z;
	d.d[3] = b[0]; d.d[4] = b[1]; d.d[5] = b[2];
	d.d[6] = jx[0]; d.d[7] = jx[1]; d.d[8] = jx[2];
	d.d[9] = jy[0]; d.d[10] = jy[1]; d.d[11] = jy[2];
	d.d[12] = jz[0]; d.d[13] = jz[1]; d.d[14] = jz[2];
	d.rec[0] = adj_key<S>(Bk, tp);
	d.rec[1] = __float_as_uint((float)tp.fx);
	d.rec[2] = __float_as_uint((float)tp.fy);
	d.rec[3] = __float_as_uint((float)tp.fz);
}

// adj_step_reverse from the stored u, T and dT/df: (ax, au) the adjoint after the step on entry, before
// it on return; returns aT
__device__ __forceinline__ V3 adj_reverse_stored(const AdjGrid &G, double h, double kq, const AdjStored &d, V3 &ax,
                                                 V3 &au, double &gradScale) {
#pragma clang fp contract(on)
	const V3 u = v3(d.d[0], d.d[1], d.d[2]);
	const V3 T = v3(d.d[3], d.d[4], d.d[5]);
	const V3 B = vmul(T, G.scale);
	const V3 t = vmul(B, kq);
	const double D = 1 + vdot(t, t);
	const V3 s2 = vmul(t, 2 / D);
	const V3 v = vadd(u, vcross(u, t));
	const V3 up = vadd(u, vcross(v, s2));
	const double r1 = __builtin_sqrt(vdot(up, up));
	const V3 n1 = vmul(up, 1 / r1);
	const double r2 = __builtin_sqrt(vdot(n1, n1));
	const V3 un = vmul(n1, 1 / r2);
	V3 an1 = vmul(vsub(au, vmul(un, vdot(un, au))), 1 / r2);
	V3 aup = vmul(vsub(an1, vmul(n1, vdot(n1, an1))), 1 / r1);
	aup = vadd(aup, vmul(ax, h / 2));
	V3 ax1 = ax;
	V3 aU = aup;
	const V3 av = vcross(s2, aup);
	const V3 as2 = vcross(aup, v);
	aU = vadd(aU, vadd(av, vcross(t, av)));
	V3 at = vcross(av, u);
	at = vadd(at, vsub(vmul(as2, 2 / D), vmul(t, 4 / (D * D) * vdot(t, as2))));
	const V3 aB = vmul(at, kq);
	gradScale += vdot(T, aB);
	const V3 aT = vmul(aB, G.scale);
	const double gx = aT.x * d.d[6] + aT.y * d.d[7] + aT.z * d.d[8];
	const double gy = aT.x * d.d[9] + aT.y * d.d[10] + aT.z * d.d[11];
	const double gz = aT.x * d.d[12] + aT.y * d.d[13] + aT.z * d.d[14];
	ax1 = vadd(ax1, v3(gx * G.invSpacing[0], gy * G.invSpacing[1], gz * G.invSpacing[2]));
	ax = ax1;
	au = vadd(aU, vmul(ax1, h / 2));
	return aT;
}

// One lane per particle, kAdjGroup particles per workgroup; each step's record goes to its tile's range
// at the group's offset (counts, from the scans) plus its rank among the group's records of the tile.
// one wave per SIMD (n / 64 waves in all): the kernel may use the whole register file, and holds a run
// of L steps' stored values in it (the step and reverse-step arithmetic, FP64 at one wave per SIMD,
// is what bounds it, not the run's gathers)
template <int S, int L>
__global__ void __launch_bounds__(kAdjGroup) __attribute__((amdgpu_waves_per_eu(1, 1))) k_adj_reverse(AdjArgs A, AdjBlock Bk) {
	extern __shared__ uint32_t lds[];
	const int T = Bk.ntiles;
	uint32_t *base = lds, *rank = lds + T;
	const uint32_t *off = Bk.counts + (size_t)blockIdx.x * T;
	for (int t = threadIdx.x; t < T; t += blockDim.x) {
		base[t] = Bk.tileStart[t] + off[t];
		rank[t] = 0u;
	}
	__syncthreads();
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	const bool live = i < A.n;
	const size_t n = A.n;
	double gs = 0;
	if (live) {
		V3 ax, au;
		if (Bk.first) {
			ax = A.gxout ? v3(A.gxout[3 * i], A.gxout[3 * i + 1], A.gxout[3 * i + 2]) : v3(0, 0, 0);
			au = v3(A.guout[3 * i], A.guout[3 * i + 1], A.guout[3 * i + 2]);
		} else {
			ax = v3(Bk.carry[i], Bk.carry[n + i], Bk.carry[2 * n + i]);
			au = v3(Bk.carry[3 * n + i], Bk.carry[4 * n + i], Bk.carry[5 * n + i]);
		}
		const double kq = A.kq[i];
		const int m = Bk.s1 - Bk.s0;
		for (int sb = (m - 1) / L; sb >= 0; sb--) {
			// steps [sb L, sb L + len) of the block, recomputed from their sub-checkpoint (the step
			// guards, uniform branches, also keep every step in a basic block of its own: one block of
			// L branch-free steps does not fit in the registers)
			const int len = min(L, m - sb * L);
			const double *c = Bk.sub + (size_t)sb * 6 * n + i;
			V3 x = v3(c[0], c[n], c[2 * n]);
			V3 u = v3(c[3 * n], c[4 * n], c[5 * n]);
			AdjStored d[L];
			V3 x1 = vadd(x, vmul(u, A.h / 2));
			TriPoint tp;
			adj_cell(A.G, x1, tp);
			float cv[24];
			adj_gather(A.G, tp, cv);
#pragma unroll
			for (int k = 0; k < L; k++) {
				if (k < len) {
					adj_stored<S>(Bk, tp, cv, u, d[k]);
					adj_push(A.G, A.h, kq, v3(d[k].d[3], d[k].d[4], d[k].d[5]), x1, x, u);
					if (k + 1 < len) {
						x1 = vadd(x, vmul(u, A.h / 2));
						adj_cell(A.G, x1, tp);
						adj_gather(A.G, tp, cv);
					}
				}
			}
#pragma unroll
			for (int k = L - 1; k >= 0; k--) {
				if (k < len) {
					const V3 aT = adj_reverse_stored(A.G, A.h, kq, d[k], ax, au, gs);
					const uint32_t t = d[k].rec[0] >> (3 * S);
					uint4 *pr = reinterpret_cast<uint4 *>(Bk.srec + base[t] + atomicAdd(&rank[t], 1u));
					pr[0] = make_uint4(d[k].rec[0], d[k].rec[1], d[k].rec[2], d[k].rec[3]);
					pr[1] = make_uint4(__float_as_uint((float)aT.x), __float_as_uint((float)aT.y),
					                   __float_as_uint((float)aT.z), 0u);
				}
			}
		}
		if (Bk.last) {
			A.gradX0[3 * i] = ax.x; A.gradX0[3 * i + 1] = ax.y; A.gradX0[3 * i + 2] = ax.z;
			A.gradU0[3 * i] = au.x; A.gradU0[3 * i + 1] = au.y; A.gradU0[3 * i + 2] = au.z;
		} else {
			Bk.carry[i] = ax.x; Bk.carry[n + i] = ax.y; Bk.carry[2 * n + i] = ax.z;
			Bk.carry[3 * n + i] = au.x; Bk.carry[4 * n + i] = au.y; Bk.carry[5 * n + i] = au.z;
		}
	}
#pragma unroll
	for (int o = 32; o > 0; o >>= 1)
		gs += __shfl_xor(gs, o);
	if ((threadIdx.x & 63) == 0)
		unsafeAtomicAdd(A.gradScale, gs);
}

// accumulation: kDepPerGroup records of one tile per workgroup, summed into the tile's 9^3 points in
// LDS (FP64), then added into the grid gradient (points on tile faces are shared with the
// neighbouring tiles, and a tile with many records with the tile's other workgroups).  The grid
// is launched with ntiles + ceil(m / kDepPerGroup) workgroups, an upper bound of wgStart[ntiles].
template <int S>
__global__ void __launch_bounds__(Tile<S>::threads) k_dep_tiles(AdjArgs A, AdjBlock Bk) {
	using Ti = Tile<S>;
	constexpr int kTilePts = Ti::pts, kTilePtsAll = Ti::ptsAll, kTileCells = Ti::cells, kTileShift = Ti::shift;
	__shared__ double acc[kTilePtsAll * 3];
	const uint32_t g = blockIdx.x;
	const int T = Bk.ntiles;
	if (g >= Bk.wgStart[T])
		return;
	// the tile: the last t with wgStart[t] <= g (tiles without records have wgStart[t] == wgStart[t + 1])
	int lo = 0, hi = T - 1;
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if (Bk.wgStart[mid] <= g)
			lo = mid;
		else
			hi = mid - 1;
	}
	const int t = lo;
	const uint32_t r0 = Bk.tileStart[t] + (g - Bk.wgStart[t]) * (uint32_t)kDepPerGroup;
	const uint32_t r1 = min(r0 + (uint32_t)kDepPerGroup, Bk.tileStart[t + 1]);
	for (int k = threadIdx.x; k < kTilePtsAll * 3; k += blockDim.x)
		acc[k] = 0.0;
	__syncthreads();
	const int cx[8] = {0, 1, 0, 0, 1, 0, 1, 1}, cy[8] = {0, 0, 1, 0, 0, 1, 1, 1}, cz[8] = {0, 0, 0, 1, 1, 1, 0, 1};
	constexpr uint32_t m = kTileCells - 1;
	for (uint32_t rb = r0 + threadIdx.x; rb < r1; rb += kDepBatch * blockDim.x) {
		uint4 p0[kDepBatch], p1[kDepBatch];
#pragma unroll
		for (int q = 0; q < kDepBatch; q++) {
			const uint32_t r = rb + q * blockDim.x;
			if (r < r1) {
				const uint4 *ps = reinterpret_cast<const uint4 *>(Bk.srec + r);
				p0[q] = ps[0];
				p1[q] = ps[1];
			}
		}
#pragma unroll
		for (int q = 0; q < kDepBatch; q++) {
			if (rb + q * blockDim.x >= r1)
				continue;
			const uint32_t k = p0[q].x;
			const int lx = (k >> (2 * kTileShift)) & m, ly = (k >> kTileShift) & m, lz = k & m;
			const double fx = __uint_as_float(p0[q].y), fy = __uint_as_float(p0[q].z), fz = __uint_as_float(p0[q].w);
			const double ax = __uint_as_float(p1[q].x), ay = __uint_as_float(p1[q].y), az = __uint_as_float(p1[q].z);
#pragma unroll
			for (int c = 0; c < 8; c++) {
				const double w = (cx[c] ? fx : 1 - fx) * (cy[c] ? fy : 1 - fy) * (cz[c] ? fz : 1 - fz);
				double *a = acc + 3 * (((lx + cx[c]) * kTilePts + (ly + cy[c])) * kTilePts + (lz + cz[c]));
				atomicAdd(a, w * ax);
				atomicAdd(a + 1, w * ay);
				atomicAdd(a + 2, w * az);
			}
		}
	}
	__syncthreads();
	const int tz = t % Bk.ntz, ty = (t / Bk.ntz) % Bk.nty, tx = t / (Bk.ntz * Bk.nty);
	for (int k = threadIdx.x; k < kTilePtsAll * 3; k += blockDim.x) {
		const double v = acc[k];
		if (v == 0.0)
			continue;
		const int p = k / 3, comp = k % 3;
		const int lz = p % kTilePts, ly = (p / kTilePts) % kTilePts, lx = p / (kTilePts * kTilePts);
		const int gx = (tx * kTileCells + lx) % A.G.nx, gy = (ty * kTileCells + ly) % A.G.ny,
		          gz = (tz * kTileCells + lz) % A.G.nz;
		unsafeAtomicAdd(A.gradG + 3 * (((size_t)gx * A.G.ny + gy) * A.G.nz + gz) + comp, v);
	}
}

} // namespace crp
