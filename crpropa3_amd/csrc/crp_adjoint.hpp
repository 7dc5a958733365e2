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
// Forward stores the state every K steps (checkpoints); backward walks the
// segments from last to first, recomputes the segment's K states into a
// per-particle scratch area, and applies the reverse step K times.
#pragma once

#include "crp_device.hpp"

namespace crp {

struct AdjGrid {
	const float *g; // (nx*ny + iy)*nz + iz, 3 floats per point
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

__device__ __forceinline__ V3 adj_interp(const AdjGrid &G, const TriPoint &tp) {
	double b[3] = {0, 0, 0};
#pragma unroll
	for (int c = 0; c < 8; c++) {
		double w, a, bb, cc;
		adj_weight(tp, c, w, a, bb, cc);
		const float *v = G.g + 3 * adj_corner(G, tp, c);
		b[0] += w * (double)v[0];
		b[1] += w * (double)v[1];
		b[2] += w * (double)v[2];
	}
	return v3(b[0], b[1], b[2]);
}

__device__ __forceinline__ V3 vadd(const V3 &a, const V3 &b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 vsub(const V3 &a, const V3 &b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 vmul(const V3 &a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ double vdot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 vcross(const V3 &a, const V3 &b) {
	return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// one forward step of the differentiable model
__device__ __forceinline__ void adj_step(const AdjGrid &G, double h, double kq, V3 &x, V3 &u) {
	const V3 x1 = vadd(x, vmul(u, h / 2));
	TriPoint tp;
	adj_cell(G, x1, tp);
	const V3 B = vmul(adj_interp(G, tp), G.scale);
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

// reverse of one step: (x, u) is the state before the step; (ax, au) the
// adjoint after it on entry and before it on return.
__device__ __forceinline__ void adj_step_reverse(const AdjGrid &G, double h, double kq, const V3 &x, const V3 &u, V3 &ax,
                                                 V3 &au, double *__restrict__ gradG, double &gradScale) {
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

} // namespace crp
