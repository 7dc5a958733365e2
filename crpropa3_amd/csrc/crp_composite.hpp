// Field decorators on the device (included by crp_kernels.hip): PeriodicMagneticField,
// MagneticFieldList and MagneticFieldEvolution (src/magneticField/MagneticField.cpp:41-79,
// include/crpropa/magneticField/MagneticField.h:47-103) over any of the device fields.
//
// The host compiles a tree of decorators into a short program of wave-uniform operations over
// four position slots and four accumulators; the kernels run it with scalar branches.  The
// program reproduces the reference's evaluation order exactly:
//   * MagneticFieldList::getField: b = (0,0,0); b += field_i->getField(position) in list order
//     (a nested list is summed on its own from zero, then added);
//   * PeriodicMagneticField::getField: n = floor((position - origin) / extends),
//     p = position - origin - n * extends, reflected (p = extends - p) where |n| is odd, then the
//     wrapped field at p;
//   * MagneticFieldEvolution overrides only getField(position, z) = field->getField(position) *
//     pow(1 + z, m); its one-argument getField is the base class's zero.  Decorators call their
//     children's one-argument getField, and the propagators the two-argument one of the top
//     field, so z matters only when the top field is an evolution: its factor is formed once per
//     particle (bind_lane) from the candidate's redshift, correctly rounded.
#pragma once

#include "crp_device.hpp"
#include "crp_jf12.hpp"

namespace crp {

constexpr int kCompMaxLeaves = 8, kCompMaxOps = 48, kCompMaxPeriodic = 8, kCompSlots = 4;

// CL_GRID_ANY: a grid or modulated field whose Grid3f or Grid1f is not trilinear (FieldGridAny; `mod`
// holds the modulation when its values are set)
enum CompLeafKind { CL_NONE = 0, CL_UNIFORM = 1, CL_GRID = 2, CL_MODULATED = 3, CL_PLANEWAVE = 4, CL_JF12 = 5,
                    CL_GRID_ANY = 6 };

struct CompLeaf {
	int kind;
	FieldUniform uniform;
	GridDesc grid;                // CL_GRID, CL_MODULATED: the Grid3f
	const float *cells;           // its cell table (periodic grids) or null
	const unsigned char *cells16; // CL_GRID: FP16 brick table, read by the mixed path when set
	GridDesc mod;                 // CL_MODULATED: the Grid1f
	PlaneWaveDesc pw;
	FieldJF12 jf12;
};

// CO_LEAF      acc[c] = leaf a at pos[b]
// CO_LEAF_ADD  acc[c] += leaf a at pos[b]
// CO_PERIODIC  pos[c] = periodic map a of pos[b]
// CO_ZERO      acc[a] = (0, 0, 0)
// CO_ADD       acc[b] += acc[a]
// CO_EVOLVE    acc[a] *= (1 + z)^m of the lane's particle
enum CompOpKind { CO_LEAF = 0, CO_LEAF_ADD = 1, CO_PERIODIC = 2, CO_ZERO = 3, CO_ADD = 4, CO_EVOLVE = 5 };
struct CompOp {
	int op, a, b, c;
};
struct CompPeriodic {
	double origin[3], extends[3];
	int reflective;
	int pad;
};
struct CompositeDesc {
	int nops;
	int evolve; // the top field is a MagneticFieldEvolution with exponent evoM
	double evoM;
	CompOp ops[kCompMaxOps];
	CompPeriodic per[kCompMaxPeriodic];
	CompLeaf leaves[kCompMaxLeaves];
};

// A trivially copyable value read from immutable device memory through the constant address space,
// so that wave-uniform reads become scalar loads.
template <class T> __device__ __forceinline__ T ld_const(const T *p) {
	static_assert(sizeof(T) % 4 == 0, "ld_const: size must be a multiple of 4");
	typedef __attribute__((address_space(4))) const uint32_t cu32;
	cu32 *src = (cu32 *)p;
	T out;
	uint32_t *dst = reinterpret_cast<uint32_t *>(&out);
#pragma unroll
	for (unsigned i = 0; i < sizeof(T) / 4; i++)
		dst[i] = src[i];
	return out;
}

template <class T> __device__ __forceinline__ T slot_get(const T &s0, const T &s1, const T &s2, const T &s3, int i) {
	return i == 0 ? s0 : (i == 1 ? s1 : (i == 2 ? s2 : s3));
}
template <class T> __device__ __forceinline__ void slot_set(T &s0, T &s1, T &s2, T &s3, int i, const T &v) {
	if (i == 0)
		s0 = v;
	else if (i == 1)
		s1 = v;
	else if (i == 2)
		s2 = v;
	else
		s3 = v;
}

// PeriodicMagneticField::getField's position map (src/magneticField/MagneticField.cpp:41-58)
__device__ __forceinline__ V3 periodic_map(const CompPeriodic &c, const V3 &x) {
	const double px[3] = {x.x, x.y, x.z};
	double q[3];
#pragma unroll
	for (int a = 0; a < 3; a++) {
		const double n = __builtin_floor((px[a] - c.origin[a]) / c.extends[a]);
		q[a] = (px[a] - c.origin[a]) - n * c.extends[a];
		if (c.reflective && ((long long)__builtin_fabs(n) % 2) == 1)
			q[a] = c.extends[a] - q[a];
	}
	return v3(q[0], q[1], q[2]);
}

struct FieldComposite {
	const CompositeDesc *d; // device memory, immutable while kernels run
	double zf;              // (1 + z)^m of the lane's particle, for a top-level evolution

	__device__ __forceinline__ void bind_lane(double z) {
		zf = 1.0;
		if (ld_const(&d->evolve))
			zf = crp_pow_cr(1 + z, ld_const(&d->evoM));
	}

	// JF12 and the grids read tricubic or nearest neighbour share one out-of-line call site
	__device__ static __attribute__((noinline)) V3 leaf_eval_call(const CompLeaf *Lp, const V3 &x) {
		if (ld_const(&Lp->kind) == CL_JF12) {
			const FieldJF12 j = ld_const(&Lp->jf12);
			return j.turbulent.interp == kInterpTrilinear ? j.eval(x) : j.eval_any(x);
		}
		const GridDesc mod = ld_const(&Lp->mod);
		return FieldGridAny{ld_const(&Lp->grid), mod, mod.values != nullptr}.eval(x);
	}
	__device__ static __attribute__((noinline)) float3 leaf_evalf_call(const CompLeaf *Lp, const V3 &x) {
		if (ld_const(&Lp->kind) == CL_JF12) {
			const FieldJF12 j = ld_const(&Lp->jf12);
			return j.turbulent.interp == kInterpTrilinear ? j.evalf(x) : FieldJF12Any{j}.evalf(x);
		}
		const GridDesc mod = ld_const(&Lp->mod);
		return FieldGridAny{ld_const(&Lp->grid), mod, mod.values != nullptr}.evalf(x);
	}

	__device__ static V3 leaf_eval(const CompLeaf *Lp, const V3 &x) {
		const int kind = ld_const(&Lp->kind);
		switch (kind) {
		case CL_UNIFORM: return ld_const(&Lp->uniform).eval(x);
		case CL_GRID: {
			const float *cells = ld_const(&Lp->cells);
			if (cells)
				return FieldGridCells{ld_const(&Lp->grid), cells}.eval(x);
			return FieldGrid{ld_const(&Lp->grid)}.eval(x);
		}
		case CL_MODULATED: {
			const float *cells = ld_const(&Lp->cells);
			if (cells)
				return FieldModulated<FieldGridCells>{FieldGridCells{ld_const(&Lp->grid), cells}, ld_const(&Lp->mod)}.eval(x);
			return FieldModulated<FieldGrid>{FieldGrid{ld_const(&Lp->grid)}, ld_const(&Lp->mod)}.eval(x);
		}
		case CL_PLANEWAVE: return FieldPlaneWave{ld_const(&Lp->pw)}.eval(x);
		case CL_JF12:
		case CL_GRID_ANY: return leaf_eval_call(Lp, x);
		default: return v3(0.0, 0.0, 0.0);
		}
	}

	__device__ static float3 leaf_evalf(const CompLeaf *Lp, const V3 &x) {
		const int kind = ld_const(&Lp->kind);
		switch (kind) {
		case CL_UNIFORM: return ld_const(&Lp->uniform).evalf(x);
		case CL_GRID: {
			const unsigned char *c16 = ld_const(&Lp->cells16);
			if (c16)
				return FieldGridCells16{ld_const(&Lp->grid), c16}.evalf(x);
			const float *cells = ld_const(&Lp->cells);
			if (cells)
				return FieldGridCells{ld_const(&Lp->grid), cells}.evalf(x);
			return FieldGrid{ld_const(&Lp->grid)}.evalf(x);
		}
		case CL_MODULATED: {
			const float *cells = ld_const(&Lp->cells);
			if (cells)
				return FieldModulated<FieldGridCells>{FieldGridCells{ld_const(&Lp->grid), cells}, ld_const(&Lp->mod)}.evalf(x);
			return FieldModulated<FieldGrid>{FieldGrid{ld_const(&Lp->grid)}, ld_const(&Lp->mod)}.evalf(x);
		}
		case CL_PLANEWAVE: return FieldPlaneWave{ld_const(&Lp->pw)}.evalf(x);
		case CL_JF12:
		case CL_GRID_ANY: return leaf_evalf_call(Lp, x);
		default: return make_float3(0.f, 0.f, 0.f);
		}
	}

	// getField(position, z) of the top field; out of line, so that the several evaluations of a
	// Cash-Karp trial share one copy of the program interpreter
	__device__ __attribute__((noinline)) V3 eval_prog(const V3 &p, double zfac) const {
		const V3 zero = v3(0.0, 0.0, 0.0);
		V3 P0 = p, P1 = zero, P2 = zero, P3 = zero;
		V3 A0 = zero, A1 = zero, A2 = zero, A3 = zero;
		const int nops = ld_const(&d->nops);
		for (int k = 0; k < nops; k++) {
			const CompOp o = ld_const(&d->ops[k]);
			switch (o.op) {
			case CO_LEAF:
			case CO_LEAF_ADD: {
				const V3 v = leaf_eval(&d->leaves[o.a], slot_get(P0, P1, P2, P3, o.b));
				if (o.op == CO_LEAF) {
					slot_set(A0, A1, A2, A3, o.c, v);
				} else {
					const V3 s = slot_get(A0, A1, A2, A3, o.c);
					slot_set(A0, A1, A2, A3, o.c, v3(s.x + v.x, s.y + v.y, s.z + v.z));
				}
				break;
			}
			case CO_PERIODIC:
				slot_set(P0, P1, P2, P3, o.c, periodic_map(ld_const(&d->per[o.a]), slot_get(P0, P1, P2, P3, o.b)));
				break;
			case CO_ZERO: slot_set(A0, A1, A2, A3, o.a, zero); break;
			case CO_ADD: {
				const V3 s = slot_get(A0, A1, A2, A3, o.b), v = slot_get(A0, A1, A2, A3, o.a);
				slot_set(A0, A1, A2, A3, o.b, v3(s.x + v.x, s.y + v.y, s.z + v.z));
				break;
			}
			default: { // CO_EVOLVE: Vector3d * double
				const V3 s = slot_get(A0, A1, A2, A3, o.a);
				slot_set(A0, A1, A2, A3, o.a, v3(s.x * zfac, s.y * zfac, s.z * zfac));
				break;
			}
			}
		}
		return A0;
	}

	// mixed precision: the same program on the FP32 evaluations (positions stay FP64)
	__device__ __attribute__((noinline)) float3 evalf_prog(const V3 &p, float zfac) const {
		const V3 zp = v3(0.0, 0.0, 0.0);
		const float3 zero = make_float3(0.f, 0.f, 0.f);
		V3 P0 = p, P1 = zp, P2 = zp, P3 = zp;
		float3 A0 = zero, A1 = zero, A2 = zero, A3 = zero;
		const int nops = ld_const(&d->nops);
		for (int k = 0; k < nops; k++) {
			const CompOp o = ld_const(&d->ops[k]);
			switch (o.op) {
			case CO_LEAF:
			case CO_LEAF_ADD: {
				const float3 v = leaf_evalf(&d->leaves[o.a], slot_get(P0, P1, P2, P3, o.b));
				if (o.op == CO_LEAF) {
					slot_set(A0, A1, A2, A3, o.c, v);
				} else {
					const float3 s = slot_get(A0, A1, A2, A3, o.c);
					slot_set(A0, A1, A2, A3, o.c, make_float3(s.x + v.x, s.y + v.y, s.z + v.z));
				}
				break;
			}
			case CO_PERIODIC:
				slot_set(P0, P1, P2, P3, o.c, periodic_map(ld_const(&d->per[o.a]), slot_get(P0, P1, P2, P3, o.b)));
				break;
			case CO_ZERO: slot_set(A0, A1, A2, A3, o.a, zero); break;
			case CO_ADD: {
				const float3 s = slot_get(A0, A1, A2, A3, o.b), v = slot_get(A0, A1, A2, A3, o.a);
				slot_set(A0, A1, A2, A3, o.b, make_float3(s.x + v.x, s.y + v.y, s.z + v.z));
				break;
			}
			default: {
				const float3 s = slot_get(A0, A1, A2, A3, o.a);
				slot_set(A0, A1, A2, A3, o.a, make_float3(s.x * zfac, s.y * zfac, s.z * zfac));
				break;
			}
			}
		}
		return A0;
	}

	__device__ __forceinline__ V3 eval(const V3 &p) const { return eval_prog(p, zf); }
	__device__ __forceinline__ float3 evalf(const V3 &p) const { return evalf_prog(p, (float)zf); }
};

// Per-lane binding of the candidate's redshift: only the decorator program uses it.
template <class F> struct UsesRedshift {
	static constexpr bool value = false;
};
template <> struct UsesRedshift<FieldComposite> {
	static constexpr bool value = true;
};
template <class F> __device__ __forceinline__ void bind_redshift(F &, double) {}
__device__ __forceinline__ void bind_redshift(FieldComposite &f, double z) { f.bind_lane(z); }

} // namespace crp
