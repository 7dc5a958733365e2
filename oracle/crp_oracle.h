/*
 * ORACLE -- test infrastructure, not product code.
 *
 * A plain-C restatement of the reference's charged-particle propagation hot
 * path (CRPropa3 PropagationBP / PropagationCK over Grid3f trilinear fields,
 * PlaneWaveTurbulence and uniform fields, MaximumTrajectoryLength, and the
 * MT19937 Random used to draw inputs).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it, as the checker.
 *
 * Every function cites the reference file:line whose arithmetic it restates,
 * in the same operation order (see crp_oracle.c).  libm selects the
 * transcendental functions: ORC_LIBM_GLIBC calls the host's cos()/pow()
 * exactly as the reference does; ORC_LIBM_CR uses the correctly rounded
 * implementations that the GPU kernels use.
 */
#ifndef CRP_ORACLE_H
#define CRP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/crpropa3_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_LIBM_GLIBC = 0, ORC_LIBM_CR = 1 };
enum { ORC_FIELD_NONE = 0, ORC_FIELD_UNIFORM = 1, ORC_FIELD_GRID = 2, ORC_FIELD_PLANEWAVE = 3, ORC_FIELD_MODULATED = 4,
       ORC_FIELD_JF12 = 5, ORC_FIELD_PERIODIC = 6, ORC_FIELD_LIST = 7, ORC_FIELD_EVOLUTION = 8 };
enum { ORC_PROP_BP = 0, ORC_PROP_CK = 1 };
/* modules after the propagator, for orc_propagator.order */
enum { ORC_MOD_BOX = 1, ORC_MOD_SPHERE = 2, ORC_MOD_MAXLEN = 3, ORC_MOD_OBSERVER = 4 };

typedef struct orc_field_s {
	int type;
	double B0[3]; /* uniform */
	/* Grid3f: float32 values, AoS (x,y,z) per point, index (ix*Ny + iy)*Nz + iz */
	const float *grid;
	int nx, ny, nz;
	double origin[3], spacing[3];
	int reflective, clip;
	/* plane waves: Nm modes, xi/kappa AoS 3*Nm */
	int nm;
	const double *k, *Ak, *xi, *kappa, *beta;
	/* ModulatedMagneticFieldGrid: the Grid3f above times a Grid1f (one float per point) */
	const float *mod;
	int mnx, mny, mnz;
	double morigin[3], mspacing[3];
	int mreflective, mclip;
	/* JF12Field: the model's constants; the turbulent Grid3f is the grid above, the striated Grid1f
	 * the modulation grid's slots */
	const crp_jf12_params *jf12;
	/* decorators (src/magneticField/MagneticField.cpp:41-79): PeriodicMagneticField(children[0], extends,
	 * origin, reflective) and MagneticFieldEvolution(children[0], m) wrap one field, MagneticFieldList
	 * sums children[0..nchildren) in order */
	const struct orc_field_s *const *children;
	int nchildren;
	double extends[3], porigin[3];
	int preflective;
	double evo_m;
	/* Grid::setInterpolationType (include/crpropa/Grid.h:24-28,255-264) of the Grid3f (interp) and of
	 * the Grid1f modulation (minterp): CRP_GRID_TRILINEAR, CRP_GRID_TRICUBIC, CRP_GRID_NEAREST_NEIGHBOUR */
	int interp, minterp;
} orc_field;

/* JF12Field / PlanckJF12bField constructor values (src/magneticField/JF12Field.cpp:8-79,348-363) */
void orc_jf12_defaults(int planck, crp_jf12_params *p);

typedef struct {
	int type; /* ORC_PROP_BP / ORC_PROP_CK */
	double tolerance, minStep, maxStep;
	double maxTrajectoryLength; /* <= 0: no MaximumTrajectoryLength module */
	/* Observer with one ObserverTimeEvolution after the length limit (nObs = 0: none):
	 * detection lengths obsTimes[nObs], deactivate on detection */
	const double *obsTimes;
	int nObs, obsDeactivate;
	/* PeriodicBox(origin, size) right after the propagator (box = 0: none) */
	int box;
	double boxOrigin[3], boxSize[3];
	/* SphericalBoundary(center, radius) after the box (sphere = 0: none); limitStep with margin */
	int sphere, sphereLimitStep;
	double sphereCenter[3], sphereRadius, sphereMargin;
	/* the order in which the configured modules run after the propagator (ModuleList order);
	 * nOrder = 0: box, sphere, length limit, observer */
	int order[4], nOrder;
} orc_propagator;

/* Particle state, structure of arrays.  pos/dir are AoS (3 per particle). */
typedef struct {
	double *pos, *dir;
	double *energy, *charge, *redshift;
	double *current_step, *next_step, *trajectory_length, *time;
	int32_t *active;
	double *prev_pos, *prev_dir; /* may be NULL */
	/* observer: detections so far per particle ("DetectionIndex"), and the state at detection k
	 * of particle i: snap[(k * n + i) * 9 + c], c = pos x,y,z, dir x,y,z, length, time, nextStep
	 * (NULL when the propagator has no observer) */
	int32_t *det_index;
	double *snap;
	size_t n; /* particles in the state (snapshot stride) */
	double *source_pos, *created_pos; /* [3n], shifted by PeriodicBox (may be NULL) */
} orc_state;

/* MT19937 as in src/Random.cpp */
typedef struct {
	uint32_t state[624];
	uint32_t *pNext;
	int left;
} orc_rng;

void orc_rng_seed(orc_rng *r, uint32_t seed);
uint32_t orc_rng_randint(orc_rng *r);
double orc_rng_rand(orc_rng *r);
double orc_rng_uniform(orc_rng *r, double min, double max);
double orc_rng_norm(orc_rng *r, double mean, double variance);
void orc_rng_vector(orc_rng *r, double out[3]);
double orc_rng_powerlaw(orc_rng *r, double index, double min, double max);
size_t orc_rng_bin(orc_rng *r, const double *cdf, size_t n);

/* Turbulent grid coefficients before the inverse FFT: SimpleGridTurbulence::initTurbulence
 * (kind ORC_TURB_SIMPLE: amplitude randNorm() * k^(alpha/2), 4 draws per mode) and
 * GridTurbulence::initTurbulence (kind ORC_TURB_GENERAL: amplitude sqrt(E(k * 2 pi / spacing)),
 * E(kh = k lBendover) = kh^q / (1 + kh^2)^((s + q)/2 + 1), 2 draws per mode).  bk holds
 * 3 components x n x n x (n/2 + 1) complex floats (re, im); returns the number of modes inside
 * [spacing/lMax, spacing/lMin]. */
enum { ORC_TURB_SIMPLE = 0, ORC_TURB_GENERAL = 1, ORC_TURB_GENERAL_POWERLAW = 2 };
/* ORC_TURB_GENERAL_POWERLAW: GridTurbulence with a SimpleTurbulenceSpectrum, E(k) = k^(-s-2) */
size_t orc_turbulence_coeffs(int n, double spacing, double lmin, double lmax, int kind, double alpha, double lbendover,
                             double sindex, double qindex, uint32_t seed, float *bk);

/* DiffusionSDE (src/module/DiffusionSDE.cpp): stochastic transport along the field lines.
 * rng ORC_SDE_MT: the reference's stream, one MT19937 seeded with `seed` and shared by the particles
 * in index order (each particle's steps before the next one's), glibc libm;
 * rng ORC_SDE_PHILOX: the device's stream, Philox4x32-10 per (particle, step) with this package's
 * correctly rounded cos/sin/pow and double-double log; sde_step[i] counts particle i's steps and
 * sde_stream[i] names its random stream (NULL: the index i). */
enum { ORC_SDE_MT = 0, ORC_SDE_PHILOX = 1 };
typedef struct {
	double tolerance, minStep, maxStep, epsilon, alpha, scale;
	int advection;
	double adv[3];              /* UniformAdvectionField */
	double maxTrajectoryLength; /* MaximumTrajectoryLength after the step (<= 0: none) */
	int rng;
	uint32_t seed;
	/* advection kind when advection != 0: 1 uniform (adv), 2 ConstantSphericalAdvectionField (origin,
	 * vWind = par[0]), 3 SphericalAdvectionField (origin, radius, vMax, tau, alpha = par[0..3]);
	 * src/advectionField/AdvectionField.cpp:37-130 */
	int adv_kind;
	double adv_origin[3], adv_par[4];
	/* the advection field as a list of terms (crp_sde_advection_kind, parameters in the C ABI's order),
	 * their velocities summed in order from 0 (AdvectionFieldList::getField,
	 * src/advectionField/AdvectionField.cpp:12-17); adv_n = 0: the single field above */
	int adv_n;
	int adv_list_kind[8];
	double adv_list_par[8][9];
} orc_sde;
/* the advection velocity of p's term list (adv_n terms, summed from 0) at n positions */
void orc_adv_eval(const orc_sde *p, const double *pos, size_t n, double *v, int libm);
long long orc_sde_propagate(const orc_field *f, const orc_sde *p, orc_state *s, size_t n, long long maxSteps,
                            uint32_t *sde_step, const uint64_t *sde_stream);
/* ModuleList::run with OpenMP: one MT19937 per thread (seed + thread number), glibc libm (CPU baseline) */
long long orc_sde_propagate_threads(const orc_field *f, const orc_sde *p, orc_state *s, size_t n, long long maxSteps,
                                    int nthreads);

/* crp_atan2_quick / crp_atan2_accurate per argument pair (tests) */
void orc_atan2_phases(const double *y, const double *x, size_t n, int *decided, double *quick, double *accurate);
/* crp_exp_quick / crp_exp_accurate per argument (tests) */
void orc_exp_phases(const double *x, size_t n, int *decided, double *quick, double *accurate);
/* crp_sincos_fast and crp_log_fast of x[i] (tests) */
void orc_fast_batch(const double *x, size_t n, double *sin_out, double *cos_out, double *log_out);
/* correctly rounded exp(x[i]) and atan2(y[i], x[i]) of crp_crmath.h (tests) */
void orc_exp_atan2_batch(const double *x, const double *y, size_t n, double *exp_out, double *atan2_out);

/* Philox4x32-10 of one counter and key (known-answer tests); out[4] */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* correctly rounded sin and the rounded double-double log of this package, per element */
void orc_sin_log_batch(const double *x, size_t n, double *sin_out, double *log_out);

/* the two phases of the correctly rounded cos separately: quick[i] valid where decided[i] != 0 */
void orc_cos_phases(const double *x, size_t n, double *quick, int32_t *decided, double *accurate);

/* batch helpers for the tests */
void orc_rng_randint_batch(uint32_t seed, size_t n, uint32_t *out);
void orc_noise_grid(uint32_t seed, int nx, int ny, int nz, double amp, float *out);
void orc_sample_particles(uint32_t seed, size_t n, double emin, double emax, const double *weights, size_t nw,
                          int32_t *species, double *energy, double *dir);

/* fields: MagneticField::getField(position) at m positions (AoS) */
void orc_field_eval(const orc_field *f, const double *pos, size_t m, double *B, int libm);
/* periodicBoundary / reflectiveBoundary (include/crpropa/Grid.h:36-46): the neighbour index of
 * tricubic interpolation continued periodically or by mirroring */
int orc_periodic_boundary(int index, int n);
int orc_reflective_boundary(int index, int n);
/* MagneticField::getField(position, z), the call the propagators make, with z[i] per position */
void orc_field_eval_z(const orc_field *f, const double *pos, const double *z, size_t m, double *B, int libm);
void orc_planewave_init(double brms, double lmin, double lmax, double lbendover, double sindex, double qindex,
                        int simple_spectrum, int nm, int seed, double *k, double *Ak, double *xi, double *kappa,
                        double *beta);

/* single-step functions (op 0 BP dY, 1 BP tryStep, 2 CK dYdt, 3 CK tryStep) on n phase points */
void orc_step_batch(const orc_field *f, int op, size_t n, const double *x, const double *u, const double *h,
                    const double *z, const double *q, const double *m, double *ox, double *ou, double *ex, double *eu,
                    int libm);

/* propagation: maxSteps process() calls per particle (stopping early when
 * inactive), OpenMP over particles with nthreads (0: default). Returns the
 * total number of accepted steps and writes rejected trial counts to *trials. */
long long orc_propagate(const orc_field *f, const orc_propagator *p, orc_state *s, size_t n, long long maxSteps,
                        int libm, int nthreads, long long *trials);

/* FP64 reference gradient of the differentiable fixed-step Boris model (the
 * model of the GPU adjoint: PropagationBP::dY with B = scale * FP64 trilinear of
 * a periodic float32 grid, both normalisations).  Stores every state (no
 * checkpoints), serial.  x0/u0/x_out/u_out/g* are AoS (n,3); kq[i] = q c h/(2E).
 * grad_grid (3*nx*ny*nz) and *grad_scale are accumulated. */
void orc_adjoint_forward(const float *grid, int nx, int ny, int nz, const double origin[3], const double spacing[3],
                         double scale, double h, int nsteps, size_t n, const double *x0, const double *u0,
                         const double *kq, double *x_out, double *u_out);
void orc_adjoint_backward(const float *grid, int nx, int ny, int nz, const double origin[3], const double spacing[3],
                          double scale, double h, int nsteps, size_t n, const double *x0, const double *u0,
                          const double *kq, const double *gx_out, const double *gu_out, double *grad_grid,
                          double *grad_x0, double *grad_u0, double *grad_scale);

/* elementary functions, for the math tests */
double orc_cos(double x, int libm);
double orc_pow(double x, double y, int libm);
void orc_cos_batch(const double *x, size_t n, double *out, int libm);
void orc_pow_batch(const double *x, double y, size_t n, double *out, int libm);

#ifdef __cplusplus
}
#endif
#endif
