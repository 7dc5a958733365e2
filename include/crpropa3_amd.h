/*
 * crpropa3_amd -- C ABI of the MI355X batched charged-particle propagator.
 *
 * Plain C types only (pointers, sizes, doubles); no C++ exception crosses this
 * boundary.  Every function returns CRP_OK (0) or a negative error code; the
 * message of the last error on the calling thread is crp_last_error().
 *
 * Units are SI, as in the reference (include/crpropa/Units.h): metres, joules,
 * tesla, coulomb, seconds.  Host pointers are read or written synchronously
 * unless a function says it takes device pointers.
 *
 * Each entry point names the reference interface it stands in for.
 */
#ifndef CRPROPA3_AMD_H
#define CRPROPA3_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CRP_ABI_VERSION 12

enum crp_status {
	CRP_OK = 0,
	CRP_E_INVALID = -1,  /* bad argument (the reference throws std::runtime_error) */
	CRP_E_HIP = -2,      /* HIP runtime error */
	CRP_E_NOMEM = -3,    /* allocation failed */
	CRP_E_STATE = -4,    /* object in the wrong state for the call */
	CRP_E_UNBOUNDED = -5 /* run would never end (no step budget, no length limit) */
};

/* Arithmetic of a propagation run. */
enum crp_precision {
	CRP_PRECISION_FP64 = 0,  /* reference arithmetic, bit for bit (FP32 trilinear as Grid3f does) */
	CRP_PRECISION_MIXED = 1  /* FP32 field + rotation, FP64 position and step control */
};

typedef struct crp_field crp_field_t;
typedef struct crp_propagator crp_propagator_t;
typedef struct crp_batch crp_batch_t;

/* Counters of one crp_propagate call. */
typedef struct {
	unsigned long long accepted_steps; /* process() calls that moved a particle */
	unsigned long long rejected_trials; /* adaptive trials thrown away (step shrunk) */
	unsigned long long field_evals;     /* getField evaluations of the reference's algorithm (a rejected
	                                       Cash-Karp trial's retry reads its first-stage value from a register,
	                                       not from the grid, and is counted as the reference evaluates it) */
	unsigned long long finished;        /* particles made inactive by the length limit */
	double kernel_ms;                   /* device time of the propagation kernels (HIP events) */
	int launches;                       /* kernel launches */
} crp_run_stats;

/* ---- runtime ---------------------------------------------------------- */
const char *crp_last_error(void);
int crp_abi_version(void);
/* Select the HIP device for objects created afterwards on this thread. */
int crp_set_device(int device);
int crp_get_device_count(int *count);
/* Wait for all work on the current device (hipDeviceSynchronize). */
int crp_device_synchronize(void);

/* ---- fields (MagneticField::getField, include/crpropa/magneticField/MagneticField.h:22-32) */

/* UniformMagneticField(B) -- include/crpropa/magneticField/MagneticField.h (UniformMagneticField) */
int crp_field_create_uniform(const double B[3], crp_field_t **out);

/* MagneticFieldGrid(ref_ptr<Grid3f>) -- src/magneticField/MagneticFieldGrid.cpp:5-19 with
 * Grid<Vector3f> (include/crpropa/Grid.h:158-582).  values: host float32, AoS (x,y,z) per point,
 * index (ix*ny + iy)*nz + iz (Grid.h:338-340).  origin/spacing as GridProperties; reflective,
 * clip_volume as Grid::setReflective / setClipVolume.  Trilinear interpolation only. */
int crp_field_create_grid3f(const float *values, int nx, int ny, int nz, const double origin[3],
                            const double spacing[3], int reflective, int clip_volume, crp_field_t **out);

/* Same, with values already on the device (borrowed: the caller keeps them alive). */
int crp_field_create_grid3f_device(const float *d_values, int nx, int ny, int nz, const double origin[3],
                                   const double spacing[3], int reflective, int clip_volume, crp_field_t **out);

/* ModulatedMagneticFieldGrid (src/magneticField/MagneticFieldGrid.cpp:20-55): the Grid3f above times a
 * Grid1f `modulation` (one float per point, index (ix * mny + iy) * mnz + iz) with its own origin,
 * spacing, size, reflective and clipVolume flags; B = Grid3f(pos) * Grid1f(pos), each interpolated as
 * the reference does.  (The reference's constructor makes the Grid3f periodic and the Grid1f
 * reflective; the flags here are taken as given.) */
int crp_field_create_modulated_grid3f(const float *values, int nx, int ny, int nz, const double origin[3],
                                      const double spacing[3], int reflective, int clip_volume, const float *modulation,
                                      int mnx, int mny, int mnz, const double morigin[3], const double mspacing[3],
                                      int mreflective, int mclip_volume, crp_field_t **out);

/* PlaneWaveTurbulence::getField -- src/magneticField/turbulentField/PlaneWaveTurbulence.cpp:231-240.
 * The mode table as built by the PlaneWaveTurbulence constructor (:99-178): k[nm], Ak[nm],
 * xi[3*nm], kappa[3*nm] (AoS), beta[nm]; host pointers. */
int crp_field_create_planewave(int nm, const double *k, const double *Ak, const double *xi, const double *kappa,
                               const double *beta, crp_field_t **out);

/* Device layout of a grid field.  POINTS reads the Grid3f values where they are (8 gathers of
 * 12 bytes per evaluation); CELLS (the default for periodic grids) adds a cell-major table with
 * the 8 corner values of every cell in one 128-byte line (32 floats per cell of device memory).
 * Both give identical results.  No reference counterpart: a device storage choice. */
enum crp_grid_layout { CRP_GRID_LAYOUT_POINTS = 0, CRP_GRID_LAYOUT_CELLS = 1 };
int crp_field_set_grid_layout(crp_field_t *f, int layout);
int crp_field_get_grid_layout(const crp_field_t *f, int *layout);

/* Reduced grid storage for the mixed-precision path: FP16_BRICK adds a cell table of FP16 corner
 * values normalised by the FP32 scale of each 8^3-point brick (64 bytes per cell: 48 B values +
 * 4 B scale); mixed-precision propagators then read it instead of the FP32 values.  The FP64 path
 * never uses it.  FP32 frees it.  No reference counterpart (the reference stores FP32 only). */
enum crp_grid_storage { CRP_GRID_STORAGE_FP32 = 0, CRP_GRID_STORAGE_FP16_BRICK = 1 };
int crp_field_set_grid_storage(crp_field_t *f, int storage);

/* Grid::setInterpolationType (include/crpropa/Grid.h:24-28,255-264) and the dispatch of
 * Grid::interpolate (Grid.h:318-335) for one grid of a field:
 *   TRILINEAR          trilinearInterpolate (Grid.h:540-582), the default;
 *   TRICUBIC           Grid3f: tricubicInterpolate(Vector3f) with CubicInterpolate (Grid.h:408-492), the
 *                      Catmull-Rom cubic along z, then y, then x over the 4x4x4 neighbouring points in
 *                      FP32, in the operation order of the reference's SIMD build (SIMD_EXTENSIONS "avx";
 *                      its "none" build rejects tricubic Grid3f); Grid1f: tricubicInterpolate(double)
 *                      (Grid.h:499-536) in FP64, the result narrowed to float;
 *   NEAREST_NEIGHBOUR  closestValue (Grid.h:383-405).
 * `grid` selects the grid: CRP_GRID_VALUES the Grid3f of a grid or modulated field, or the turbulent
 * Grid3f of a JF12 field (JF12Field.cpp:340); CRP_GRID_MODULATION the Grid1f of a modulated field.
 * (JF12's striated Grid1f is always read with closestValue, as src/magneticField/JF12Field.cpp:300-301
 * does.)  A non-trilinear Grid3f
 * is read from its points: its cell table and FP16 storage are released, and
 * crp_field_set_grid_layout(CELLS) / crp_field_set_grid_storage(FP16_BRICK) fail until it is
 * trilinear again (a periodic grid then gets its cell table back).  The mixed-precision path reads
 * non-trilinear grids with the FP64 path's evaluation, whose result is FP32 already.  (The adjoint,
 * crp_adjoint_*, differentiates trilinear interpolation only.)  Unknown types fail with
 * CRP_E_INVALID, as Grid::setInterpolationType throws. */
enum crp_grid_interpolation { CRP_GRID_TRILINEAR = 0, CRP_GRID_TRICUBIC = 1, CRP_GRID_NEAREST_NEIGHBOUR = 2 };
enum crp_grid_select { CRP_GRID_VALUES = 0, CRP_GRID_MODULATION = 1 };
int crp_field_set_grid_interpolation(crp_field_t *f, int grid, int type);
int crp_field_get_grid_interpolation(const crp_field_t *f, int grid, int *type);

/* B at n host positions with the given precision's field evaluation (MIXED: the FP32 evaluation
 * of the mixed path, FP16 storage if set). */
int crp_field_eval_ex(const crp_field_t *f, const double *pos, size_t n, double *B, int precision);

int crp_field_destroy(crp_field_t *f);

/* B at n host positions (AoS xyz), evaluated on the device with the FP64 path. */
int crp_field_eval(const crp_field_t *f, const double *pos, size_t n, double *B);

/* Field decorators (src/magneticField/MagneticField.cpp:41-79, include/crpropa/magneticField/MagneticField.h:47-103).
 * They borrow the fields they wrap: those must outlive the decorator.  Decorators nest (a list of
 * periodic fields, a periodic list, ...) up to 3 levels of lists or periodic maps, at most 8 fields
 * below a tree; deeper trees are rejected with CRP_E_INVALID.
 *   PeriodicMagneticField(field, extends, origin, reflective): origin may be NULL (0, 0, 0).
 *   MagneticFieldList() + addField(field): crp_field_create_list, then crp_field_list_add in order
 *     (at most 8 fields per list).
 *   MagneticFieldEvolution(field, m): getField(position, z) = field->getField(position) * (1+z)^m;
 *     its one-argument getField (crp_field_eval) is zero, as in the reference; the propagators use
 *     the candidate's redshift (crp_host_state.redshift), and pow is correctly rounded. */
int crp_field_create_periodic(const crp_field_t *field, const double extends[3], const double origin[3], int reflective,
                              crp_field_t **out);
int crp_field_create_list(crp_field_t **out);
int crp_field_list_add(crp_field_t *list, const crp_field_t *field);
int crp_field_create_evolution(const crp_field_t *field, double m, crp_field_t **out);

/* MagneticField::getField(position, z) at n host positions with their redshifts z[n] (FP64): the call
 * PropagationBP/CK::getFieldAtPosition(pos, z) makes (src/module/PropagationBP.cpp:137-149).  Only a
 * top-level MagneticFieldEvolution reads z; for every other field this is crp_field_eval. */
int crp_field_eval_z(const crp_field_t *f, const double *pos, const double *z, size_t n, double *B);

/* crp_field_eval_ex without any device allocation: positions and results are staged through a
 * fixed per-device buffer of the library (CRP_EVAL_NOALLOC_MAX positions, 6 MiB, held until the
 * process ends) instead of hipMalloc'd buffers; same kernels, same bits.  n above
 * CRP_EVAL_NOALLOC_MAX is rejected with CRP_E_INVALID.  Calls that overlap on one GPU (several host
 * threads) share that buffer and race.  No counterpart in the reference: for callers that cannot
 * allocate device memory (per-candidate getField calls in tight loops, harnesses that forbid it). */
#define CRP_EVAL_NOALLOC_MAX 131072
int crp_field_eval_noalloc(const crp_field_t *f, const double *pos, size_t n, double *B, int precision);

/* JF12Field / PlanckJF12bField (src/magneticField/JF12Field.cpp): the model's constants as the
 * reference's constructors set them (JF12Field.cpp:8-79, 348-363) and its switches. */
typedef struct {
	int use_regular, use_disk, use_toroidal_halo, use_x, use_striated, use_turbulent;
	double r_arms[8], b_disk[8], sin_pitch, cos_pitch, tan90_minus_pitch;
	double b_ring, h_disk, w_disk, b_north, b_south, r_north, r_south, w_halo, z0;
	double b_x, sin_theta_x0, cos_theta_x0, tan_theta_x0, r_xc, r_x, sqrt_beta;
	double b_disk_turb[8], b_disk_turb5, z_disk_turb, b_halo_turb, r_halo_turb, z_halo_turb;
} crp_jf12_params;
/* the constructor's values (planck = 0: JF12Field, 1: PlanckJF12bField), computed with the host's
 * libm like the reference's constructor; regular field on, striated and turbulent off */
int crp_jf12_default_params(int planck, crp_jf12_params *out);
/* JF12Field::getField (JF12Field.cpp:230-346) on the device: the regular field, times
 * (1 + sqrt(beta) x the striated Grid1f's closest value) when use_striated, plus the turbulent
 * Grid3f's trilinear value times the turbulent strength when use_turbulent.  The grids are host
 * arrays (one float per point; (x,y,z) floats per point), NULL when unused. */
int crp_field_create_jf12(const crp_jf12_params *p, const float *striated, const int striated_n[3],
                          const double striated_origin[3], const double striated_spacing[3], int striated_reflective,
                          const float *turbulent, const int turbulent_n[3], const double turbulent_origin[3],
                          const double turbulent_spacing[3], int turbulent_reflective, int turbulent_clip,
                          crp_field_t **out);
/* JF12Field::getRegularField, getStriatedField, getTurbulentField (JF12Field.cpp:198-341): the same
 * model with other switches p over base's device grids (borrowed: base must outlive the view; no grid
 * is copied).  A switch whose grid base does not hold is rejected with CRP_E_INVALID. */
int crp_field_create_jf12_view(const crp_field_t *base, const crp_jf12_params *p, crp_field_t **out);

/* ---- propagators ------------------------------------------------------ */

/* PropagationBP(field, tolerance, minStep, maxStep) -- src/module/PropagationBP.cpp:43-59,162-183.
 * Fixed-step form PropagationBP(field, step) = (field, 0.42, step, step).  field may be NULL
 * (B = 0, src/module/PropagationBP.cpp:137-149).  Validation as setTolerance / setMinimumStep /
 * setMaximumStep: tolerance in [0,1], 0 <= minStep <= maxStep. */
int crp_propagator_create_bp(const crp_field_t *field, double tolerance, double min_step, double max_step,
                             crp_propagator_t **out);

/* PropagationCK(field, tolerance, minStep, maxStep) -- src/module/PropagationCK.cpp:63-75,160-179. */
int crp_propagator_create_ck(const crp_field_t *field, double tolerance, double min_step, double max_step,
                             crp_propagator_t **out);

/* DiffusionSDE(field, tolerance, minStep, maxStep, epsilon) -- src/module/DiffusionSDE.cpp:21-46,
 * 272-315: stochastic transport along the field lines, alpha = 1/3 and scale = 1 until
 * crp_propagator_set_sde.  Validation and messages as the reference's setters (tolerance and
 * epsilon in [0,1], 0 <= minStep <= maxStep).  Runs in FP64 (crp_propagator_set_precision
 * refuses CRP_PRECISION_MIXED); the fused modules (length limit, periodic box, spherical boundary,
 * observer, module order) apply after each DiffusionSDE step as after a BP/CK step.  Random
 * numbers: Philox4x32-10 with key (seed, 0x43525053) and counter (stream low word, step, block,
 * stream high word) per particle (crp_host_state.sde_step / sde_stream).  A step whose trials
 * would need more than 2^20 sub-steps (minStep far below the field's scale) stops that particle
 * where it is (inactive; its step count advances) and crp_propagate returns CRP_E_STATE after the
 * run, with every other particle's progress in the batch.  The environment variable
 * CRP_SDE_MAX_HALVINGS (1..20), read at each call, lowers the limit to 2^value sub-steps (tests). */
int crp_propagator_create_sde(const crp_field_t *field, double tolerance, double min_step, double max_step,
                              double epsilon, crp_propagator_t **out);
/* setAlpha / setScale / setEpsilon (src/module/DiffusionSDE.cpp:294-314) and the random seed */
int crp_propagator_set_sde(crp_propagator_t *p, double alpha, double scale, double epsilon, uint32_t seed);
/* UniformAdvectionField(v) of DiffusionSDE(field, advectionField, ...) -- src/advectionField/
 * AdvectionField.cpp:37-50; v == NULL removes it */
int crp_propagator_set_sde_advection(crp_propagator_t *p, const double v[3]);
/* ConstantSphericalAdvectionField(origin, vWind): v = vWind (x - origin) / |x - origin|
 * (src/advectionField/AdvectionField.cpp:52-56) */
int crp_propagator_set_sde_constant_spherical_advection(crp_propagator_t *p, const double origin[3], double v_wind);
/* SphericalAdvectionField(origin, radius, vMax, tau, alpha): v = vMax (1 - exp(-R^alpha / tau)) along
 * x - origin inside the radius, 0 outside (AdvectionField.cpp:90-120) */
int crp_propagator_set_sde_spherical_advection(crp_propagator_t *p, const double origin[3], double radius,
                                               double v_max, double tau, double alpha);
/* Any advection field of src/advectionField/AdvectionField.cpp as a list of terms whose velocities are
 * summed in order from 0 (AdvectionFieldList::getField, :12-17; a single term is the field itself).
 * crp_propagator_add_sde_advection appends one term (at most CRP_SDE_ADVECTION_MAX); the setters above
 * replace the list with one term; crp_propagator_clear_sde_advection empties it.  Parameters per kind:
 *   UNIFORM              v[3]                                         (:29-39)
 *   CONSTANT_SPHERICAL   origin[3], vWind                             (:49-62)
 *   SPHERICAL            origin[3], radius, vMax, tau, alpha          (:93-123)
 *   CARTESIAN_SHOCK      compressionRatio, vUp, lShock  (OneDimensionalCartesianShock, :182-199)
 *   SPHERICAL_SHOCK_1D   rShock, vUp, compressionRatio, lShock, coolUpstream (0/1)
 *                        (OneDimensionalSphericalShock, :246-275)
 *   OBLIQUE_SHOCK        compressionRatio, vXUp, vY, lShock  (ObliqueAdvectionShock, :353-372)
 *   SPHERICAL_SHOCK      origin[3], r0, v0, lambda, rRot, vPhi  (SphericalAdvectionShock with setRRot /
 *                        setAzimuthalSpeed, :432-466)
 * Deterministic forms where the reference calls glibc: tanh from the correctly rounded exp (not correctly
 * rounded itself), SphericalAdvectionShock's e_phi = (-sin phi, cos phi, 0), phi = atan2(y, x), as
 * (-y, x, 0) / sqrt(x^2 + y^2); exp and pow are correctly rounded. */
enum crp_sde_advection_kind {
	CRP_ADV_UNIFORM = 1,
	CRP_ADV_CONSTANT_SPHERICAL = 2,
	CRP_ADV_SPHERICAL = 3,
	CRP_ADV_CARTESIAN_SHOCK = 4,
	CRP_ADV_SPHERICAL_SHOCK_1D = 5,
	CRP_ADV_OBLIQUE_SHOCK = 6,
	CRP_ADV_SPHERICAL_SHOCK = 7
};
#define CRP_SDE_ADVECTION_MAX 8
int crp_propagator_add_sde_advection(crp_propagator_t *p, int kind, const double *par, int npar);
int crp_propagator_clear_sde_advection(crp_propagator_t *p);

/* MaximumTrajectoryLength(maxLength) fused after each step -- src/module/BreakCondition.cpp:42-64
 * (no observer positions).  max_length <= 0 removes it. */
int crp_propagator_set_max_trajectory_length(crp_propagator_t *p, double max_length);
int crp_propagator_set_precision(crp_propagator_t *p, int precision);
int crp_propagator_destroy(crp_propagator_t *p);

/* The reference's public per-step functions, on the device with the propagator's field (FP64, the
 * kernels' arithmetic), for n phase points; positions/directions/outputs are AoS [3n], the other
 * arrays [n]; z may be NULL (0).  For callers of the single-step API of
 *   PropagationBP::dY(pos, dir, step, z, q, m)              include/crpropa/module/PropagationBP.h:89
 *   PropagationBP::tryStep(y, out, error, h, p, z, q, m)     include/crpropa/module/PropagationBP.h:116
 *     (error = Y(errorEstimation(out.x, x after two h/2 steps, h)): all six components equal S)
 *   PropagationCK::dYdt(y, p, z) -> (velocity, du/dt)        include/crpropa/module/PropagationCK.h:74
 *   PropagationCK::tryStep(y, out, error, t, p, z)           include/crpropa/module/PropagationCK.h:76
 * (src/module/PropagationBP.cpp:8-39, src/module/PropagationCK.cpp:28-61).  m is the mass E/c^2 (BP),
 * E the energy (CK); t = step / c. */
int crp_bp_dy(const crp_propagator_t *p, size_t n, const double *pos, const double *dir, const double *step,
              const double *z, const double *q, const double *m, double *out_pos, double *out_dir);
int crp_bp_try_step(const crp_propagator_t *p, size_t n, const double *pos, const double *dir, const double *h,
                    const double *z, const double *q, const double *m, double *out_pos, double *out_dir,
                    double *err_pos, double *err_dir);
int crp_ck_dydt(const crp_propagator_t *p, size_t n, const double *x, const double *u, const double *z,
                const double *q, const double *E, double *dxdt, double *dudt);
int crp_ck_try_step(const crp_propagator_t *p, size_t n, const double *x, const double *u, const double *t,
                    const double *z, const double *q, const double *E, double *out_x, double *out_u, double *err_x,
                    double *err_u);

/* ---- particle batches: device-resident structure of arrays ------------ */
/* The fields of Candidate/ParticleState that the path reads or writes
 * (include/crpropa/Candidate.h, include/crpropa/ParticleState.h). */
typedef struct {
	double *pos;               /* [3n] current.position (AoS on the host side) */
	double *dir;               /* [3n] current.direction */
	double *energy;            /* [n] current.energy */
	double *charge;            /* [n] current.charge (Z * eplus) */
	double *current_step;      /* [n] */
	double *next_step;         /* [n] */
	double *trajectory_length; /* [n] */
	double *time;              /* [n] */
	int32_t *active;           /* [n] */
	double *prev_pos;          /* [3n] previous.position, written on download (may be NULL) */
	double *prev_dir;          /* [3n] previous.direction (may be NULL) */
	int32_t *detection_index;  /* [n] observer detections so far, the candidate property
	                              "DetectionIndex" of src/module/Observer.cpp:282-290 (may be NULL: 0) */
	double *source_pos;        /* [3n] source.position, moved by a periodic box (may be NULL: the
	                              current position on upload, not written on download) */
	double *created_pos;       /* [3n] created.position, likewise */
	uint32_t *sde_step;        /* [n] DiffusionSDE steps taken, the particle's random-number counter
	                              (may be NULL: 0 on upload) */
	const uint64_t *sde_stream; /* [n] DiffusionSDE random stream of each particle, e.g. its serial
	                              number (upload only; may be NULL: the particle's index) */
	const double *redshift;    /* [n] candidate redshift, the z of getField(position, z) (upload only;
	                              may be NULL: 0) */
} crp_host_state;

int crp_batch_create(size_t capacity, crp_batch_t **out);
int crp_batch_destroy(crp_batch_t *b);
int crp_batch_size(const crp_batch_t *b, size_t *n);
/* Copy n particles from host arrays to the device (pos/dir AoS are transposed to SoA). */
int crp_batch_upload(crp_batch_t *b, size_t n, const crp_host_state *s);
int crp_batch_download(const crp_batch_t *b, crp_host_state *s);
/* Device-to-device copy of the whole state (dst gets src's size). */
int crp_batch_copy(crp_batch_t *dst, const crp_batch_t *src);

/* ModuleList::run(candidates) for [propagator, MaximumTrajectoryLength] (src/ModuleList.cpp.in:97-151):
 * at most max_steps process() calls per active particle (max_steps < 0: until every particle is
 * inactive, which needs a trajectory-length limit and minStep > 0).  Synchronous. */
int crp_propagate(const crp_propagator_t *p, crp_batch_t *b, long long max_steps, crp_run_stats *stats);

/* PeriodicBox(origin, size) fused right after the propagator (src/module/Boundary.cpp:16-27): after
 * each accepted step n = floor((pos - origin) / size) per axis and, if n != 0, the current, previous,
 * source and created positions move by -n * size.  size all 0 removes it. */
int crp_propagator_set_periodic_box(crp_propagator_t *p, const double origin[3], const double size[3]);

/* SphericalBoundary(center, radius) fused after the box (src/module/Boundary.cpp:208-223): after each
 * accepted step d = |pos - center|; d >= radius rejects the particle (inactive); with limit_step the
 * next step is limited to radius - d + margin (the reference's defaults: limit_step 1, margin 0.1 kpc).
 * center NULL removes it. */
int crp_propagator_set_spherical_boundary(crp_propagator_t *p, const double center[3], double radius,
                                          double margin, int limit_step);

/* The order of the fused modules after the propagator, as in the ModuleList (src/ModuleList.cpp.in:54-58):
 * a permutation of the configured modules, CRP_MODULE_* values.  Positions and rejections do not depend
 * on it; an observer's detections see nextStep limited only by the boundary and length limit listed
 * before it.  The periodic box, if listed, must come first.  n = 0: box, sphere, length limit, observer. */
enum crp_module {
	CRP_MODULE_PERIODIC_BOX = 1,
	CRP_MODULE_SPHERICAL_BOUNDARY = 2,
	CRP_MODULE_MAX_TRAJECTORY_LENGTH = 3,
	CRP_MODULE_OBSERVER = 4
};
int crp_propagator_set_module_order(crp_propagator_t *p, const int *modules, int n);

/* Observer with one ObserverTimeEvolution, fused after the propagator and the length limit
 * (src/module/Observer.cpp:26-57, 266-306): after each accepted step a particle whose trajectory
 * length reached times[k] (k = its detections so far) is detected (at most once per step), its state
 * is recorded as detection k, and nextStep is limited towards times[k+1] (not for the last two
 * times, as the reference); otherwise nextStep is limited to land on times[k].  deactivate != 0
 * deactivates on detection (Observer::setDeactivateOnDetection, default true in the reference).
 * ntimes = 0 removes the observer. */
int crp_propagator_set_observer(crp_propagator_t *p, const double *times, int ntimes, int deactivate);
/* Recorded detections: out[(k * n + i) * 9 + c] for k < ntimes, c = position x,y,z, direction x,y,z,
 * trajectory length, time, nextStep after the observer; NaN where particle i has fewer than k+1
 * detections. */
int crp_batch_detections(const crp_batch_t *b, int ntimes, double *out);

/* ---- launch configuration, hipGraph replay and autotuning ------------------
 * No reference counterpart (the reference loops over candidates with OpenMP,
 * src/ModuleList.cpp.in:116-125).  Results do not depend on any of these
 * settings: every particle's arithmetic is the same whichever lane, launch or
 * graph replay runs it. */
typedef struct {
	int block;                  /* threads per workgroup: 64, 128 or 256 (0: 256) */
	int blocks_per_cu;          /* resident workgroups per CU (0: the occupancy limit) */
	long long steps_per_launch; /* split a call into launches of at most this many process() calls per
	                               particle (<= 0: one launch) */
	int use_graph;              /* capture the call's launch sequence in a hipGraph and replay it while the
	                               propagator, field, batch size and step budget stay the same */
	double particles_per_lane;  /* output only: batch size / lanes launched, from the last call */
} crp_launch_config;

int crp_propagator_set_launch(crp_propagator_t *p, const crp_launch_config *cfg);
int crp_propagator_get_launch(const crp_propagator_t *p, crp_launch_config *cfg);
/* Time short runs (probe_steps process() calls per particle, on a device copy of b, which is left
 * unchanged) over block sizes, workgroups per CU and steps per launch; keep the fastest as the
 * propagator's launch configuration and return it in *chosen (may be NULL). */
int crp_propagator_autotune(crp_propagator_t *p, const crp_batch_t *b, long long probe_steps,
                            crp_launch_config *chosen);

/* ---- adjoint of the fixed-step Boris push through a periodic Grid3f -------
 * No reference counterpart: the reference has no derivatives.  The forward model is
 * PropagationBP::dY (src/module/PropagationBP.cpp:19-39) with both direction normalisations,
 * B = scale * trilinear(grid) evaluated in FP64 (a differentiable model of Grid.h:540-582).
 * All pointers are device pointers; work is queued on `stream` (a hipStream_t, NULL = default)
 * and the calls do not synchronise.  Positions/directions are AoS (n,3) FP64; kq[i] = q c h / (2 E). */
typedef struct {
	const float *grid;  /* device, (ix*ny + iy)*nz + iz, 3 floats per point */
	int nx, ny, nz;
	double origin[3], spacing[3];
	double scale;       /* field = scale * grid */
	double step;        /* h */
	int nsteps;
	int checkpoint_every; /* K */
} crp_adjoint_config;

/* Device workspace sizes: checkpoints ((ceil(nsteps/K)) * 6 * n doubles, then, for grids of at most
 * 8 GiB of cells, a table of the 8 corner values of every cell in one 128-byte line that the forward
 * builds and the backward reads) and the backward's scratch.
 * For grids of at most 8192 tiles of 8^3 cells (or of 16^3 cells, up to about 320^3 points) the
 * backward recomputes several segments of every
 * particle at once, keeping the state every 8 steps, and sorts one deposit record per step by tile
 * before accumulating the grid gradient tile by tile in LDS; its scratch (two buffer sets of at most
 * 4 GiB each, less for small batches) holds those states and records.  Larger grids use per-step atomics and K * 6 * n doubles of scratch.  The
 * environment variable CRP_ADJOINT_TILE_SHIFT=4 selects 16^3-cell tiles where 8^3 would do (tests).  The
 * environment variable CRP_ADJOINT_BLOCK_BYTES, read at each call, overrides the per-block budget
 * in bytes (0 selects the per-step atomics). */
int crp_adjoint_workspace(const crp_adjoint_config *cfg, size_t n, size_t *checkpoint_bytes, size_t *scratch_bytes);
/* Forward: final positions/directions and checkpoints every K steps. */
int crp_adjoint_forward(const crp_adjoint_config *cfg, size_t n, const double *x0, const double *u0,
                        const double *kq, double *checkpoints, double *x_out, double *u_out, void *stream);
/* Backward: given dL/dx_out (may be NULL) and dL/du_out, ADD dL/dgrid (3*nx*ny*nz doubles) and
 * dL/dscale (one double) into the outputs, and WRITE dL/dx0, dL/du0. */
int crp_adjoint_backward(const crp_adjoint_config *cfg, size_t n, const double *kq, const double *checkpoints,
                         const double *grad_x_out, const double *grad_u_out, double *scratch, double *grad_grid,
                         double *grad_x0, double *grad_u0, double *grad_scale, void *stream);

/* ---------------------------------------------------------------- turbulent grids
 * SimpleGridTurbulence::initTurbulence (src/magneticField/turbulentField/SimpleGridTurbulence.cpp:18-113)
 * and GridTurbulence::initTurbulence (src/magneticField/turbulentField/GridTurbulence.cpp:29-124,
 * 147-185) on the device, up to (not including) the final scaleGrid to Brms: random k-space
 * coefficients for every mode with spacing/lMax <= |k| <= spacing/lMin, one in-place
 * complex-to-real FFT per component (hipFFT), Grid3f layout.  The random numbers are the caller's
 * MT19937 Random::randInt() outputs, in the reference's order: 4 per mode for
 * CRP_TURBULENCE_SIMPLE (orientation, randNorm, phase), 2 for the general spectra. */
enum crp_turbulence_kind {
	CRP_TURBULENCE_SIMPLE = 0,          /* amplitude randNorm() * k^(alpha/2) */
	CRP_TURBULENCE_GENERAL = 1,         /* amplitude sqrt(E(k 2pi/spacing)), TurbulenceSpectrum */
	CRP_TURBULENCE_GENERAL_POWERLAW = 2 /* the same with SimpleTurbulenceSpectrum's E(k) = k^(-s-2) */
};
typedef struct {
	int n;                            /* grid points per axis (cubic grid), 2..2048 */
	double spacing;                   /* equal on all axes */
	double lmin, lmax;                /* 2 spacing <= lmin <= lmax <= n spacing */
	int kind;                         /* crp_turbulence_kind */
	double alpha;                     /* SIMPLE: <B^2(k)> ~ k^alpha */
	double lbendover, sindex, qindex; /* GENERAL: TurbulenceSpectrum parameters; POWERLAW: sindex */
} crp_turbulence_desc;

/* Number of modes inside the turbulence range (draws needed: 4 or 2 per mode). */
int crp_turbulence_mode_count(const crp_turbulence_desc *d, uint64_t *modes);
/* The unnormalised grid: grid[3 * ((ix * n + iy) * n + iz) + c], n^3 x 3 floats, host or device
 * memory.  ndraws must equal draws-per-mode x crp_turbulence_mode_count. */
int crp_turbulence_generate(const crp_turbulence_desc *d, const uint32_t *draws, uint64_t ndraws, float *grid);

/* ---------------------------------------------------------------- measurement
 * No counterpart in the reference: the rate of random line reads of this GPU (lines per second),
 * each lane reading bytes_per_line (64, 96 or 128) from a random 128-byte line of a table of
 * table_bytes; the ceiling of the cell-table field evaluation (96 bytes per evaluation). */
int crp_measure_gather(uint64_t table_bytes, int bytes_per_line, double *lines_per_s);

#ifdef __cplusplus
}
#endif
#endif
