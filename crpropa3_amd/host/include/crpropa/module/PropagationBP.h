// PropagationBP: Boris push with fixed or adaptive steps, same constructors,
// defaults, setters and errors as the reference
// (include/crpropa/module/PropagationBP.h, src/module/PropagationBP.cpp:43-208).
// process() advances one candidate by one step on the GPU; ModuleList::run
// advances whole candidate vectors in one batched device run.
#ifndef CRPROPA_AMD_PROPAGATIONBP_H
#define CRPROPA_AMD_PROPAGATIONBP_H

#include "crpropa/Module.h"
#include "crpropa/Units.h"
#include "crpropa/magneticField/MagneticField.h"

#include <string>
#include <vector>

struct crp_propagator;
struct crp_batch;

namespace crpropa {

/* Shared implementation of the GPU propagators (PropagationBP, PropagationCK, DiffusionSDE). */
class DevicePropagator : public Module {
public:
	enum Kind { BP = 0, CK = 1, SDE = 2 };
	enum Precision { FP64 = 0, MIXED = 1 };
	DevicePropagator(Kind kind) : kind(kind), tolerance(0), minStep(0), maxStep(0), precision(FP64) {}
	virtual ~DevicePropagator();

	void process(Candidate *candidate) const;

	void setField(ref_ptr<MagneticField> f) { field = f; }
	ref_ptr<MagneticField> getField() const { return field; }
	void setTolerance(double tol);
	void setMinimumStep(double minStep);
	void setMaximumStep(double maxStep);
	double getTolerance() const { return tolerance; }
	double getMinimumStep() const { return minStep; }
	double getMaximumStep() const { return maxStep; }
	/* FP64 (reference arithmetic) or MIXED (FP32 field and rotation). */
	void setPrecision(Precision p) { precision = p; }
	Precision getPrecision() const { return precision; }
	Kind getKind() const { return kind; }

	/* Launch configuration of the device runs: block size (64, 128, 256), resident workgroups
	 * per CU (0: occupancy limit), process() calls per launch (0: one launch per run) and hipGraph
	 * replay.  Results do not depend on it. */
	void setLaunchConfig(int block, int blocksPerCU = 0, long long stepsPerLaunch = 0, bool useGraph = false);
	/* Tune the launch configuration on the first batched run (probeSteps process() calls per
	 * candidate on a device copy of the batch) and keep it for later runs. */
	void setAutotune(bool on, long long probeSteps = 32);
	bool getAutotune() const { return autotune; }
	/* block, blocks per CU, steps per launch, graph flag, particles per lane of the last run */
	std::vector<double> getLaunchConfig() const;
	/* tune on this batch if autotuning is on and not yet done (used by ModuleList::run) */
	void tuneOn(crp_propagator *p, const crp_batch *b) const;
	void noteParticlesPerLane(double v) const { lastPPL = v; }

	/* Creates the device propagator for the current parameters (caller destroys).
	 * maxLength > 0 fuses MaximumTrajectoryLength. Throws std::runtime_error. */
	crp_propagator *makeDevicePropagator(double maxLength) const;

	/* getFieldAtPosition (src/module/PropagationBP.cpp:137-149, src/module/PropagationCK.cpp:146-158):
	 * field->getField(pos, z), zero without a field; an exception of the field is reported on stderr
	 * and gives zero. */
	Vector3d getFieldAtPosition(Vector3d pos, double z) const;

protected:
	const char *name() const { return kind == BP ? "PropagationBP" : (kind == CK ? "PropagationCK" : "DiffusionSDE"); }
	Kind kind;
	ref_ptr<MagneticField> field;
	double tolerance, minStep, maxStep;
	Precision precision;
	int launchBlock = 256, launchBlocksPerCU = 0;
	long long launchStepsPerLaunch = 0;
	bool launchGraph = false;
	bool autotune = false;
	long long autotuneProbe = 32;
	mutable bool tuned = false;
	mutable double lastPPL = 0;
};

class ParticleState;

/* Phase point (position, direction) of the single-step functions (include/crpropa/module/PropagationBP.h:24-46) */
class PhasePoint {
public:
	Vector3d x, u;
	PhasePoint() {}
	PhasePoint(const Vector3d &x, const Vector3d &u) : x(x), u(u) {}
	PhasePoint(double f) : x(Vector3d(f, f, f)), u(Vector3d(f, f, f)) {}
	PhasePoint operator*(double f) const { return PhasePoint(x * f, u * f); }
	PhasePoint &operator+=(const PhasePoint &y) {
		x += y.x;
		u += y.u;
		return *this;
	}
};

class PropagationBP : public DevicePropagator {
public:
	typedef PhasePoint Y;
	/* fixed step: tolerance 0.42, minStep = maxStep = fixedStep */
	PropagationBP(ref_ptr<MagneticField> field = NULL, double fixedStep = 1. * kpc);
	PropagationBP(ref_ptr<MagneticField> field, double tolerance, double minStep, double maxStep);
	std::string getDescription() const;

	/* The reference's single-step functions (include/crpropa/module/PropagationBP.h:89-116), run on the
	 * device with this propagator's field (the kernels' FP64 code): one Boris step; the error
	 * estimate |x1 - x2| / (0.75 step); a full step, two half steps and the error as Y(estimate). */
	Y dY(Vector3d pos, Vector3d dir, double step, double z, double q, double m) const;
	double errorEstimation(const Vector3d x1, const Vector3d x2, double step) const;
	void tryStep(const Y &y, Y &out, Y &error, double h, ParticleState &p, double z, double q, double m) const;
};

} // namespace crpropa

#endif
