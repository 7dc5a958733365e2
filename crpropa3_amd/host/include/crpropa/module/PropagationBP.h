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

class PropagationBP : public DevicePropagator {
public:
	/* fixed step: tolerance 0.42, minStep = maxStep = fixedStep */
	PropagationBP(ref_ptr<MagneticField> field = NULL, double fixedStep = 1. * kpc);
	PropagationBP(ref_ptr<MagneticField> field, double tolerance, double minStep, double maxStep);
	std::string getDescription() const;
};

} // namespace crpropa

#endif
