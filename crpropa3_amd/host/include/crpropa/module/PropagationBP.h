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

struct crp_propagator;

namespace crpropa {

/* Shared implementation of the two GPU propagators. */
class DevicePropagator : public Module {
public:
	enum Kind { BP = 0, CK = 1 };
	enum Precision { FP64 = 0, MIXED = 1 };
	DevicePropagator(Kind kind) : kind(kind), tolerance(0), minStep(0), maxStep(0), precision(FP64) {}

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

	/* Creates the device propagator for the current parameters (caller destroys).
	 * maxLength > 0 fuses MaximumTrajectoryLength. Throws std::runtime_error. */
	crp_propagator *makeDevicePropagator(double maxLength) const;

protected:
	const char *name() const { return kind == BP ? "PropagationBP" : "PropagationCK"; }
	Kind kind;
	ref_ptr<MagneticField> field;
	double tolerance, minStep, maxStep;
	Precision precision;
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
