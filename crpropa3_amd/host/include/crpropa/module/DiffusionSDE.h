// DiffusionSDE: stochastic transport along the magnetic field lines, same
// constructors, defaults, setters and errors as the reference
// (include/crpropa/module/DiffusionSDE.h:55-64, src/module/DiffusionSDE.cpp:21-46,
// 272-405); runs on the GPU (one device step per process() call, whole candidate
// vectors per ModuleList::run).
//
// Random numbers: the reference draws from the per-thread MT19937 of
// Random::instance(); on the device each candidate has its own Philox4x32-10
// stream, named by its serial number and keyed by the module's seed (drawn from
// Random::instance() at construction, so seeding it first makes runs
// reproducible; setSeed fixes it).  The candidate keeps its step count
// (Candidate::getDiffusionSteps), so the numbers do not depend on how a run is
// split into calls or batches.
//
// A step whose trials would need more than 2^20 sub-steps (minStep far below the
// field's scale, where the reference's loop would run for minutes) stops that
// candidate where it is (inactive); process() and ModuleList::run raise
// std::runtime_error after every candidate's state has been updated.
#ifndef CRPROPA_AMD_DIFFUSIONSDE_H
#define CRPROPA_AMD_DIFFUSIONSDE_H

#include "crpropa/advectionField/AdvectionField.h"
#include "crpropa/module/PropagationBP.h"

#include <cstdint>

namespace crpropa {

class DiffusionSDE : public DevicePropagator {
public:
	DiffusionSDE(ref_ptr<MagneticField> magneticField, double tolerance = 1e-4, double minStep = 10 * pc,
	             double maxStep = 1 * kpc, double epsilon = 0.1);
	DiffusionSDE(ref_ptr<MagneticField> magneticField, ref_ptr<AdvectionField> advectionField,
	             double tolerance = 1e-4, double minStep = 10 * pc, double maxStep = 1 * kpc, double epsilon = 0.1);

	void setEpsilon(double e);
	void setAlpha(double alpha);
	void setScale(double scale);
	void setMagneticField(ref_ptr<MagneticField> f) { setField(f); }
	void setAdvectionField(ref_ptr<AdvectionField> f) { advectionField = f; }
	double getEpsilon() const { return epsilon; }
	double getAlpha() const { return alpha; }
	double getScale() const { return scale; }
	ref_ptr<MagneticField> getMagneticField() const { return getField(); }
	ref_ptr<AdvectionField> getAdvectionField() const { return advectionField; }
	Vector3d getMagneticFieldAtPosition(Vector3d pos, double z) const;
	Vector3d getAdvectionFieldAtPosition(Vector3d pos, double t) const;
	/* seed of the device random streams */
	void setSeed(uint32_t s) { seed = s; }
	uint32_t getSeed() const { return seed; }
	std::string getDescription() const;

	/* the diffusion parameters on a device propagator (used by makeDevicePropagator) */
	void configureDevice(crp_propagator *p) const;

private:
	double epsilon = 0.1, alpha = 1. / 3., scale = 1.;
	ref_ptr<AdvectionField> advectionField;
	uint32_t seed = 0;
};

} // namespace crpropa

#endif
