// Reference-side binding of the GPU propagator: modules with the interface of the reference's own
// Module (include/crpropa/Module.h:26, plugin-template/myPlugin.h:13-19) that hand candidates to
// the C ABI of include/crpropa3_amd.h.  Compiled against the reference's headers and linked with its
// library and libcrpropa3_amd.so (crpropa3_amd/plugin/Makefile).
//
//   ref_ptr<MagneticFieldGrid> field = ...;                      // the reference's field
//   ref_ptr<crpropa_amd::GpuPropagationCK> ck = new crpropa_amd::GpuPropagationCK(field, 1e-4, 1 * kpc, 1 * Mpc);
//   sim->add(ck);                                                 // per-candidate process(), or
//   ck->run(candidates, 100 * Mpc);                               // one batched device run
//
// Fields with a device counterpart: MagneticFieldGrid (Grid3f trilinear, tricubic or nearest
// neighbour; periodic, reflective or clipVolume) and UniformMagneticField; any other field is rejected
// with std::runtime_error.
#ifndef CRPROPA3_AMD_GPUPROPAGATION_H
#define CRPROPA3_AMD_GPUPROPAGATION_H

#include "crpropa/Candidate.h"
#include "crpropa/Module.h"
#include "crpropa/magneticField/MagneticField.h"

#include "crpropa3_amd.h"

#include <mutex>
#include <vector>

namespace crpropa_amd {

/* The device field of a reference field (caller destroys with crp_field_destroy). */
crp_field_t *makeDeviceField(crpropa::ref_ptr<crpropa::MagneticField> field);

class GpuPropagation : public crpropa::Module {
public:
	~GpuPropagation();
	/* PropagationBP/CK::process (src/module/PropagationBP.cpp:62-124, src/module/PropagationCK.cpp:77-136):
	 * one step of this candidate on the device (correct, one launch per call). */
	void process(crpropa::Candidate *candidate) const;
	/* ModuleList::run(candidates) for [this propagator, MaximumTrajectoryLength(maxLength)]
	 * (src/ModuleList.cpp.in:97-151): every active candidate propagated until the length limit in one
	 * batched device run (maxLength > 0), or maxSteps steps (maxLength <= 0).  Rejected candidates
	 * are deactivated and flagged "Rejected" as MaximumTrajectoryLength does. */
	void run(const std::vector<crpropa::ref_ptr<crpropa::Candidate> > &candidates, double maxLength,
	         long long maxSteps = -1) const;

protected:
	GpuPropagation(crpropa::ref_ptr<crpropa::MagneticField> field);
	crp_field_t *field_;
	crp_propagator_t *prop_;
	crpropa::ref_ptr<crpropa::MagneticField> hostField_;

private:
	void propagate(const std::vector<crpropa::Candidate *> &c, double maxLength, long long maxSteps) const;
	mutable std::mutex mtx_;
};

/* PropagationBP(field, fixedStep) / (field, tolerance, minStep, maxStep) on the device */
class GpuPropagationBP : public GpuPropagation {
public:
	GpuPropagationBP(crpropa::ref_ptr<crpropa::MagneticField> field, double fixedStep);
	GpuPropagationBP(crpropa::ref_ptr<crpropa::MagneticField> field, double tolerance, double minStep, double maxStep);
	std::string getDescription() const;
};

/* PropagationCK(field, tolerance, minStep, maxStep) on the device */
class GpuPropagationCK : public GpuPropagation {
public:
	GpuPropagationCK(crpropa::ref_ptr<crpropa::MagneticField> field, double tolerance = 1e-4,
	                 double minStep = 0.1 * crpropa::kpc, double maxStep = 1 * crpropa::Gpc);
	std::string getDescription() const;
};

} // namespace crpropa_amd

#endif
