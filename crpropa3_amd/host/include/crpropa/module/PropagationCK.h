// PropagationCK: adaptive Cash-Karp, same constructor and defaults as the
// reference (include/crpropa/module/PropagationCK.h:67-68,
// src/module/PropagationCK.cpp:63-200); runs on the GPU.
#ifndef CRPROPA_AMD_PROPAGATIONCK_H
#define CRPROPA_AMD_PROPAGATIONCK_H

#include "crpropa/module/PropagationBP.h"

namespace crpropa {

class PropagationCK : public DevicePropagator {
public:
	typedef PhasePoint Y;
	PropagationCK(ref_ptr<MagneticField> field = NULL, double tolerance = 1e-4, double minStep = (0.1 * kpc),
	              double maxStep = (1 * Gpc));
	std::string getDescription() const;

	/* The reference's single-step functions (include/crpropa/module/PropagationCK.h:74-76), on the
	 * device with this propagator's field: the derivative (velocity, du/dt) of a phase point for the
	 * particle's charge and energy, and one Cash-Karp trial of duration t with its error estimate. */
	Y dYdt(const Y &y, ParticleState &p, double z) const;
	void tryStep(const Y &y, Y &out, Y &error, double t, ParticleState &p, double z) const;
};

} // namespace crpropa

#endif
