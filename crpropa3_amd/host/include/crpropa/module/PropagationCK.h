// PropagationCK: adaptive Cash-Karp, same constructor and defaults as the
// reference (include/crpropa/module/PropagationCK.h:67-68,
// src/module/PropagationCK.cpp:63-200); runs on the GPU.
#ifndef CRPROPA_AMD_PROPAGATIONCK_H
#define CRPROPA_AMD_PROPAGATIONCK_H

#include "crpropa/module/PropagationBP.h"

namespace crpropa {

class PropagationCK : public DevicePropagator {
public:
	PropagationCK(ref_ptr<MagneticField> field = NULL, double tolerance = 1e-4, double minStep = (0.1 * kpc),
	              double maxStep = (1 * Gpc));
	std::string getDescription() const;
};

} // namespace crpropa

#endif
