// PeriodicBox with the reference's interface (include/crpropa/module/Boundary.h:12-38,
// src/module/Boundary.cpp:8-41): a particle leaving the box re-enters on the opposite
// side; its previous, source and created positions move with it.  In a ModuleList
// [propagator, PeriodicBox, ...] the box runs inside the GPU kernel
// (crp_propagator_set_periodic_box); anywhere else process() runs on the host.
#ifndef CRPROPA_AMD_BOUNDARY_H
#define CRPROPA_AMD_BOUNDARY_H

#include "crpropa/Module.h"

#include <string>

namespace crpropa {

class PeriodicBox : public Module {
public:
	PeriodicBox() : origin(0.), size(0.) {}
	PeriodicBox(Vector3d origin, Vector3d size) : origin(origin), size(size) {}
	void process(Candidate *candidate) const;
	void setOrigin(Vector3d o) { origin = o; }
	void setSize(Vector3d s) { size = s; }
	Vector3d getOrigin() const { return origin; }
	Vector3d getSize() const { return size; }
	std::string getDescription() const;

private:
	Vector3d origin, size;
};

} // namespace crpropa

#endif
