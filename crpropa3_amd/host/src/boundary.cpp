// PeriodicBox and SphericalBoundary on the host (src/module/Boundary.cpp:8-41, 208-247).
// The same arithmetic runs inside the GPU kernel when ModuleList fuses them.
#include "crpropa/module/Boundary.h"
#include "crpropa/Units.h"

#include <sstream>

namespace crpropa {

void PeriodicBox::process(Candidate *c) const {
	Vector3d pos = c->current.getPosition();
	Vector3d n = ((pos - origin) / size).floor();
	if ((n.x == 0) && (n.y == 0) && (n.z == 0))
		return; // inside the box
	c->current.setPosition(pos - n * size);
	c->previous.setPosition(c->previous.getPosition() - n * size);
	c->source.setPosition(c->source.getPosition() - n * size);
	c->created.setPosition(c->created.getPosition() - n * size);
}

std::string PeriodicBox::getDescription() const {
	std::stringstream s;
	s << "Periodic box: origin " << origin / Mpc << " Mpc, ";
	s << "size " << size / Mpc << " Mpc";
	return s.str();
}

SphericalBoundary::SphericalBoundary() : center(0.), radius(0), margin(0.1 * kpc), limitStep(true) {}

SphericalBoundary::SphericalBoundary(Vector3d c, double r) : center(c), radius(r), margin(0.1 * kpc), limitStep(true) {}

void SphericalBoundary::process(Candidate *c) const {
	const double d = (c->current.getPosition() - center).getR();
	if (d >= radius)
		reject(c);
	if (limitStep)
		c->limitNextStep(radius - d + margin);
}

std::string SphericalBoundary::getDescription() const {
	std::stringstream s;
	s << "Spherical Boundary: radius " << radius / Mpc << " Mpc, around " << center / Mpc << " Mpc, Flag: '"
	  << rejectFlagKey << "' -> '" << rejectFlagValue << "', MakeInactive: " << (makeRejectedInactive ? "yes" : "no");
	if (rejectAction.valid())
		s << ", Action: " << rejectAction->getDescription();
	return s.str();
}

} // namespace crpropa
