// PeriodicBox and SphericalBoundary with the reference's interface
// (include/crpropa/module/Boundary.h:12-38,127-156; src/module/Boundary.cpp:8-41,208-247).
// PeriodicBox: a particle leaving the box re-enters on the opposite side; its previous,
// source and created positions move with it.  SphericalBoundary: a particle at or beyond
// the radius is rejected; with limitStep the next step is limited to radius - d + margin.
// In a ModuleList [propagator, (PeriodicBox), then SphericalBoundary, MaximumTrajectoryLength
// and Observer in any order] they run inside the GPU kernel (crp_propagator_set_periodic_box,
// crp_propagator_set_spherical_boundary); anywhere else process() runs on the host.
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

class SphericalBoundary : public AbstractCondition {
public:
	SphericalBoundary();
	SphericalBoundary(Vector3d center, double radius);
	void process(Candidate *candidate) const;
	void setCenter(Vector3d c) { center = c; }
	void setRadius(double r) { radius = r; }
	void setMargin(double m) { margin = m; }
	void setLimitStep(bool b) { limitStep = b; }
	Vector3d getCenter() const { return center; }
	double getRadius() const { return radius; }
	double getMargin() const { return margin; }
	bool getLimitStep() const { return limitStep; }
	std::string getDescription() const;

private:
	Vector3d center;
	double radius, margin;
	bool limitStep;
};

} // namespace crpropa

#endif
