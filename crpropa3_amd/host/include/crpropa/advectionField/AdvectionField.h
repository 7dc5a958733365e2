// AdvectionField and UniformAdvectionField with the reference's interface
// (include/crpropa/advectionField/AdvectionField.h:20-58,
// src/advectionField/AdvectionField.cpp:33-50).  DiffusionSDE runs on the GPU
// with a UniformAdvectionField (or none); other advection fields are host-only.
#ifndef CRPROPA_AMD_ADVECTIONFIELD_H
#define CRPROPA_AMD_ADVECTIONFIELD_H

#include "crpropa/Referenced.h"
#include "crpropa/Units.h"
#include "crpropa/Vector3.h"

#include <sstream>
#include <string>

namespace crpropa {

class AdvectionField : public Referenced {
public:
	virtual ~AdvectionField() {}
	virtual Vector3d getField(const Vector3d &position, const double &time = 0) const = 0;
	virtual double getDivergence(const Vector3d &position, const double &time = 0) const = 0;
	/* the constant velocity if the field is uniform (the form the device runs); false otherwise */
	virtual bool uniformValue(Vector3d &) const { return false; }
};

class UniformAdvectionField : public AdvectionField {
	Vector3d value;

public:
	UniformAdvectionField(const Vector3d &value) : value(value) {}
	Vector3d getField(const Vector3d &, const double & = 0) const { return value; }
	double getDivergence(const Vector3d &, const double & = 0) const { return 0.; }
	bool uniformValue(Vector3d &v) const {
		v = value;
		return true;
	}
	std::string getDescription() const {
		std::stringstream s;
		s << "v0: " << value / km * sec << " km/s, ";
		return s.str();
	}
};

} // namespace crpropa

#endif
