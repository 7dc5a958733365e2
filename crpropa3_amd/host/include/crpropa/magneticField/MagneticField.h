// Magnetic-field interface with the reference's signatures
// (include/crpropa/magneticField/MagneticField.h:22-32).  Fields that the GPU
// evaluates return a device descriptor from deviceField(); any other field
// (e.g. one defined in Python) returns nullptr and is rejected by the
// propagators with std::runtime_error.
#ifndef CRPROPA_AMD_MAGNETICFIELD_H
#define CRPROPA_AMD_MAGNETICFIELD_H

#include "crpropa/Referenced.h"
#include "crpropa/Vector3.h"

#include <mutex>

struct crp_field;

namespace crpropa {

class MagneticField : public Referenced {
public:
	virtual ~MagneticField() {}
	virtual Vector3d getField(const Vector3d &position) const { return Vector3d(0.); }
	virtual Vector3d getField(const Vector3d &position, double z) const { return getField(position); }
	/* Device descriptor of this field, or nullptr if the GPU cannot evaluate it. */
	virtual crp_field *deviceField() const { return nullptr; }
};

/* UniformMagneticField (include/crpropa/magneticField/MagneticField.h) */
class UniformMagneticField : public MagneticField {
public:
	UniformMagneticField(const Vector3d &value) : value(value), dev(nullptr) {}
	~UniformMagneticField();
	Vector3d getField(const Vector3d &position) const { return value; }
	crp_field *deviceField() const;

private:
	Vector3d value;
	mutable crp_field *dev;
	mutable std::mutex mtx;
};

/* Evaluates any device-capable field at host positions on the GPU. */
void evaluateOnDevice(const MagneticField &f, const double *pos, size_t n, double *B);

} // namespace crpropa

#endif
