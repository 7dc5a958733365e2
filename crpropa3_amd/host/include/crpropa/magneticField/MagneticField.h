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
#include <vector>

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

/* Field decorators (include/crpropa/magneticField/MagneticField.h:47-103,
 * src/magneticField/MagneticField.cpp:5-79), evaluated on the device: the wrapped fields must have
 * a device implementation.  Each keeps a device decorator over its children's current device
 * fields and rebuilds it when a child's device field or a setting changes. */
class DecoratorField : public MagneticField {
public:
	~DecoratorField();
	crp_field *deviceField() const;

protected:
	DecoratorField() : dev(nullptr) {}
	void invalidate();
	virtual int create(const std::vector<crp_field *> &children, crp_field **out) const = 0;
	virtual std::vector<ref_ptr<MagneticField> > childFields() const = 0;

private:
	mutable crp_field *dev;
	mutable std::vector<crp_field *> devChildren;
	mutable std::mutex mtx;
};

/* PeriodicMagneticField: the field continued periodically (or reflected) outside the cell
 * [origin, origin + extends] */
class PeriodicMagneticField : public DecoratorField {
	ref_ptr<MagneticField> field;
	Vector3d origin, extends;
	bool reflective;

public:
	PeriodicMagneticField(ref_ptr<MagneticField> field, const Vector3d &extends);
	PeriodicMagneticField(ref_ptr<MagneticField> field, const Vector3d &extends, const Vector3d &origin,
	                      bool reflective);
	Vector3d &getOrigin();
	void setOrigin(const Vector3d &origin);
	Vector3d &getExtends();
	void setExtends(const Vector3d &origin);
	bool isReflective();
	void setReflective(bool reflective);
	Vector3d getField(const Vector3d &position) const;

protected:
	int create(const std::vector<crp_field *> &children, crp_field **out) const;
	std::vector<ref_ptr<MagneticField> > childFields() const { return {field}; }
};

/* MagneticFieldList: the sum of its fields, in the order they were added */
class MagneticFieldList : public DecoratorField {
	std::vector<ref_ptr<MagneticField> > fields;

public:
	void addField(ref_ptr<MagneticField> field);
	Vector3d getField(const Vector3d &position) const;

protected:
	int create(const std::vector<crp_field *> &children, crp_field **out) const;
	std::vector<ref_ptr<MagneticField> > childFields() const { return fields; }
};

/* MagneticFieldEvolution: field * (1 + z)^m.  As in the reference only getField(position, z) is
 * overridden: through a MagneticField reference the one-argument call gives zero, a direct call
 * with one argument takes z = 0. */
class MagneticFieldEvolution : public DecoratorField {
	ref_ptr<MagneticField> field;
	double m;

public:
	MagneticFieldEvolution(ref_ptr<MagneticField> field, double m);
	Vector3d getField(const Vector3d &position, double z = 0) const;

protected:
	int create(const std::vector<crp_field *> &children, crp_field **out) const;
	std::vector<ref_ptr<MagneticField> > childFields() const { return {field}; }
};

/* Evaluates any device-capable field at host positions on the GPU. */
void evaluateOnDevice(const MagneticField &f, const double *pos, size_t n, double *B);

} // namespace crpropa

#endif
