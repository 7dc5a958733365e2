// AdvectionField, UniformAdvectionField, ConstantSphericalAdvectionField and
// SphericalAdvectionField with the reference's interface
// (include/crpropa/advectionField/AdvectionField.h:23-135,
// src/advectionField/AdvectionField.cpp:33-170).  DiffusionSDE runs on the GPU
// with any of these three (or none); other advection fields raise
// std::runtime_error when a DiffusionSDE that uses them is run.
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
	/* the form the device runs: 1 uniform (a = velocity), 2 constant spherical wind (a = origin,
	 * par[0] = vWind), 3 SphericalAdvectionField (a = origin, par = radius, vMax, tau, alpha); 0: none */
	virtual int deviceForm(Vector3d &, double *) const { return 0; }
};

class UniformAdvectionField : public AdvectionField {
	Vector3d value;

public:
	UniformAdvectionField(const Vector3d &value) : value(value) {}
	Vector3d getField(const Vector3d &, const double & = 0) const { return value; }
	double getDivergence(const Vector3d &, const double & = 0) const { return 0.; }
	int deviceForm(Vector3d &a, double *) const {
		a = value;
		return 1;
	}
	std::string getDescription() const {
		std::stringstream s;
		s << "v0: " << value / km * sec << " km/s, ";
		return s.str();
	}
};

/* ConstantSphericalAdvectionField (AdvectionField.cpp:52-83): a radial wind of constant speed */
class ConstantSphericalAdvectionField : public AdvectionField {
	Vector3d origin;
	double vWind;

public:
	ConstantSphericalAdvectionField(const Vector3d origin, double vWind) : origin(origin), vWind(vWind) {}
	Vector3d getField(const Vector3d &position, const double & = 0) const {
		Vector3d Pos = position - origin;
		return Pos.getUnitVector() * vWind;
	}
	double getDivergence(const Vector3d &position, const double & = 0) const {
		double R = (position - origin).getR();
		return 2 * vWind / R;
	}
	void setOrigin(const Vector3d o) { origin = o; }
	void setVWind(double v) { vWind = v; }
	Vector3d getOrigin() const { return origin; }
	double getVWind() const { return vWind; }
	int deviceForm(Vector3d &a, double *par) const {
		a = origin;
		par[0] = vWind;
		return 2;
	}
	std::string getDescription() const {
		std::stringstream s;
		s << "Origin: " << origin / kpc << " kpc, ";
		s << "v0: " << vWind / km * sec << " km/s, ";
		return s.str();
	}
};

/* SphericalAdvectionField (AdvectionField.cpp:86-170): v(R) = vMax (1 - exp(-R^alpha / tau)) inside the
 * radius, 0 outside */
class SphericalAdvectionField : public AdvectionField {
	Vector3d origin;
	double radius, vMax, tau, alpha;

public:
	SphericalAdvectionField(const Vector3d origin, double radius, double vMax, double tau, double alpha)
	    : origin(origin), radius(radius), vMax(vMax), tau(tau), alpha(alpha) {}
	Vector3d getField(const Vector3d &position, const double & = 0) const {
		Vector3d Pos = position - origin;
		double R = Pos.getR();
		if (R > radius)
			return Vector3d(0.);
		return Pos.getUnitVector() * getV(R);
	}
	double getDivergence(const Vector3d &position, const double & = 0) const {
		double R = (position - origin).getR();
		if (R > radius)
			return 0.;
		return 2 * vMax / R * (1 - (1 - alpha * (std::pow(R, alpha) / (2 * tau))) * std::exp(-(std::pow(R, alpha) / tau)));
	}
	double getV(const double &r) const { return vMax * (1 - std::exp(-(std::pow(r, alpha) / tau))); }
	void setOrigin(const Vector3d o) { origin = o; }
	void setRadius(double r) { radius = r; }
	void setVMax(double v) { vMax = v; }
	void setTau(double t) { tau = t; }
	void setAlpha(double a) { alpha = a; }
	Vector3d getOrigin() const { return origin; }
	double getRadius() const { return radius; }
	double getVMax() const { return vMax; }
	double getTau() const { return tau; }
	double getAlpha() const { return alpha; }
	int deviceForm(Vector3d &a, double *par) const {
		a = origin;
		par[0] = radius; par[1] = vMax; par[2] = tau; par[3] = alpha;
		return 3;
	}
	std::string getDescription() const {
		std::stringstream s;
		s << "Origin: " << origin / kpc << " kpc, ";
		s << "Radius: " << radius / kpc << " kpc, ";
		s << "vMax: " << vMax / km * sec << " km/s, ";
		s << "tau: " << tau << ", ";
		s << "alpha: " << alpha << "\n";
		return s.str();
	}
};

} // namespace crpropa

#endif
