// The advection fields of the reference (include/crpropa/advectionField/AdvectionField.h,
// src/advectionField/AdvectionField.cpp) with its interface: AdvectionFieldList, UniformAdvectionField,
// ConstantSphericalAdvectionField, SphericalAdvectionField, OneDimensionalCartesianShock,
// OneDimensionalSphericalShock, ObliqueAdvectionShock and SphericalAdvectionShock.  getField and
// getDivergence evaluate on the host, as the reference's do; DiffusionSDE runs on the GPU with any of
// them (a list of lists raises std::runtime_error when such a DiffusionSDE is run).
#ifndef CRPROPA_AMD_ADVECTIONFIELD_H
#define CRPROPA_AMD_ADVECTIONFIELD_H

#include "crpropa/Referenced.h"
#include "crpropa/Units.h"
#include "crpropa/Vector3.h"

#include <cmath>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace crpropa {

/* one term of the device's advection field: crp_sde_advection_kind and its parameters */
struct DeviceAdvectionTerm {
	int kind;
	std::vector<double> par;
};

class AdvectionField : public Referenced {
public:
	virtual ~AdvectionField() {}
	virtual Vector3d getField(const Vector3d &position, const double &time = 0) const = 0;
	virtual double getDivergence(const Vector3d &position, const double &time = 0) const = 0;
	/* the terms the device sums for this field (in order); a field without a device form throws */
	virtual void deviceTerms(std::vector<DeviceAdvectionTerm> &) const {
		throw std::runtime_error("DiffusionSDE: this advection field has no GPU implementation");
	}
};

/* AdvectionFieldList (AdvectionField.cpp:8-25): the sum of its fields, from 0, in order */
class AdvectionFieldList : public AdvectionField {
	std::vector<ref_ptr<AdvectionField>> fields;

public:
	void addField(ref_ptr<AdvectionField> field) { fields.push_back(field); }
	Vector3d getField(const Vector3d &position, const double & = 0) const {
		Vector3d b(0.);
		for (size_t i = 0; i < fields.size(); i++)
			b += fields[i]->getField(position);
		return b;
	}
	double getDivergence(const Vector3d &position, const double & = 0) const {
		double D = 0.;
		for (size_t i = 0; i < fields.size(); i++)
			D += fields[i]->getDivergence(position);
		return D;
	}
	void deviceTerms(std::vector<DeviceAdvectionTerm> &out) const {
		for (const ref_ptr<AdvectionField> &f : fields) {
			if (dynamic_cast<const AdvectionFieldList *>(f.get()))
				throw std::runtime_error("DiffusionSDE: an AdvectionFieldList inside an AdvectionFieldList has no GPU "
				                         "implementation (the sums associate differently)");
			f->deviceTerms(out);
		}
	}
};

class UniformAdvectionField : public AdvectionField {
	Vector3d value;

public:
	UniformAdvectionField(const Vector3d &value) : value(value) {}
	Vector3d getField(const Vector3d &, const double & = 0) const { return value; }
	double getDivergence(const Vector3d &, const double & = 0) const { return 0.; }
	void deviceTerms(std::vector<DeviceAdvectionTerm> &out) const { out.push_back({1, {value.x, value.y, value.z}}); }
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
	void deviceTerms(std::vector<DeviceAdvectionTerm> &out) const {
		out.push_back({2, {origin.x, origin.y, origin.z, vWind}});
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
	void deviceTerms(std::vector<DeviceAdvectionTerm> &out) const {
		out.push_back({3, {origin.x, origin.y, origin.z, radius, vMax, tau, alpha}});
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

/* OneDimensionalCartesianShock (AdvectionField.cpp:182-241): v_x = a - b tanh(x / lShock) */
class OneDimensionalCartesianShock : public AdvectionField {
	double compressionRatio, vUp, lShock;

public:
	OneDimensionalCartesianShock(double compressionRatio, double vUp, double lShock)
	    : compressionRatio(compressionRatio), vUp(vUp), lShock(lShock) {}
	Vector3d getField(const Vector3d &position, const double & = 0) const {
		double x = position.x;
		double vDown = vUp / compressionRatio;
		double a = (vUp + vDown) * 0.5;
		double b = (vUp - vDown) * 0.5;
		Vector3d v(0.);
		v.x = a - b * std::tanh(x / lShock);
		return v;
	}
	double getDivergence(const Vector3d &position, const double & = 0) const {
		double x = position.x;
		double vDown = vUp / compressionRatio;
		double b = (vUp - vDown) * 0.5;
		return -b / lShock * (1 - std::tanh(x / lShock) * std::tanh(x / lShock));
	}
	void setComp(double r) { compressionRatio = r; }
	void setVup(double v) { vUp = v; }
	void setShockwidth(double w) { lShock = w; }
	double getComp() const { return compressionRatio; }
	double getVup() const { return vUp; }
	double getShockwidth() const { return lShock; }
	void deviceTerms(std::vector<DeviceAdvectionTerm> &out) const { out.push_back({4, {compressionRatio, vUp, lShock}}); }
	std::string getDescription() const {
		std::stringstream s;
		s << "Shock width: " << lShock / km << " km, ";
		s << "Vup: " << vUp / km * sec << " km/s, ";
		s << "Compression: " << compressionRatio;
		return s.str();
	}
};

/* OneDimensionalSphericalShock (AdvectionField.cpp:246-349): radial, the profile a - b tanh((r - rShock) /
 * lShock), times (rShock / r)^2 outside the shock, or everywhere without upstream cooling */
class OneDimensionalSphericalShock : public AdvectionField {
	double compressionRatio, vUp, lShock, rShock;
	bool coolUpstream;

public:
	OneDimensionalSphericalShock(double rShock, double vUp, double compressionRatio, double lShock, bool coolUpstream)
	    : compressionRatio(compressionRatio), vUp(vUp), lShock(lShock), rShock(rShock), coolUpstream(coolUpstream) {}
	Vector3d getField(const Vector3d &position, const double & = 0) const {
		double r = position.getR();
		Vector3d e_r = position.getUnitVector();
		double vDown = vUp / compressionRatio;
		double a = (vUp + vDown) * 0.5;
		double b = (vUp - vDown) * 0.5;
		double v;
		if (coolUpstream) {
			if (r <= rShock)
				v = a - b * std::tanh((r - rShock) / lShock);
			else
				v = (a - b * std::tanh((r - rShock) / lShock)) * (rShock / r) * (rShock / r);
		} else
			v = (a - b * std::tanh((r - rShock) / lShock)) * (rShock / r) * (rShock / r);
		return e_r * v;
	}
	double getDivergence(const Vector3d &position, const double & = 0) const {
		double r = position.getR();
		double vDown = vUp / compressionRatio;
		double a = (vUp + vDown) * 0.5;
		double b = (vUp - vDown) * 0.5;
		double c = std::tanh((r - rShock) / lShock);
		if (coolUpstream && r <= rShock)
			return 2 * a / r - 2 * b / r * c - b / lShock * (1 - c * c);
		return -(rShock / r) * (rShock / r) * b / lShock * (1 - c * c);
	}
	void setComp(double r) { compressionRatio = r; }
	void setVup(double v) { vUp = v; }
	void setShockwidth(double w) { lShock = w; }
	void setShockRadius(double r) { rShock = r; }
	void setCooling(bool c) { coolUpstream = c; }
	double getComp() const { return compressionRatio; }
	double getVup() const { return vUp; }
	double getShockwidth() const { return lShock; }
	double getShockRadius() const { return rShock; }
	bool getCooling() const { return coolUpstream; }
	void deviceTerms(std::vector<DeviceAdvectionTerm> &out) const {
		out.push_back({5, {rShock, vUp, compressionRatio, lShock, coolUpstream ? 1.0 : 0.0}});
	}
	std::string getDescription() const {
		std::stringstream s;
		s << "Shock width: " << lShock / km << " km, ";
		s << "Shock radius: " << rShock / km << " km, ";
		s << "Vup: " << vUp / km * sec << " km/s, ";
		s << "Comp: " << compressionRatio;
		return s.str();
	}
};

/* ObliqueAdvectionShock (AdvectionField.cpp:353-428): v_x = a - b tanh(x / lShock), v_y = vY */
class ObliqueAdvectionShock : public AdvectionField {
	double compressionRatio, vXUp, vY, lShock;

public:
	ObliqueAdvectionShock(double compressionRatio, double vXUp, double vY, double lShock)
	    : compressionRatio(compressionRatio), vXUp(vXUp), vY(vY), lShock(lShock) {}
	Vector3d getField(const Vector3d &position, const double & = 0) const {
		double x = position.x;
		double vXDown = vXUp / compressionRatio;
		double a = (vXUp + vXDown) * 0.5;
		double b = (vXUp - vXDown) * 0.5;
		Vector3d v(0.);
		v.x = a - b * std::tanh(x / lShock);
		v.y = vY;
		return v;
	}
	double getDivergence(const Vector3d &position, const double & = 0) const {
		double x = position.x;
		double vXDown = vXUp / compressionRatio;
		double b = (vXUp - vXDown) * 0.5;
		return -b / lShock * (1 - std::tanh(x / lShock) * std::tanh(x / lShock));
	}
	void setComp(double r) { compressionRatio = r; }
	void setVup(double v) { vXUp = v; }
	void setVy(double v) { vY = v; }
	void setShockwidth(double w) { lShock = w; }
	double getComp() const { return compressionRatio; }
	double getVup() const { return vXUp; }
	double getVy() const { return vY; }
	double getShockwidth() const { return lShock; }
	void deviceTerms(std::vector<DeviceAdvectionTerm> &out) const {
		out.push_back({6, {compressionRatio, vXUp, vY, lShock}});
	}
	std::string getDescription() const {
		std::stringstream s;
		s << "Shock width: " << lShock / km << " km, ";
		s << "Vx_up: " << vXUp / km * sec << " km/s, ";
		s << "Vy: " << vY / km * sec << " km/s, ";
		s << "Comp: " << compressionRatio;
		return s.str();
	}
};

/* SphericalAdvectionShock (AdvectionField.cpp:432-532): v_r e_r + v_phi (r_rot / r) e_phi with
 * v_r = v0 (1 + ((r0 / 2r)^2 - 1) g(r)), g = 1 / (1 + exp(-(r - r0) / lambda)) */
class SphericalAdvectionShock : public AdvectionField {
	Vector3d origin;
	double r_0, v_0, lambda, r_rot, v_phi;

public:
	SphericalAdvectionShock(const Vector3d origin, double r_0, double v_0, double lambda)
	    : origin(origin), r_0(r_0), v_0(v_0), lambda(lambda), r_rot(r_0), v_phi(0.) {}
	Vector3d getField(const Vector3d &pos, const double & = 0) const {
		Vector3d R = pos - origin;
		Vector3d e_r = R.getUnitVector();
		Vector3d e_phi = R.getUnitVectorPhi();
		double r = R.getR();
		double q = r_0 / (2 * r);
		double v_r = v_0 * (1 + (q * q - 1) * g(r)); // the reference's pow(q, 2.), which GCC writes q * q
		double v_p = v_phi * (r_rot / r);
		return e_r * v_r + e_phi * v_p;
	}
	double getDivergence(const Vector3d &pos, const double & = 0) const {
		double r = (pos - origin).getR();
		double q = r_0 / (2 * r);
		double d1 = 2. / r * (1 - g(r));
		double d2 = (q * q - 1) * g_prime(r);
		return v_0 * (d1 + d2);
	}
	double g(double r) const {
		double a = (r - r_0) / lambda;
		return 1. / (1 + std::exp(-a));
	}
	double g_prime(double r) const {
		double a = (r - r_0) / lambda;
		return 1. / (2 * lambda * (1 + std::cosh(-a)));
	}
	void setOrigin(const Vector3d o) { origin = o; }
	void setR0(double r) { r_0 = r; }
	void setV0(double v) { v_0 = v; }
	void setLambda(double l) { lambda = l; }
	void setRRot(double r) { r_rot = r; }
	void setAzimuthalSpeed(double v) { v_phi = v; }
	Vector3d getOrigin() const { return origin; }
	double getR0() const { return r_0; }
	double getV0() const { return v_0; }
	double getLambda() const { return lambda; }
	double getRRot() const { return r_rot; }
	double getAzimuthalSpeed() const { return v_phi; }
	void deviceTerms(std::vector<DeviceAdvectionTerm> &out) const {
		out.push_back({7, {origin.x, origin.y, origin.z, r_0, v_0, lambda, r_rot, v_phi}});
	}
	std::string getDescription() const {
		std::stringstream s;
		s << "Origin: " << origin / kpc << " kpc, ";
		s << "r0 (shock radius): " << r_0 / kpc << " kpc, ";
		s << "r_rot (norm. azimuthal velocity): " << r_rot / kpc << " kpc, ";
		s << "v0 (maximum radial speed): " << v_0 / km * sec << " km/s, ";
		s << "v_phi (azimuthal speed @ r_rot): " << v_phi / km * sec << " km/s, ";
		s << "lambda: " << lambda / pc << " pc";
		return s.str();
	}
};

} // namespace crpropa

#endif
