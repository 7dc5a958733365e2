// Turbulence spectra with the reference's interface
// (include/crpropa/magneticField/turbulentField/TurbulentField.h:17-92 and the
// SimpleTurbulenceSpectrum of SimpleGridTurbulence.h:21-41).
#ifndef CRPROPA_AMD_TURBULENTFIELD_H
#define CRPROPA_AMD_TURBULENTFIELD_H

#include "crpropa/magneticField/MagneticField.h"

#include <cmath>
#include <stdexcept>

namespace crpropa {

/* A(k) ~ k^q / (1 + k^2)^((s + q)/2 + 1) with k in units of 1/lBendover */
class TurbulenceSpectrum : public Referenced {
public:
	TurbulenceSpectrum(double Brms, double lMin, double lMax, double lBendover = 1, double sIndex = (5. / 3.),
	                   double qIndex = 4)
	    : Brms(Brms), sIndex(sIndex), qIndex(qIndex), lBendover(lBendover), lMin(lMin), lMax(lMax) {
		if (lMin > lMax)
			throw std::runtime_error("TurbulenceSpectrum: lMin > lMax");
		if (lMin <= 0)
			throw std::runtime_error("TurbulenceSpectrum: lMin <= 0");
	}
	virtual ~TurbulenceSpectrum() {}
	double getBrms() const { return Brms; }
	double getLmin() const { return lMin; }
	double getLmax() const { return lMax; }
	double getLbendover() const { return lBendover; }
	double getSindex() const { return sIndex; }
	double getQindex() const { return qIndex; }
	virtual double energySpectrum(double k) const {
		double kHat = k * lBendover;
		return std::pow(kHat, qIndex) / std::pow(1.0 + kHat * kHat, (sIndex + qIndex) / 2.0 + 1.0);
	}

private:
	const double Brms, sIndex, qIndex, lBendover, lMin, lMax;
};

/* E(k) = k^(-s-2), bend-over at 1000 lMax, q = 0 */
class SimpleTurbulenceSpectrum : public TurbulenceSpectrum {
public:
	SimpleTurbulenceSpectrum(double Brms, double lMin, double lMax, double sIndex = 5. / 3)
	    : TurbulenceSpectrum(Brms, lMin, lMax, 1000 * lMax, sIndex, 0) {}
	double energySpectrum(double k) const { return std::pow(k, -getSindex() - 2); }
};

class TurbulentField : public MagneticField {
public:
	/* Keeps its own copy of the spectrum parameters (the reference holds a reference). */
	TurbulentField(const TurbulenceSpectrum &spectrum)
	    : brms(spectrum.getBrms()), lmin(spectrum.getLmin()), lmax(spectrum.getLmax()) {}
	double getBrms() const { return brms; }

protected:
	double brms, lmin, lmax;
};

} // namespace crpropa

#endif
