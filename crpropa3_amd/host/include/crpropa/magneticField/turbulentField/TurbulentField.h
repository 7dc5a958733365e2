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
	/* l_c = 4 pi / ((s + 2) s) N lBendover, N = G((s+q)/2) / (2 G((s-1)/2) G((q+1)/2))
	 * (TurbulentField.h:33-39, 82-91) */
	virtual double getCorrelationLength() const {
		return 4 * M_PI / ((sIndex + 2.0) * sIndex) * spectrumNormalization() * lBendover;
	}

protected:
	double spectrumNormalization() const {
		return std::tgamma((sIndex + qIndex) / 2.0) /
		       (2.0 * std::tgamma((sIndex - 1) / 2.0) * std::tgamma((qIndex + 1) / 2.0));
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
	/* Harari et al. JHEP03(2002)045 (SimpleGridTurbulence.h:39-51) */
	double getCorrelationLength() const { return turbulentCorrelationLength(getLmin(), getLmax(), getSindex()); }
	static double turbulentCorrelationLength(double lMin, double lMax, double s) {
		double r = lMin / lMax;
		return lMax / 2 * (s - 1) / s * (1 - std::pow(r, s)) / (1 - std::pow(r, s - 1));
	}
};

class TurbulentField : public MagneticField {
public:
	/* Keeps its own copy of the spectrum (the reference holds a reference to the caller's). */
	TurbulentField(const TurbulenceSpectrum &spectrum)
	    : brms(spectrum.getBrms()), lmin(spectrum.getLmin()), lmax(spectrum.getLmax()),
	      lc(spectrum.getCorrelationLength()) {}
	double getBrms() const { return brms; }
	virtual double getCorrelationLength() const { return lc; }

protected:
	double brms, lmin, lmax, lc;
};

} // namespace crpropa

#endif
