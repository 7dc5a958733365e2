// Grid-based turbulent fields with the reference's interface
// (include/crpropa/magneticField/turbulentField/GridTurbulence.h:18-74,
// SimpleGridTurbulence.h:56-112).  The random k-space field and its inverse FFT
// are computed on the GPU (crp_turbulence_generate, hipFFT) from the same
// MT19937 draws the reference takes; the normalisation to Brms uses the host
// grid tools in the reference's summation order.  FFTW is not needed.
#ifndef CRPROPA_AMD_GRIDTURBULENCE_H
#define CRPROPA_AMD_GRIDTURBULENCE_H

#include "crpropa/Grid.h"

#include <array>
#include <string>
#include "crpropa/magneticField/turbulentField/TurbulentField.h"

namespace crpropa {

class GridTurbulence : public TurbulentField {
public:
	/* TurbulenceSpectrum: amplitude sqrt(E(k)); SimpleTurbulenceSpectrum: E(k) = k^(-s-2) */
	GridTurbulence(const TurbulenceSpectrum &spectrum, const GridProperties &gridProp, unsigned int seed = 0);
	Vector3d getField(const Vector3d &pos) const;
	const ref_ptr<Grid3f> &getGrid() const { return gridPtr; }
	crp_field *deviceField() const { return gridPtr->deviceField(); }

	static void checkGridRequirements(ref_ptr<Grid3f> grid, double lMin, double lMax);
	Vector3f getMeanFieldVector() const;
	double getMeanFieldStrength() const;
	double getRmsFieldStrength() const;
	std::array<float, 3> getRmsFieldStrengthPerAxis() const;
	void dumpToFile(std::string filename) const;

protected:
	/* for SimpleGridTurbulence, which fills the grid itself */
	GridTurbulence(const TurbulenceSpectrum &spectrum, const GridProperties &gridProp, unsigned int seed, bool fill);
	unsigned int seed;
	ref_ptr<Grid3f> gridPtr;
};

class SimpleGridTurbulence : public GridTurbulence {
public:
	SimpleGridTurbulence(const SimpleTurbulenceSpectrum &spectrum, const GridProperties &gridProp,
	                     unsigned int seed = 0);
	/* random field with <B^2(k)> ~ k^alpha between lMin and lMax, normalised to Brms; seed 0 seeds
	 * from /dev/urandom as Random() does */
	static void initTurbulence(ref_ptr<Grid3f> grid, double Brms, double lMin, double lMax, double alpha, int seed);
};

/* the reference's deprecated free functions (SimpleGridTurbulence.h:68-106) */
double turbulentCorrelationLength(double lMin, double lMax, double alpha = -11 / 3.);
void initTurbulence(ref_ptr<Grid3f> grid, double Brms, double lMin, double lMax, double alpha = -11 / 3.,
                    int seed = 0);

} // namespace crpropa

#endif
