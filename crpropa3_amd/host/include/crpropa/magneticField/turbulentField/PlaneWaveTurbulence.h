// PlaneWaveTurbulence with the reference's constructor
// (src/magneticField/turbulentField/PlaneWaveTurbulence.cpp:79-229): Nm modes,
// log-spaced k in [2 pi/lMax, 2 pi/lMin], amplitudes from the spectrum,
// random directions/polarisations/phases from a seeded Random.  The field sum
// B = sum_i xi_i A_i cos(k_i kappa_i.x + beta_i) is evaluated on the GPU.
#ifndef CRPROPA_AMD_PLANEWAVETURBULENCE_H
#define CRPROPA_AMD_PLANEWAVETURBULENCE_H

#include "crpropa/magneticField/turbulentField/TurbulentField.h"

#include <mutex>
#include <vector>

namespace crpropa {

class PlaneWaveTurbulence : public TurbulentField {
public:
	/* seed 0 draws a seed from /dev/urandom, as Random() does. */
	PlaneWaveTurbulence(const TurbulenceSpectrum &spectrum, int Nm = 64, int seed = 0);
	~PlaneWaveTurbulence();
	Vector3d getField(const Vector3d &pos) const;
	crp_field *deviceField() const;

	int getNm() const { return Nm; }
	const std::vector<double> &getK() const { return k; }
	const std::vector<double> &getAk() const { return Ak; }
	const std::vector<double> &getBeta() const { return beta; }
	const std::vector<Vector3d> &getXi() const { return xi; }
	const std::vector<Vector3d> &getKappa() const { return kappa; }

private:
	int Nm;
	std::vector<Vector3d> xi, kappa;
	std::vector<double> phi, costheta, beta, Ak, k;
	mutable crp_field *dev;
	mutable std::mutex mtx;
};

} // namespace crpropa

#endif
