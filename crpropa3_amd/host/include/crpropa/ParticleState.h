// Particle state and particle ids, with the reference's interface
// (include/crpropa/ParticleState.h, include/crpropa/ParticleID.h).
#ifndef CRPROPA_AMD_PARTICLESTATE_H
#define CRPROPA_AMD_PARTICLESTATE_H

#include "crpropa/Units.h"
#include "crpropa/Vector3.h"

#include <string>

namespace crpropa {

/* PDG nuclear code 10LZZZAAAI: nucleusId(A, Z) = 1000000000 + Z*10000 + A*10
 * (src/ParticleID.cpp:11-25); throws std::runtime_error for Z < 0, A < 1, A < Z. */
int nucleusId(int a, int z);
int chargeNumber(int id);
int massNumber(int id);
bool isNucleus(int id);

class ParticleState {
public:
	ParticleState(int id = 0, double energy = 0, Vector3d position = Vector3d(0, 0, 0),
	              Vector3d direction = Vector3d(-1, 0, 0));

	void setPosition(const Vector3d &pos) { position = pos; }
	const Vector3d &getPosition() const { return position; }
	/* normalises: direction = dir / dir.getR() (src/ParticleState.cpp:30-32) */
	void setDirection(const Vector3d &dir) { direction = dir / dir.getR(); }
	const Vector3d &getDirection() const { return direction; }
	void setEnergy(double e) { energy = e < 0 ? 0. : e; }
	double getEnergy() const { return energy; }
	double getRigidity() const;
	void setId(int newId);
	int getId() const { return id; }
	double getCharge() const { return charge; }
	double getMass() const { return pmass; }
	Vector3d getVelocity() const { return direction * c_light; }
	Vector3d getMomentum() const { return direction * (energy / c_light); }
	std::string getDescription() const;

	/* Sets the charge directly (for particles defined by charge only). */
	void setCharge(double q) { charge = q; }
	/* Stores a direction that is already a unit vector (a device result) without
	 * normalising it again. */
	void setDirectionExact(const Vector3d &dir) { direction = dir; }

private:
	int id;
	double energy;
	Vector3d position;
	Vector3d direction;
	double pmass;
	double charge;
};

} // namespace crpropa

#endif
