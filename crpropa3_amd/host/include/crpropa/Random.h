// Mersenne Twister MT19937 with the reference's seeding and conversions
// (include/crpropa/Random.h, src/Random.cpp:80-205,249-401), so that seeded
// inputs (particle energies/directions, plane-wave mode tables) are the same
// numbers the reference draws.
#ifndef CRPROPA_AMD_RANDOM_H
#define CRPROPA_AMD_RANDOM_H

#include "crpropa/Vector3.h"

#include <cstdint>
#include <vector>

namespace crpropa {

class Random {
public:
	Random();                      // seeded from /dev/urandom
	Random(uint32_t seed);
	void seed(uint32_t oneSeed);
	void seed();

	uint32_t randInt();
	double rand();                            // [0, 1]
	double randExc();                         // [0, 1)
	double randDblExc();                      // (0, 1)
	double randUniform(double min, double max);
	double randNorm(double mean = 0.0, double variance = 1.0);
	Vector3d randVector();
	double randPowerLaw(double index, double min, double max);
	size_t randBin(const std::vector<double> &cdf);

	static Random &instance(); // per-thread instance

private:
	void initialize(uint32_t s);
	void reload();
	void seedArray(const uint32_t *bigSeed, uint32_t length);
	enum { N = 624, M = 397 };
	uint32_t state[N];
	uint32_t *pNext;
	int left;
};

} // namespace crpropa

#endif
