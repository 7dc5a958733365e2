// Host-side core types: particle ids, ParticleState, Candidate, Random.
#include "crpropa/Candidate.h"
#include "crpropa/ParticleState.h"
#include "crpropa/Random.h"

#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <limits>
#include <sstream>
#include <stdexcept>

namespace crpropa {

// ------------------------------------------------------------------ particle ids
int nucleusId(int a, int z) {
	if (z < 0)
		throw std::runtime_error("crpropa::Nucleus: no nucleus with Z < 0, A=" + std::to_string(a) +
		                         " Z=" + std::to_string(z));
	if (a < 1)
		throw std::runtime_error("crpropa::Nucleus: no nucleus with A < 1, A=" + std::to_string(a) +
		                         " Z=" + std::to_string(z));
	if (a < z)
		throw std::runtime_error("crpropa::Nucleus: no nucleus with A < Z, A=" + std::to_string(a) +
		                         " Z=" + std::to_string(z));
	return 1000000000 + z * 10000 + a * 10;
}

static bool isNuclearCode(int id) { return std::abs(id) >= 1000000000 && std::abs(id) < 1100000000; }

bool isNucleus(int id) {
	// the neutron (2112) and the proton (2212) count as nuclei, as in the PDG scheme
	return id == 2112 || std::abs(id) == 2212 || isNuclearCode(id);
}

int chargeNumber(int id) {
	if (isNuclearCode(id))
		return (std::abs(id) / 10000) % 1000;
	if (std::abs(id) == 2212)
		return 1;
	return 0;
}

int massNumber(int id) {
	if (id == 2112 || std::abs(id) == 2212)
		return 1;
	if (isNuclearCode(id))
		return (std::abs(id) / 10) % 1000;
	return 0;
}

/* Electric charge in units of eplus of the common non-nuclear PDG codes. */
static int pdgChargeNumber(int id) {
	int a = std::abs(id);
	int s = id < 0 ? -1 : 1;
	switch (a) {
	case 11: case 13: case 15: return -s;        // e-, mu-, tau-
	case 211: case 321: case 2212: return s;     // pi+, K+, p
	case 24: return s;                           // W+
	case 2224: return 2 * s;                     // Delta++
	default: return 0;                           // photon, neutrinos, neutron, ...
	}
}

ParticleState::ParticleState(int id, double E, Vector3d pos, Vector3d dir)
    : id(0), energy(0.), position(0.), direction(0.), pmass(0.), charge(0.) {
	setId(id);
	setEnergy(E);
	setPosition(pos);
	setDirection(dir);
}

void ParticleState::setId(int newId) {
	id = newId;
	if (isNuclearCode(id)) {
		charge = chargeNumber(id) * eplus;
		if (id < 0)
			charge *= -1;
		// mass: A amu - Z m_e (the reference reads a tabulated mass; BP/CK do not use it)
		pmass = massNumber(id) * amu - chargeNumber(id) * mass_electron;
	} else {
		charge = pdgChargeNumber(id) * eplus;
		pmass = (std::abs(id) == 11) ? mass_electron : (std::abs(id) == 2212 ? mass_proton : 0.0);
	}
}

double ParticleState::getRigidity() const { return std::fabs(energy / charge); }

std::string ParticleState::getDescription() const {
	std::stringstream ss;
	ss << "Particle " << id << ", E = " << energy / EeV << " EeV, x = " << position / Mpc << " Mpc, p = "
	   << direction;
	return ss.str();
}

// ------------------------------------------------------------------ Candidate
static std::atomic<uint64_t> g_nextSerial(0);

void Candidate::setNextSerialNumber(uint64_t snr) { g_nextSerial = snr; }
uint64_t Candidate::getNextSerialNumber() { return g_nextSerial; }

Candidate::Candidate(int id, double E, Vector3d pos, Vector3d dir, double z, double weight, std::string tagOrigin)
    : parent(nullptr), active(true), weight(weight), redshift(z), trajectoryLength(0), currentStep(0), nextStep(0),
      tagOrigin(tagOrigin), time(0) {
	ParticleState state(id, E, pos, dir);
	source = state;
	created = state;
	previous = state;
	current = state;
	serialNumber = ++g_nextSerial;
}

Candidate::Candidate(const ParticleState &state)
    : source(state), created(state), current(state), previous(state), parent(nullptr), active(true), weight(1),
      redshift(0), trajectoryLength(0), currentStep(0), nextStep(0), tagOrigin("PRIM"), time(0) {
	serialNumber = ++g_nextSerial;
}

ref_ptr<Candidate> Candidate::clone(bool recursive) const {
	ref_ptr<Candidate> c = new Candidate();
	c->source = source;
	c->created = created;
	c->current = current;
	c->previous = previous;
	c->properties = properties;
	c->active = active;
	c->redshift = redshift;
	c->weight = weight;
	c->trajectoryLength = trajectoryLength;
	c->time = time;
	c->currentStep = currentStep;
	c->nextStep = nextStep;
	if (recursive) {
		c->secondaries.reserve(secondaries.size());
		for (const ref_ptr<Candidate> &s : secondaries) {
			ref_ptr<Candidate> sc = s->clone(true);
			sc->parent = c.get();
			c->secondaries.push_back(sc);
		}
	}
	return c;
}

const std::string &Candidate::getProperty(const std::string &name) const {
	auto it = properties.find(name);
	if (it == properties.end())
		throw std::runtime_error("Unknown candidate property: " + name);
	return it->second;
}

std::string Candidate::getDescription() const {
	std::stringstream ss;
	ss << "CosmicRay at z = " << redshift << "\n  source:  " << source.getDescription()
	   << "\n  current: " << current.getDescription();
	return ss.str();
}

// ------------------------------------------------------------------ Random (MT19937)
Random::Random() { seed(); }
Random::Random(uint32_t s) { seed(s); }

void Random::initialize(uint32_t s) {
	state[0] = s;
	for (int i = 1; i < N; i++)
		state[i] = 1812433253U * (state[i - 1] ^ (state[i - 1] >> 30)) + (uint32_t)i;
}

static inline uint32_t mt_twist(uint32_t m, uint32_t s0, uint32_t s1) {
	return m ^ (((s0 & 0x80000000U) | (s1 & 0x7fffffffU)) >> 1) ^ ((0U - (s1 & 1U)) & 0x9908b0dfU);
}

void Random::reload() {
	uint32_t *p = state;
	for (int i = N - M; i--; ++p)
		*p = mt_twist(p[M], p[0], p[1]);
	for (int i = M; --i; ++p)
		*p = mt_twist(p[M - N], p[0], p[1]);
	*p = mt_twist(p[M - N], p[0], state[0]);
	left = N;
	pNext = state;
}

void Random::seed(uint32_t s) {
	initialize(s);
	reload();
}

void Random::seedArray(const uint32_t *bigSeed, uint32_t length) {
	initialize(19650218U);
	int i = 1;
	uint32_t j = 0;
	for (int k = (N > (int)length ? N : (int)length); k; --k) {
		state[i] = (state[i] ^ ((state[i - 1] ^ (state[i - 1] >> 30)) * 1664525U)) + bigSeed[j] + j;
		++i;
		++j;
		if (i >= N) {
			state[0] = state[N - 1];
			i = 1;
		}
		if (j >= length)
			j = 0;
	}
	for (int k = N - 1; k; --k) {
		state[i] = (state[i] ^ ((state[i - 1] ^ (state[i - 1] >> 30)) * 1566083941U)) - (uint32_t)i;
		++i;
		if (i >= N) {
			state[0] = state[N - 1];
			i = 1;
		}
	}
	state[0] = 0x80000000U;
	reload();
}

void Random::seed() {
	uint32_t big[N];
	FILE *f = std::fopen("/dev/urandom", "rb");
	if (f && std::fread(big, sizeof(uint32_t), N, f) == (size_t)N) {
		std::fclose(f);
		seedArray(big, N);
		return;
	}
	if (f)
		std::fclose(f);
	seed((uint32_t)std::time(nullptr));
}

uint32_t Random::randInt() {
	if (left == 0)
		reload();
	--left;
	uint32_t s = *pNext++;
	s ^= (s >> 11);
	s ^= (s << 7) & 0x9d2c5680U;
	s ^= (s << 15) & 0xefc60000U;
	return s ^ (s >> 18);
}

double Random::rand() { return double(randInt()) * (1.0 / 4294967295.0); }
double Random::randExc() { return double(randInt()) * (1.0 / 4294967296.0); }
double Random::randDblExc() { return (double(randInt()) + 0.5) * (1.0 / 4294967296.0); }
double Random::randUniform(double min, double max) { return min + (max - min) * rand(); }

double Random::randNorm(double mean, double variance) {
	double r = std::sqrt(-2.0 * std::log(1.0 - randDblExc())) * variance;
	double phi = 2.0 * 3.14159265358979323846264338328 * randExc();
	return mean + r * std::cos(phi);
}

Vector3d Random::randVector() {
	double z = randUniform(-1.0, 1.0);
	double t = randUniform(-1.0 * M_PI, M_PI);
	double r = std::sqrt(1 - z * z);
	return Vector3d(r * std::cos(t), r * std::sin(t), z);
}

double Random::randPowerLaw(double index, double min, double max) {
	if ((min < 0) || (max < min))
		throw std::runtime_error("Power law distribution only possible for 0 <= min <= max");
	if (std::abs(index + 1.0) < std::numeric_limits<double>::epsilon()) {
		double part1 = std::log(max);
		double part2 = std::log(min);
		return std::exp((part1 - part2) * rand() + part2);
	}
	double part1 = std::pow(max, index + 1);
	double part2 = std::pow(min, index + 1);
	double ex = 1 / (index + 1);
	return std::pow((part1 - part2) * rand() + part2, ex);
}

size_t Random::randBin(const std::vector<double> &cdf) {
	double v = rand() * cdf.back();
	size_t lo = 0, hi = cdf.size();
	while (lo < hi) {
		size_t mid = lo + (hi - lo) / 2;
		if (cdf[mid] < v)
			lo = mid + 1;
		else
			hi = mid;
	}
	return lo;
}

Random &Random::instance() {
	static thread_local Random r;
	return r;
}

} // namespace crpropa
