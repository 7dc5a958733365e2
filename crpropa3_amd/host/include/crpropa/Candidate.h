// Candidate: the passive state object of one cosmic ray, with the reference's
// interface (include/crpropa/Candidate.h; semantics of src/Candidate.cpp:9-104).
#ifndef CRPROPA_AMD_CANDIDATE_H
#define CRPROPA_AMD_CANDIDATE_H

#include "crpropa/ParticleState.h"
#include "crpropa/Referenced.h"

#include <map>
#include <string>
#include <vector>

namespace crpropa {

class Candidate : public Referenced {
public:
	ParticleState source;
	ParticleState created;
	ParticleState current;
	ParticleState previous;
	std::vector<ref_ptr<Candidate>> secondaries;
	typedef std::map<std::string, std::string> PropertyMap;
	PropertyMap properties;
	Candidate *parent;

	Candidate(int id = 0, double energy = 0, Vector3d position = Vector3d(0, 0, 0),
	          Vector3d direction = Vector3d(-1, 0, 0), double z = 0, double weight = 1.,
	          std::string tagOrigin = "PRIM");
	Candidate(const ParticleState &state);

	bool isActive() const { return active; }
	void setActive(bool b) { active = b; }
	void setTrajectoryLength(double a) { trajectoryLength = a; }
	double getTrajectoryLength() const { return trajectoryLength; }
	double getVelocity() const { return c_light; }
	void setRedshift(double z) { redshift = z; }
	double getRedshift() const { return redshift; }
	void setWeight(double w) { weight = w; }
	double getWeight() const { return weight; }
	/* currentStep = step; trajectoryLength += step; time += step / c (src/Candidate.cpp:92-96) */
	void setCurrentStep(double step) {
		currentStep = step;
		trajectoryLength += step;
		time += step / getVelocity();
	}
	double getCurrentStep() const { return currentStep; }
	void setNextStep(double step) { nextStep = step; }
	double getNextStep() const { return nextStep; }
	/* nextStep = min(nextStep, step) (src/Candidate.cpp:102-104) */
	void limitNextStep(double step) { nextStep = step < nextStep ? step : nextStep; }
	/* Restores the step bookkeeping computed on the device. */
	void setStepState(double current, double next, double length, double t) {
		currentStep = current;
		nextStep = next;
		trajectoryLength = length;
		time = t;
	}
	void setTime(double t) { time = t; }
	double getTime() const { return time; }
	void setTagOrigin(const std::string &t) { tagOrigin = t; }
	std::string getTagOrigin() const { return tagOrigin; }

	void setProperty(const std::string &name, const std::string &value) { properties[name] = value; }
	const std::string &getProperty(const std::string &name) const;
	bool hasProperty(const std::string &name) const { return properties.count(name) != 0; }
	bool removeProperty(const std::string &name) { return properties.erase(name) != 0; }

	/* copy of the state and properties (src/Candidate.cpp:206-230); a new serial number */
	ref_ptr<Candidate> clone(bool recursive = false) const;

	uint64_t getSerialNumber() const { return serialNumber; }
	void setSerialNumber(uint64_t snr) { serialNumber = snr; }
	/* the first ancestor's serial number, and the parent's (src/Candidate.cpp:240-252) */
	uint64_t getSourceSerialNumber() const { return parent ? parent->getSourceSerialNumber() : serialNumber; }
	uint64_t getCreatedSerialNumber() const { return parent ? parent->getSerialNumber() : serialNumber; }
	static void setNextSerialNumber(uint64_t snr);
	static uint64_t getNextSerialNumber();
	/* active again at the source state, length and time 0 (src/Candidate.cpp:264-270) */
	void restart() {
		setActive(true);
		setTrajectoryLength(0);
		setTime(0);
		previous = source;
		current = source;
	}
	std::string getDescription() const;
	/* DiffusionSDE steps taken on the device: where the candidate's random numbers continue in its
	 * stream (the stream is named by the serial number) */
	uint32_t getDiffusionSteps() const { return diffusionSteps; }
	void setDiffusionSteps(uint32_t n) { diffusionSteps = n; }

private:
	uint32_t diffusionSteps = 0;
	bool active;
	double weight;
	double redshift;
	double trajectoryLength;
	double currentStep;
	double nextStep;
	std::string tagOrigin;
	double time;
	uint64_t serialNumber;
};

typedef std::vector<ref_ptr<Candidate>> candidate_vector_t;

} // namespace crpropa

#endif
