// Observer with the reference's interface (include/crpropa/module/Observer.h:20-75,
// 237-360; src/module/Observer.cpp:13-90, 266-420) for the time-evolution
// feature.  In a ModuleList [propagator, (MaximumTrajectoryLength), Observer
// with one ObserverTimeEvolution] the detection test runs inside the GPU kernel
// (crp_propagator_set_observer); the detection action is then called on the
// host, candidate by candidate and in detection order, with a copy of the
// candidate in its state at detection.  Anywhere else process() runs on the host
// exactly as the reference's.
#ifndef CRPROPA_AMD_OBSERVER_H
#define CRPROPA_AMD_OBSERVER_H

#include "crpropa/Module.h"

#include <string>
#include <vector>

namespace crpropa {

enum DetectionState { DETECTED, VETO, NOTHING };

class ObserverFeature : public Referenced {
public:
	virtual ~ObserverFeature() {}
	virtual DetectionState checkDetection(Candidate *candidate) const { return NOTHING; }
	virtual void onDetection(Candidate *candidate) const {}
	virtual std::string getDescription() const { return "ObserverFeature"; }
};

class ObserverTimeEvolution : public ObserverFeature {
public:
	ObserverTimeEvolution();
	/* times min + i * dist, i < numb */
	ObserverTimeEvolution(double min, double dist, double numb);
	/* numb times from min to max, log- or lin-spaced */
	ObserverTimeEvolution(double min, double max, double numb, bool log);
	ObserverTimeEvolution(const std::vector<double> &detList);

	DetectionState checkDetection(Candidate *candidate) const;
	void constructDetListIfEmpty();
	void clear();
	bool empty() { return detList.empty(); }
	void addTime(const double &time);
	void addTimeRange(double min, double max, double numb, bool log = false);
	void setTimes(const std::vector<double> &detList);
	void setMinimum(double min);
	void setMaximum(double max) { maximum = max; }
	void setNIntervals(int n) { nIntervals = n; }
	void setIsLogarithmicScaling(bool log) { isLogarithmicScaling = log; }
	double getMinimum() const { return minimum; }
	double getMaximum() const { return maximum; }
	int getNIntervals() const { return nIntervals; }
	bool getIsLogarithmicScaling() const { return isLogarithmicScaling; }
	double getTime(size_t index) const;
	const std::vector<double> &getTimes() const;
	std::string getDescription() const;

protected:
	int nIntervals = 0;
	bool isLogarithmicScaling = false;
	bool doDetListConstruction = true;
	double minimum = 0, maximum = 0;
	std::vector<double> detList;
	mutable std::vector<double> tempDetList;
};

class Observer : public Module {
public:
	Observer();
	void add(ObserverFeature *feature);
	void onDetection(Module *action, bool clone = false);
	void process(Candidate *candidate) const;
	std::string getDescription() const;
	void setFlag(std::string key, std::string value);
	void setDeactivateOnDetection(bool deactivate);

	// read by ModuleList to fuse the observer into the device run
	const std::vector<ref_ptr<ObserverFeature>> &getFeatures() const { return features; }
	Module *getDetectionAction() const { return detectionAction.get(); }
	bool getClone() const { return clone; }
	bool getDeactivateOnDetection() const { return makeInactive; }
	const std::string &getFlagKey() const { return flagKey; }
	const std::string &getFlagValue() const { return flagValue; }

private:
	std::vector<ref_ptr<ObserverFeature>> features;
	ref_ptr<Module> detectionAction;
	bool clone;
	bool makeInactive;
	std::string flagKey, flagValue;
};

} // namespace crpropa

#endif
