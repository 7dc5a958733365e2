// Observer and ObserverTimeEvolution on the host (src/module/Observer.cpp:13-90,
// 266-420): used when the module list is not fused into the device run, and
// for the detection lists the device run is given.
#include "crpropa/module/Observer.h"
#include "crpropa/Units.h"

#include <cmath>
#include <sstream>
#include <stdexcept>

namespace crpropa {

Observer::Observer() : clone(false), makeInactive(true) {}

void Observer::add(ObserverFeature *feature) { features.push_back(feature); }

void Observer::onDetection(Module *action, bool clone_) {
	detectionAction = action;
	clone = clone_;
}

void Observer::process(Candidate *candidate) const {
	DetectionState state = NOTHING;
	for (size_t i = 0; i < features.size(); i++) {
		DetectionState s = features[i]->checkDetection(candidate);
		if (s == VETO)
			state = VETO;
		else if ((s == DETECTED) && (state != VETO))
			state = DETECTED;
	}
	if (state == DETECTED) {
		for (size_t i = 0; i < features.size(); i++)
			features[i]->onDetection(candidate);
		if (detectionAction.valid()) {
			if (clone)
				detectionAction->process(candidate->clone(false));
			else
				detectionAction->process(candidate);
		}
		if (!flagKey.empty())
			candidate->setProperty(flagKey, flagValue);
		if (makeInactive)
			candidate->setActive(false);
	}
}

void Observer::setFlag(std::string key, std::string value) {
	flagKey = key;
	flagValue = value;
}

void Observer::setDeactivateOnDetection(bool deactivate) { makeInactive = deactivate; }

std::string Observer::getDescription() const {
	std::stringstream ss;
	ss << "Observer";
	for (size_t i = 0; i < features.size(); i++)
		ss << "\n    " << features[i]->getDescription() << "\n";
	ss << "    Flag: '" << flagKey << "' -> '" << flagValue << "'\n";
	ss << "    MakeInactive: " << (makeInactive ? "yes\n" : "no\n");
	if (detectionAction.valid())
		ss << "    Action: " << detectionAction->getDescription() << ", clone: " << (clone ? "yes" : "no");
	return ss.str();
}

ObserverTimeEvolution::ObserverTimeEvolution() {}

ObserverTimeEvolution::ObserverTimeEvolution(double min, double dist, double numb) {
	setIsLogarithmicScaling(false);
	setMaximum(min + (numb - 1) * dist);
	setMinimum(min);
	setNIntervals((int)numb);
}

ObserverTimeEvolution::ObserverTimeEvolution(double min, double max, double numb, bool log) {
	setIsLogarithmicScaling(log);
	setMinimum(min);
	setMaximum(max);
	setNIntervals((int)numb);
}

ObserverTimeEvolution::ObserverTimeEvolution(const std::vector<double> &list) { setTimes(list); }

DetectionState ObserverTimeEvolution::checkDetection(Candidate *c) const {
	if (!nIntervals)
		return NOTHING;
	const double length = c->getTrajectoryLength();
	const std::string DI = "DetectionIndex";
	size_t index = c->hasProperty(DI) ? (size_t)std::stoull(c->getProperty(DI)) : 0;
	if (index >= (size_t)nIntervals)
		return NOTHING;
	const double distance = length - getTime(index);
	if (distance < 0.) {
		c->limitNextStep(-distance);
		return NOTHING;
	}
	if ((int)index < nIntervals - 2)
		c->limitNextStep(getTime(index + 1) - length);
	c->setProperty(DI, std::to_string(index + 1));
	return DETECTED;
}

void ObserverTimeEvolution::clear() {
	doDetListConstruction = false;
	detList.clear();
	setNIntervals(0);
}

void ObserverTimeEvolution::constructDetListIfEmpty() {
	if (detList.empty() && doDetListConstruction) {
		std::vector<double> tmp;
		size_t counter = 0;
		while (getTime(counter) <= maximum) {
			tmp.push_back(getTime(counter));
			counter++;
		}
		detList.assign(tmp.begin(), tmp.end());
	}
}

void ObserverTimeEvolution::addTime(const double &time) {
	constructDetListIfEmpty();
	detList.push_back(time);
	setNIntervals(nIntervals + 1);
}

void ObserverTimeEvolution::addTimeRange(double min, double max, double numb, bool log) {
	for (size_t i = 0; i < numb; i++) {
		if (log) {
			if (min <= 0)
				throw std::runtime_error("min can not be <= 0 if log=true");
			addTime(min * std::pow(max / min, i / (numb - 1.0)));
		} else {
			addTime(min + i * (max - min) / (numb - 1.0));
		}
	}
	setNIntervals((int)detList.size());
}

void ObserverTimeEvolution::setTimes(const std::vector<double> &list) {
	detList.assign(list.begin(), list.end());
	setNIntervals((int)list.size());
	setMinimum(list.front());
	setMaximum(list.back());
	doDetListConstruction = false;
}

void ObserverTimeEvolution::setMinimum(double min) {
	if ((min <= 0) && isLogarithmicScaling)
		throw std::runtime_error("minimum can not be <= 0 if isLogarithmicScaling=true");
	minimum = min;
}

double ObserverTimeEvolution::getTime(size_t index) const {
	if (!detList.empty())
		return detList.at(index);
	if (isLogarithmicScaling)
		return minimum * std::pow(maximum / minimum, index / (nIntervals - 1.0));
	return minimum + index * (maximum - minimum) / (nIntervals - 1.0);
}

const std::vector<double> &ObserverTimeEvolution::getTimes() const {
	tempDetList.resize(nIntervals);
	for (int i = 0; i < nIntervals; i++)
		tempDetList[i] = getTime(i);
	return tempDetList;
}

std::string ObserverTimeEvolution::getDescription() const {
	std::stringstream s;
	s << "List of Detection lengths in kpc";
	for (int i = 0; i < nIntervals; i++)
		s << "  - " << getTime(i) / kpc;
	return s.str();
}

} // namespace crpropa
