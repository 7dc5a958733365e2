// Module API with the reference's interface (include/crpropa/Module.h:18-30,
// src/Module.cpp:19-37): stateless modules whose const process() is called
// once per step on a Candidate.
#ifndef CRPROPA_AMD_MODULE_H
#define CRPROPA_AMD_MODULE_H

#include "crpropa/Candidate.h"
#include "crpropa/Referenced.h"

#include <string>
#include <typeinfo>

namespace crpropa {

class Module : public Referenced {
public:
	Module() : description("Module") {}
	virtual ~Module() {}
	virtual void process(Candidate *candidate) const = 0;
	inline void process(ref_ptr<Candidate> candidate) const { process(candidate.get()); }
	virtual std::string getDescription() const { return description; }
	void setDescription(const std::string &d) { description = d; }

private:
	std::string description;
};

/* Base of break conditions: reject() flags the candidate ("Rejected" ->
 * class name) and, by default, makes it inactive (src/Module.cpp:19-37). */
class AbstractCondition : public Module {
public:
	// the flag value is the type name seen while the base is constructed, i.e. that of
	// AbstractCondition itself, for every condition (src/Module.cpp:20-23)
	AbstractCondition()
	    : makeRejectedInactive(true), rejectFlagKey("Rejected"), rejectFlagValue(typeid(*this).name()) {}
	void reject(Candidate *candidate) const {
		if (!candidate)
			return;
		if (rejectAction.valid())
			rejectAction->process(candidate);
		if (!rejectFlagKey.empty())
			candidate->setProperty(rejectFlagKey, rejectFlagValue);
		if (makeRejectedInactive)
			candidate->setActive(false);
	}
	void setMakeRejectedInactive(bool b) { makeRejectedInactive = b; }
	bool getMakeRejectedInactive() const { return makeRejectedInactive; }
	void setRejectFlag(const std::string &key, const std::string &value) {
		rejectFlagKey = key;
		rejectFlagValue = value;
	}
	const std::string &getRejectFlagKey() const { return rejectFlagKey; }
	const std::string &getRejectFlagValue() const { return rejectFlagValue; }
	void onReject(Module *action) { rejectAction = action; }
	bool hasRejectAction() const { return rejectAction.valid(); }

protected:
	bool makeRejectedInactive;
	std::string rejectFlagKey, rejectFlagValue;
	ref_ptr<Module> rejectAction;
};

} // namespace crpropa

#endif
