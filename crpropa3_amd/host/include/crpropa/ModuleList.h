// ModuleList with the reference's interface (include/crpropa/ModuleList.h,
// src/ModuleList.cpp.in:54-216).
//
// run(candidates) is the batched path: when the list is one GPU propagator
// (PropagationBP/CK), optionally followed by MaximumTrajectoryLength without
// observer positions, all active candidates are packed into one device batch
// and propagated until inactive in a single device run.  Any other module list
// runs step by step: the propagator advances all active candidates one step on
// the GPU, then the remaining modules run on each candidate on the host.
#ifndef CRPROPA_AMD_MODULELIST_H
#define CRPROPA_AMD_MODULELIST_H

#include "crpropa/Candidate.h"
#include "crpropa/Module.h"

#include <list>

namespace crpropa {

class ModuleList : public Module {
public:
	typedef std::list<ref_ptr<Module>> module_list_t;
	typedef std::vector<ref_ptr<Candidate>> candidate_vector_t;

	ModuleList() : showProgress(false), maxStepsPerRun(-1) {}
	void setShowProgress(bool show = true) { showProgress = show; }
	void add(Module *module) { modules.push_back(module); }
	void remove(std::size_t i);
	std::size_t size() const { return modules.size(); }
	ref_ptr<Module> operator[](const std::size_t i);

	void process(Candidate *candidate) const;
	void process(ref_ptr<Candidate> candidate) const { process(candidate.get()); }
	void run(Candidate *candidate, bool recursive = true, bool secondariesFirst = false);
	void run(ref_ptr<Candidate> candidate, bool recursive = true, bool secondariesFirst = false) {
		run(candidate.get(), recursive, secondariesFirst);
	}
	void run(const candidate_vector_t *candidates, bool recursive = true, bool secondariesFirst = false);
	/* Optional cap on process() calls per candidate in one run (-1: none). */
	void setMaximumSteps(long long n) { maxStepsPerRun = n; }

	std::string getDescription() const;
	void showModules() const;
	const module_list_t &getModules() const { return modules; }

	/* Statistics of the last batched run. */
	unsigned long long lastAcceptedSteps() const { return lastAccepted; }
	double lastKernelMs() const { return lastKernel; }

private:
	bool runBatched(const candidate_vector_t &cands);
	void runStepwise(const candidate_vector_t &cands);
	module_list_t modules;
	bool showProgress;
	long long maxStepsPerRun;
	unsigned long long lastAccepted = 0;
	double lastKernel = 0;
};

} // namespace crpropa

#endif
