// Host side of the propagators, the length limit and ModuleList: parameter
// handling with the reference's validation, and the packing of Candidates into
// device batches (include/crpropa3_amd.h).
#include "crpropa/ModuleList.h"
#include "crpropa/Random.h"
#include "crpropa/module/Boundary.h"
#include "crpropa/module/BreakCondition.h"
#include "crpropa/module/DiffusionSDE.h"
#include "crpropa/module/Observer.h"
#include "crpropa/module/PropagationBP.h"
#include "crpropa/module/PropagationCK.h"

#include "crpropa3_amd.h"

#include <cstring>
#include <iostream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace crpropa {

static void throwIf(int rc) {
	if (rc != CRP_OK)
		throw std::runtime_error(crp_last_error());
}

// crp_propagate, except that CRP_E_STATE after a completed run (DiffusionSDE particles stopped at the
// sub-step limit; the others have moved on) is returned in `deferred`: the caller downloads and unpacks
// the batch first, so that every candidate keeps its progress, and raises afterwards.
static void propagateKeeping(crp_propagator *p, crp_batch *b, long long steps, crp_run_stats *st,
                             std::string &deferred) {
	const int rc = crp_propagate(p, b, steps, st);
	if (rc == CRP_E_STATE)
		deferred = crp_last_error();
	else
		throwIf(rc);
}

struct PropagatorDeleter {
	void operator()(crp_propagator *p) const { crp_propagator_destroy(p); }
};
struct BatchDeleter {
	void operator()(crp_batch *b) const { crp_batch_destroy(b); }
};
typedef std::unique_ptr<crp_propagator, PropagatorDeleter> PropHandle;
typedef std::unique_ptr<crp_batch, BatchDeleter> BatchHandle;

// ------------------------------------------------------------------ packing
/* Host mirror of a candidate vector in the layout crp_batch_upload takes. */
struct Packed {
	std::vector<double> pos, dir, E, q, cur, next, len, time, ppos, pdir, z;
	std::vector<int32_t> active;
	std::vector<int32_t> det; // observer detections per candidate (empty: none tracked)
	std::vector<double> src, cre; // source / created positions (empty: not tracked)
	std::vector<uint32_t> sdeStep; // DiffusionSDE: steps taken (empty: not tracked)
	std::vector<uint64_t> sdeStream; // DiffusionSDE: random stream = serial number
	void resize(size_t n) {
		pos.resize(3 * n); dir.resize(3 * n); ppos.resize(3 * n); pdir.resize(3 * n);
		E.resize(n); q.resize(n); cur.resize(n); next.resize(n); len.resize(n); time.resize(n);
		z.resize(n);
		active.resize(n);
	}
	crp_host_state view() {
		crp_host_state s;
		std::memset(&s, 0, sizeof(s));
		s.pos = pos.data(); s.dir = dir.data(); s.energy = E.data(); s.charge = q.data();
		s.current_step = cur.data(); s.next_step = next.data(); s.trajectory_length = len.data();
		s.time = time.data(); s.active = active.data(); s.prev_pos = ppos.data(); s.prev_dir = pdir.data();
		s.detection_index = det.empty() ? nullptr : det.data();
		s.source_pos = src.empty() ? nullptr : src.data();
		s.created_pos = cre.empty() ? nullptr : cre.data();
		s.sde_step = sdeStep.empty() ? nullptr : sdeStep.data();
		s.sde_stream = sdeStream.empty() ? nullptr : sdeStream.data();
		s.redshift = z.data();
		return s;
	}
};

static void pack(const std::vector<Candidate *> &c, Packed &p, bool sde = false) {
	p.resize(c.size());
	if (sde) {
		p.sdeStep.resize(c.size());
		p.sdeStream.resize(c.size());
		for (size_t i = 0; i < c.size(); i++) {
			p.sdeStep[i] = c[i]->getDiffusionSteps();
			p.sdeStream[i] = c[i]->getSerialNumber();
		}
	}
	for (size_t i = 0; i < c.size(); i++) {
		const ParticleState &s = c[i]->current;
		const Vector3d &x = s.getPosition(), &u = s.getDirection();
		p.pos[3 * i] = x.x; p.pos[3 * i + 1] = x.y; p.pos[3 * i + 2] = x.z;
		p.dir[3 * i] = u.x; p.dir[3 * i + 1] = u.y; p.dir[3 * i + 2] = u.z;
		p.E[i] = s.getEnergy();
		p.q[i] = s.getCharge();
		p.cur[i] = c[i]->getCurrentStep();
		p.next[i] = c[i]->getNextStep();
		p.len[i] = c[i]->getTrajectoryLength();
		p.time[i] = c[i]->getTime();
		p.z[i] = c[i]->getRedshift();
		p.active[i] = c[i]->isActive() ? 1 : 0;
	}
}

/* Writes device results back; `moved[i]` tells whether candidate i took a step
 * (its previous state is then the state before its last step).  A candidate the device
 * deactivated is left active here: the caller replays the rejections in module order. */
static void unpack(const std::vector<Candidate *> &c, const Packed &p, const std::vector<char> &moved) {
	for (size_t i = 0; i < c.size(); i++) {
		if (!moved[i])
			continue;
		Candidate *cand = c[i];
		// previous = state before the last step; directions are device unit vectors, kept as they are
		cand->previous = cand->current;
		cand->previous.setPosition(Vector3d(p.ppos[3 * i], p.ppos[3 * i + 1], p.ppos[3 * i + 2]));
		cand->previous.setDirectionExact(Vector3d(p.pdir[3 * i], p.pdir[3 * i + 1], p.pdir[3 * i + 2]));
		cand->current.setPosition(Vector3d(p.pos[3 * i], p.pos[3 * i + 1], p.pos[3 * i + 2]));
		cand->current.setDirectionExact(Vector3d(p.dir[3 * i], p.dir[3 * i + 1], p.dir[3 * i + 2]));
		cand->setStepState(p.cur[i], p.next[i], p.len[i], p.time[i]);
	}
	if (!p.sdeStep.empty())
		for (size_t i = 0; i < c.size(); i++)
			c[i]->setDiffusionSteps(p.sdeStep[i]);
}

// ------------------------------------------------------------------ DevicePropagator
void DevicePropagator::setTolerance(double tol) {
	if ((tol > 1) || (tol < 0))
		throw std::runtime_error(std::string(name()) +
		                         (kind == SDE ? ": tolerance error not in range 0-1" : ": target error not in range 0-1"));
	tolerance = tol;
}

void DevicePropagator::setMinimumStep(double min) {
	if (min < 0)
		throw std::runtime_error(std::string(name()) + ": minStep < 0 ");
	if (min > maxStep)
		throw std::runtime_error(std::string(name()) + ": minStep > maxStep");
	minStep = min;
}

void DevicePropagator::setMaximumStep(double max) {
	if (max < minStep)
		throw std::runtime_error(std::string(name()) + ": maxStep < minStep");
	maxStep = max;
}

crp_propagator *DevicePropagator::makeDevicePropagator(double maxLength) const {
	crp_field *f = nullptr;
	if (field.valid()) {
		f = field->deviceField();
		if (!f)
			throw std::runtime_error(std::string(name()) +
			                         ": the magnetic field has no GPU implementation (fields defined in Python or "
			                         "other host code do not run on the device)");
	}
	crp_propagator *p = nullptr;
	if (kind == BP)
		throwIf(crp_propagator_create_bp(f, tolerance, minStep, maxStep, &p));
	else if (kind == CK)
		throwIf(crp_propagator_create_ck(f, tolerance, minStep, maxStep, &p));
	else
		throwIf(crp_propagator_create_sde(f, tolerance, minStep, maxStep, static_cast<const DiffusionSDE *>(this)->getEpsilon(),
		                                  &p));
	PropHandle h(p);
	if (kind == SDE)
		static_cast<const DiffusionSDE *>(this)->configureDevice(p);
	if (maxLength > 0)
		throwIf(crp_propagator_set_max_trajectory_length(p, maxLength));
	if (precision != FP64)
		throwIf(crp_propagator_set_precision(p, CRP_PRECISION_MIXED));
	crp_launch_config lc;
	lc.block = launchBlock;
	lc.blocks_per_cu = launchBlocksPerCU;
	lc.steps_per_launch = launchStepsPerLaunch;
	lc.use_graph = launchGraph ? 1 : 0;
	lc.particles_per_lane = 0;
	throwIf(crp_propagator_set_launch(p, &lc));
	return h.release();
}

DevicePropagator::~DevicePropagator() {}

void DevicePropagator::setLaunchConfig(int block, int blocksPerCU, long long stepsPerLaunch, bool useGraph) {
	if (block != 64 && block != 128 && block != 256)
		throw std::runtime_error(std::string(name()) + ": launch block size must be 64, 128 or 256");
	if (blocksPerCU < 0)
		throw std::runtime_error(std::string(name()) + ": blocks per CU must be >= 0");
	launchBlock = block;
	launchBlocksPerCU = blocksPerCU;
	launchStepsPerLaunch = stepsPerLaunch > 0 ? stepsPerLaunch : 0;
	launchGraph = useGraph;
}

void DevicePropagator::setAutotune(bool on, long long probeSteps) {
	if (probeSteps < 1)
		throw std::runtime_error(std::string(name()) + ": autotune probe steps must be >= 1");
	autotune = on;
	autotuneProbe = probeSteps;
	tuned = false;
}

std::vector<double> DevicePropagator::getLaunchConfig() const {
	return {(double)launchBlock, (double)launchBlocksPerCU, (double)launchStepsPerLaunch, launchGraph ? 1. : 0.,
	        lastPPL};
}

void DevicePropagator::tuneOn(crp_propagator *p, const crp_batch *b) const {
	if (autotune && !tuned) {
		crp_launch_config lc;
		throwIf(crp_propagator_autotune(p, b, autotuneProbe, &lc));
		DevicePropagator *self = const_cast<DevicePropagator *>(this);
		self->launchBlock = lc.block;
		self->launchBlocksPerCU = lc.blocks_per_cu;
		self->launchStepsPerLaunch = lc.steps_per_launch;
		tuned = true;
	}
}

/* One step of one candidate on the GPU (correct, not fast: batch of one). */
void DevicePropagator::process(Candidate *candidate) const {
	PropHandle p(makeDevicePropagator(0));
	crp_batch *b = nullptr;
	throwIf(crp_batch_create(1, &b));
	BatchHandle bh(b);
	std::vector<Candidate *> one(1, candidate);
	Packed pk;
	pack(one, pk, kind == SDE);
	pk.active[0] = 1;
	crp_host_state s = pk.view();
	throwIf(crp_batch_upload(b, 1, &s));
	std::string deferred;
	propagateKeeping(p.get(), b, 1, nullptr, deferred);
	throwIf(crp_batch_download(b, &s));
	unpack(one, pk, std::vector<char>(1, 1));
	if (!pk.active[0])
		one[0]->setActive(false);
	if (!deferred.empty())
		throw std::runtime_error(deferred);
}

PropagationBP::PropagationBP(ref_ptr<MagneticField> f, double fixedStep) : DevicePropagator(BP) {
	setField(f);
	setTolerance(0.42);
	setMaximumStep(fixedStep);
	setMinimumStep(fixedStep);
}

PropagationBP::PropagationBP(ref_ptr<MagneticField> f, double tol, double minS, double maxS) : DevicePropagator(BP) {
	setField(f);
	setTolerance(tol);
	setMaximumStep(maxS);
	setMinimumStep(minS);
}

std::string PropagationBP::getDescription() const {
	std::stringstream s;
	s << "Propagation in magnetic fields using the adaptive Boris push method.";
	s << " Target error: " << tolerance;
	s << ", Minimum Step: " << minStep / kpc << " kpc";
	s << ", Maximum Step: " << maxStep / kpc << " kpc";
	return s.str();
}

PropagationCK::PropagationCK(ref_ptr<MagneticField> f, double tol, double minS, double maxS) : DevicePropagator(CK) {
	setField(f);
	setTolerance(tol);
	setMaximumStep(maxS);
	setMinimumStep(minS);
}

// ------------------------------------------------------------------ single-step functions
Vector3d DevicePropagator::getFieldAtPosition(Vector3d pos, double z) const {
	Vector3d B(0, 0, 0);
	try {
		if (field.valid())
			B = field->getField(pos, z);
	} catch (std::exception &e) {
		std::cerr << name() << ": Exception in " << name() << "::getFieldAtPosition.\n" << e.what() << std::endl;
	}
	return B;
}

namespace {
// a device propagator of this module's kind and field for the single-step C functions
struct StepHandle {
	crp_propagator *p = nullptr;
	~StepHandle() {
		if (p)
			crp_propagator_destroy(p);
	}
};
} // namespace

PropagationBP::Y PropagationBP::dY(Vector3d pos, Vector3d dir, double step, double z, double q, double m) const {
	StepHandle h;
	h.p = makeDevicePropagator(0);
	const double x[3] = {pos.x, pos.y, pos.z}, u[3] = {dir.x, dir.y, dir.z};
	double ox[3], ou[3];
	throwIf(crp_bp_dy(h.p, 1, x, u, &step, &z, &q, &m, ox, ou));
	return Y(Vector3d(ox[0], ox[1], ox[2]), Vector3d(ou[0], ou[1], ou[2]));
}

double PropagationBP::errorEstimation(const Vector3d x1, const Vector3d x2, double step) const {
	// src/module/PropagationBP.cpp:152-159: |x1 - x2| / (step * (1 - 1/4))
	Vector3d diff = (x1 - x2);
	return diff.getR() / (step * (1 - 1 / 4.));
}

void PropagationBP::tryStep(const Y &y, Y &out, Y &error, double h, ParticleState &, double z, double q,
                            double m) const {
	StepHandle hd;
	hd.p = makeDevicePropagator(0);
	const double x[3] = {y.x.x, y.x.y, y.x.z}, u[3] = {y.u.x, y.u.y, y.u.z};
	double ox[3], ou[3], ex[3], eu[3];
	throwIf(crp_bp_try_step(hd.p, 1, x, u, &h, &z, &q, &m, ox, ou, ex, eu));
	out = Y(Vector3d(ox[0], ox[1], ox[2]), Vector3d(ou[0], ou[1], ou[2]));
	error = Y(Vector3d(ex[0], ex[1], ex[2]), Vector3d(eu[0], eu[1], eu[2]));
}

PropagationCK::Y PropagationCK::dYdt(const Y &y, ParticleState &p, double z) const {
	StepHandle h;
	h.p = makeDevicePropagator(0);
	const double x[3] = {y.x.x, y.x.y, y.x.z}, u[3] = {y.u.x, y.u.y, y.u.z};
	const double q = p.getCharge(), E = p.getEnergy();
	double dx[3], du[3];
	throwIf(crp_ck_dydt(h.p, 1, x, u, &z, &q, &E, dx, du));
	return Y(Vector3d(dx[0], dx[1], dx[2]), Vector3d(du[0], du[1], du[2]));
}

void PropagationCK::tryStep(const Y &y, Y &out, Y &error, double t, ParticleState &p, double z) const {
	StepHandle h;
	h.p = makeDevicePropagator(0);
	const double x[3] = {y.x.x, y.x.y, y.x.z}, u[3] = {y.u.x, y.u.y, y.u.z};
	const double q = p.getCharge(), E = p.getEnergy();
	double ox[3], ou[3], ex[3], eu[3];
	throwIf(crp_ck_try_step(h.p, 1, x, u, &t, &z, &q, &E, ox, ou, ex, eu));
	out = Y(Vector3d(ox[0], ox[1], ox[2]), Vector3d(ou[0], ou[1], ou[2]));
	error = Y(Vector3d(ex[0], ex[1], ex[2]), Vector3d(eu[0], eu[1], eu[2]));
}

std::string PropagationCK::getDescription() const {
	std::stringstream s;
	s << "Propagation in magnetic fields using the Cash-Karp method.";
	s << " Target error: " << tolerance;
	s << ", Minimum Step: " << minStep / kpc << " kpc";
	s << ", Maximum Step: " << maxStep / kpc << " kpc";
	return s.str();
}

// ------------------------------------------------------------------ DiffusionSDE
DiffusionSDE::DiffusionSDE(ref_ptr<MagneticField> magneticField, double tol, double minS, double maxS, double e)
    : DevicePropagator(SDE) {
	// src/module/DiffusionSDE.cpp:21-32: maxStep is set before minStep (minStep starts at 0)
	setMagneticField(magneticField);
	setMaximumStep(maxS);
	setMinimumStep(minS);
	setTolerance(tol);
	setEpsilon(e);
	setScale(1.);
	setAlpha(1. / 3.);
	seed = Random::instance().randInt();
}

DiffusionSDE::DiffusionSDE(ref_ptr<MagneticField> magneticField, ref_ptr<AdvectionField> adv, double tol, double minS,
                           double maxS, double e)
    : DiffusionSDE(magneticField, tol, minS, maxS, e) {
	setAdvectionField(adv);
}

void DiffusionSDE::setEpsilon(double e) {
	if ((e > 1) || (e < 0))
		throw std::runtime_error("DiffusionSDE: epsilon not in range 0-1");
	epsilon = e;
}

void DiffusionSDE::setAlpha(double a) {
	if ((a > 2.) || (a < 0))
		throw std::runtime_error("DiffusionSDE: alpha not in range 0-2");
	alpha = a;
}

void DiffusionSDE::setScale(double s) {
	if (s < 0)
		throw std::runtime_error("DiffusionSDE: Scale error: Scale < 0");
	scale = s;
}

Vector3d DiffusionSDE::getMagneticFieldAtPosition(Vector3d pos, double z) const {
	// src/module/DiffusionSDE.cpp:352-366: an exception of the field is reported and gives zero
	Vector3d B(0, 0, 0);
	try {
		if (field.valid())
			B = field->getField(pos, z);
	} catch (std::exception &e) {
		std::cerr << "DiffusionSDE: Exception in DiffusionSDE::getMagneticFieldAtPosition.\n" << e.what() << std::endl;
	}
	return B;
}

Vector3d DiffusionSDE::getAdvectionFieldAtPosition(Vector3d pos, double t) const {
	Vector3d v(0.);
	if (advectionField.valid())
		v = advectionField->getField(pos, t);
	return v;
}

void DiffusionSDE::configureDevice(crp_propagator *p) const {
	throwIf(crp_propagator_set_sde(p, alpha, scale, epsilon, seed));
	throwIf(crp_propagator_clear_sde_advection(p));
	if (advectionField.valid()) {
		std::vector<DeviceAdvectionTerm> terms;
		advectionField->deviceTerms(terms);
		for (const DeviceAdvectionTerm &t : terms)
			throwIf(crp_propagator_add_sde_advection(p, t.kind, t.par.data(), (int)t.par.size()));
	}
}

std::string DiffusionSDE::getDescription() const {
	// src/module/DiffusionSDE.cpp:386-405
	std::stringstream s;
	s << "minStep: " << minStep / kpc << " kpc, ";
	s << "maxStep: " << maxStep / kpc << " kpc, ";
	s << "tolerance: " << tolerance << "\n";
	if (epsilon != 0.1)
		s << "epsilon: " << epsilon << ", ";
	if (alpha != 1. / 3.)
		s << "alpha: " << alpha << "\n";
	if (scale != 1.)
		s << "D_0: " << scale * 6.1e24 << " m^2/s" << "\n";
	return s.str();
}

// ------------------------------------------------------------------ MaximumTrajectoryLength
MaximumTrajectoryLength::MaximumTrajectoryLength(double length)
    : maxLength(length) {}

void MaximumTrajectoryLength::process(Candidate *c) const {
	double length = c->getTrajectoryLength();
	Vector3d position = c->current.getPosition();
	if (observerPositions.size()) {
		bool inRange = false;
		for (size_t i = 0; i < observerPositions.size(); i++)
			if (position.getDistanceTo(observerPositions[i]) + length < maxLength)
				inRange = true;
		if (!inRange) {
			reject(c);
			return;
		}
	}
	if (length >= maxLength)
		reject(c);
	else
		c->limitNextStep(maxLength - length);
}

std::string MaximumTrajectoryLength::getDescription() const {
	std::stringstream s;
	s << "Maximum trajectory length: " << maxLength / Mpc << " Mpc, ";
	s << "Flag: '" << rejectFlagKey << "' -> '" << rejectFlagValue << "', ";
	s << "MakeInactive: " << (makeRejectedInactive ? "yes" : "no");
	return s.str();
}

// ------------------------------------------------------------------ ModuleList
void ModuleList::remove(std::size_t i) {
	if (i >= modules.size())
		throw std::runtime_error("ModuleList: index out of range");
	auto it = modules.begin();
	std::advance(it, i);
	modules.erase(it);
}

ref_ptr<Module> ModuleList::operator[](const std::size_t i) {
	if (i >= modules.size())
		throw std::runtime_error("ModuleList: index out of range");
	auto it = modules.begin();
	std::advance(it, i);
	return *it;
}

void ModuleList::process(Candidate *candidate) const {
	for (auto &m : modules)
		m->process(candidate);
}

void ModuleList::run(Candidate *candidate, bool recursive, bool secondariesFirst) {
	candidate_vector_t v(1, candidate);
	run(&v, recursive, secondariesFirst);
}

/* The fused form: a DevicePropagator, optionally a PeriodicBox right after it, then in any order at
 * most one SphericalBoundary and one MaximumTrajectoryLength (default flags: they deactivate, no
 * reject action, no observer positions) and at most one Observer holding exactly one
 * ObserverTimeEvolution.  Every module of a step runs even on a candidate deactivated earlier in
 * that step (src/ModuleList.cpp.in:54-58); the kernel applies the limits (minima, so their order
 * does not matter) and deactivations, and the observer records nextStep as limited by the modules
 * listed before it (crp_propagator_set_module_order). */
struct Fused {
	const DevicePropagator *prop = nullptr;
	const PeriodicBox *box = nullptr;
	const SphericalBoundary *sphere = nullptr;
	const MaximumTrajectoryLength *limit = nullptr;
	const Observer *obs = nullptr;
	const ObserverTimeEvolution *ote = nullptr;
	std::vector<int> order;                      // CRP_MODULE_* after the propagator, in list order
	std::vector<const AbstractCondition *> conds; // rejecting modules in list order
	size_t condsBeforeObs = 0;                   // how many of them run before the observer

	// would condition c reject a candidate at this position and trajectory length?
	bool rejects(const AbstractCondition *c, const Vector3d &pos, double len) const {
		if (c == sphere)
			return (pos - sphere->getCenter()).getR() >= sphere->getRadius();
		return len >= limit->getMaximumTrajectoryLength();
	}
};

static bool fusable(const ModuleList::module_list_t &mods, Fused &f) {
	f = Fused();
	if (mods.empty() || mods.size() > 5)
		return false;
	auto it = mods.begin();
	f.prop = dynamic_cast<const DevicePropagator *>(it->get());
	if (!f.prop)
		return false;
	// a PeriodicBox is fused only directly after the propagator (it moves the positions that the
	// boundary tests and the observer records)
	auto nx = std::next(it);
	if (nx != mods.end())
		if ((f.box = dynamic_cast<const PeriodicBox *>(nx->get()))) {
			f.order.push_back(CRP_MODULE_PERIODIC_BOX);
			++it;
		}
	for (++it; it != mods.end(); ++it) {
		if (const MaximumTrajectoryLength *l = dynamic_cast<const MaximumTrajectoryLength *>(it->get())) {
			if (f.limit || !l->getObserverPositions().empty() || !l->getMakeRejectedInactive() || l->hasRejectAction() ||
			    !(l->getMaximumTrajectoryLength() > 0))
				return false;
			f.limit = l;
			f.conds.push_back(l);
			f.order.push_back(CRP_MODULE_MAX_TRAJECTORY_LENGTH);
		} else if (const SphericalBoundary *b = dynamic_cast<const SphericalBoundary *>(it->get())) {
			if (f.sphere || !b->getMakeRejectedInactive() || b->hasRejectAction())
				return false;
			f.sphere = b;
			f.conds.push_back(b);
			f.order.push_back(CRP_MODULE_SPHERICAL_BOUNDARY);
		} else if (const Observer *o = dynamic_cast<const Observer *>(it->get())) {
			if (f.obs || o->getFeatures().size() != 1)
				return false;
			const ObserverTimeEvolution *t = dynamic_cast<const ObserverTimeEvolution *>(o->getFeatures()[0].get());
			if (!t || t->getNIntervals() <= 0)
				return false;
			f.obs = o;
			f.ote = t;
			f.condsBeforeObs = f.conds.size();
			f.order.push_back(CRP_MODULE_OBSERVER);
		} else {
			return false;
		}
	}
	if (!f.obs)
		f.condsBeforeObs = f.conds.size();
	return true;
}

bool ModuleList::runBatched(const candidate_vector_t &cands) {
	Fused f;
	if (!fusable(modules, f))
		return false;
	const DevicePropagator *prop = f.prop;
	std::vector<Candidate *> act;
	for (auto &c : cands)
		if (c.valid() && c->isActive())
			act.push_back(c.get());
	lastAccepted = 0;
	lastKernel = 0;
	if (act.empty())
		return true;
	const double maxLen = f.limit ? f.limit->getMaximumTrajectoryLength() : 0;
	long long steps = maxStepsPerRun;
	if (steps < 0 && !(maxLen > 0 && prop->getMinimumStep() > 0))
		throw std::runtime_error("ModuleList::run: the module list never deactivates candidates "
		                         "(add MaximumTrajectoryLength and use minStep > 0, or setMaximumSteps)");
	PropHandle p(prop->makeDevicePropagator(maxLen));
	std::vector<double> times;
	if (f.ote) {
		times = f.ote->getTimes();
		throwIf(crp_propagator_set_observer(p.get(), times.data(), (int)times.size(),
		                                    f.obs->getDeactivateOnDetection() ? 1 : 0));
	}
	if (f.box) {
		const Vector3d o = f.box->getOrigin(), z = f.box->getSize();
		const double oo[3] = {o.x, o.y, o.z}, zz[3] = {z.x, z.y, z.z};
		throwIf(crp_propagator_set_periodic_box(p.get(), oo, zz));
	}
	if (f.sphere) {
		const Vector3d c = f.sphere->getCenter();
		const double cc[3] = {c.x, c.y, c.z};
		throwIf(crp_propagator_set_spherical_boundary(p.get(), cc, f.sphere->getRadius(), f.sphere->getMargin(),
		                                              f.sphere->getLimitStep() ? 1 : 0));
	}
	throwIf(crp_propagator_set_module_order(p.get(), f.order.data(), (int)f.order.size()));
	crp_batch *b = nullptr;
	throwIf(crp_batch_create(act.size(), &b));
	BatchHandle bh(b);
	Packed pk;
	pack(act, pk, prop->getKind() == DevicePropagator::SDE);
	if (f.box) {
		pk.src.resize(3 * act.size());
		pk.cre.resize(3 * act.size());
		for (size_t i = 0; i < act.size(); i++) {
			const Vector3d a = act[i]->source.getPosition(), c = act[i]->created.getPosition();
			pk.src[3 * i] = a.x; pk.src[3 * i + 1] = a.y; pk.src[3 * i + 2] = a.z;
			pk.cre[3 * i] = c.x; pk.cre[3 * i + 1] = c.y; pk.cre[3 * i + 2] = c.z;
		}
	}
	std::vector<int32_t> det0;
	if (f.ote) {
		// DetectionIndex as ObserverTimeEvolution::checkDetection keeps it (src/module/Observer.cpp:282-290)
		pk.det.resize(act.size());
		for (size_t i = 0; i < act.size(); i++)
			pk.det[i] = act[i]->hasProperty("DetectionIndex") ? (int32_t)std::stol(act[i]->getProperty("DetectionIndex"))
			                                                  : 0;
		det0 = pk.det;
	}
	crp_host_state s = pk.view();
	throwIf(crp_batch_upload(b, act.size(), &s));
	prop->tuneOn(p.get(), b);
	crp_run_stats st;
	std::string deferred;
	propagateKeeping(p.get(), b, steps, &st, deferred);
	throwIf(crp_batch_download(b, &s));
	std::vector<double> snaps;
	if (f.ote) {
		snaps.resize(times.size() * act.size() * 9);
		throwIf(crp_batch_detections(b, (int)times.size(), snaps.data()));
	}
	crp_launch_config used;
	throwIf(crp_propagator_get_launch(p.get(), &used));
	prop->noteParticlesPerLane(used.particles_per_lane);
	std::vector<char> moved(act.size());
	for (size_t i = 0; i < act.size(); i++)
		moved[i] = pk.len[i] != act[i]->getTrajectoryLength() || pk.cur[i] != act[i]->getCurrentStep() ||
		           !pk.active[i];
	unpack(act, pk, moved);
	if (f.box)
		for (size_t i = 0; i < act.size(); i++) {
			act[i]->source.setPosition(Vector3d(pk.src[3 * i], pk.src[3 * i + 1], pk.src[3 * i + 2]));
			act[i]->created.setPosition(Vector3d(pk.cre[3 * i], pk.cre[3 * i + 1], pk.cre[3 * i + 2]));
		}
	if (f.ote) {
		// Observer::process after the fact (src/module/Observer.cpp:26-57): per candidate, in detection
		// order, the features' onDetection and the action run on a copy of the candidate in its state at
		// detection (position, direction, trajectory length, time, nextStep; DetectionIndex = k + 1),
		// rejected by the conditions listed before the observer that reject that state
		const size_t n = act.size();
		for (size_t i = 0; i < n; i++) {
			Candidate *cand = act[i];
			for (int32_t k = det0[i]; k < pk.det[i] && k < (int32_t)times.size(); k++) {
				const double *r = &snaps[((size_t)k * n + i) * 9];
				ref_ptr<Candidate> view = cand->clone(false);
				if (!f.obs->getClone()) {
					// the reference hands over the candidate itself: same serial number and parent
					view->setSerialNumber(cand->getSerialNumber());
					view->parent = cand->parent;
				}
				const Vector3d vpos(r[0], r[1], r[2]);
				view->current.setPosition(vpos);
				view->current.setDirectionExact(Vector3d(r[3], r[4], r[5]));
				view->setStepState(view->getCurrentStep(), r[8], r[6], r[7]);
				view->setActive(true);
				for (size_t c = 0; c < f.condsBeforeObs; c++)
					if (f.rejects(f.conds[c], vpos, r[6]))
						f.conds[c]->reject(view.get());
				view->setProperty("DetectionIndex", std::to_string(k + 1));
				if (!f.obs->getFlagKey().empty())
					view->setProperty(f.obs->getFlagKey(), f.obs->getFlagValue());
				for (const ref_ptr<ObserverFeature> &ft : f.obs->getFeatures())
					ft->onDetection(view.get());
				if (Module *a = f.obs->getDetectionAction())
					a->process(view.get());
			}
			if (pk.det[i] > det0[i]) {
				cand->setProperty("DetectionIndex", std::to_string(pk.det[i]));
				if (!f.obs->getFlagKey().empty())
					cand->setProperty(f.obs->getFlagKey(), f.obs->getFlagValue());
			}
		}
	}
	// deactivations of the last step, in module order: each condition that rejects the final state
	// rejects the candidate (flag, inactive); otherwise the observer deactivated it
	for (size_t i = 0; i < act.size(); i++) {
		Candidate *cand = act[i];
		if (pk.active[i] || !cand->isActive())
			continue;
		const Vector3d pos = cand->current.getPosition();
		for (const AbstractCondition *c : f.conds)
			if (f.rejects(c, pos, pk.len[i]))
				c->reject(cand);
		cand->setActive(false);
	}
	lastAccepted = st.accepted_steps;
	lastKernel = st.kernel_ms;
	if (!deferred.empty())
		throw std::runtime_error(deferred);
	return true;
}

void ModuleList::runStepwise(const candidate_vector_t &cands) {
	// the first module must be the GPU propagator; the others run on the host
	auto it = modules.begin();
	const DevicePropagator *prop = dynamic_cast<const DevicePropagator *>(it->get());
	if (!prop) {
		for (auto &c : cands)
			while (c.valid() && c->isActive())
				process(c.get());
		return;
	}
	PropHandle p(prop->makeDevicePropagator(0));
	crp_batch *b = nullptr;
	throwIf(crp_batch_create(cands.size() ? cands.size() : 1, &b));
	BatchHandle bh(b);
	for (;;) {
		std::vector<Candidate *> act;
		for (auto &c : cands)
			if (c.valid() && c->isActive())
				act.push_back(c.get());
		if (act.empty())
			break;
		Packed pk;
		pack(act, pk, prop->getKind() == DevicePropagator::SDE);
		crp_host_state s = pk.view();
		throwIf(crp_batch_upload(b, act.size(), &s));
		std::string deferred;
		propagateKeeping(p.get(), b, 1, nullptr, deferred);
		throwIf(crp_batch_download(b, &s));
		unpack(act, pk, std::vector<char>(act.size(), 1));
		for (Candidate *c : act)
			for (auto m = std::next(modules.begin()); m != modules.end(); ++m)
				(*m)->process(c);
		if (!deferred.empty())
			throw std::runtime_error(deferred);
	}
}

void ModuleList::run(const candidate_vector_t *candidates, bool recursive, bool secondariesFirst) {
	if (!candidates || modules.empty())
		return;
	if (!runBatched(*candidates))
		runStepwise(*candidates);
	if (recursive) {
		for (auto &c : *candidates)
			if (c.valid() && !c->secondaries.empty())
				run(&c->secondaries, recursive, secondariesFirst);
	}
}

std::string ModuleList::getDescription() const {
	std::stringstream ss;
	ss << "ModuleList\n";
	for (auto &m : modules)
		ss << "  " << m->getDescription() << "\n";
	return ss.str();
}

void ModuleList::showModules() const { std::cout << getDescription(); }

} // namespace crpropa
