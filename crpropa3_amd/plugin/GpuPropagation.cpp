// Reference-side binding of the GPU propagator (see GpuPropagation.h).
#include "GpuPropagation.h"

#include "crpropa/Units.h"
#include "crpropa/magneticField/MagneticFieldGrid.h"

#include <sstream>
#include <stdexcept>
#include <string>

namespace crpropa_amd {

using namespace crpropa;

static void check(int rc) {
	if (rc != CRP_OK)
		throw std::runtime_error(crp_last_error());
}

crp_field_t *makeDeviceField(ref_ptr<MagneticField> field) {
	crp_field_t *f = nullptr;
	if (!field.valid())
		return nullptr; // no field: B = 0 (src/module/PropagationBP.cpp:137-149)
	if (MagneticFieldGrid *mg = dynamic_cast<MagneticFieldGrid *>(field.get())) {
		ref_ptr<Grid3f> g = mg->getGrid();
		int ipol = CRP_GRID_TRILINEAR;
		switch (g->getInterpolationType()) { // Grid::setInterpolationType (include/crpropa/Grid.h:24-28)
		case TRILINEAR: ipol = CRP_GRID_TRILINEAR; break;
		case TRICUBIC: ipol = CRP_GRID_TRICUBIC; break;
		case NEAREST_NEIGHBOUR: ipol = CRP_GRID_NEAREST_NEIGHBOUR; break;
		default: throw std::runtime_error("InterpolationType: unknown interpolation type");
		}
		const Vector3d o = g->getOrigin(), s = g->getSpacing();
		const double origin[3] = {o.x, o.y, o.z}, spacing[3] = {s.x, s.y, s.z};
		// Vector3f is three packed floats: the grid's storage is the AoS layout the C ABI takes
		const float *values = reinterpret_cast<const float *>(g->getGrid().data());
		check(crp_field_create_grid3f(values, (int)g->getNx(), (int)g->getNy(), (int)g->getNz(), origin, spacing,
		                              g->isReflective() ? 1 : 0, g->getClipVolume() ? 1 : 0, &f));
		if (ipol != CRP_GRID_TRILINEAR) {
			const int rc = crp_field_set_grid_interpolation(f, CRP_GRID_VALUES, ipol);
			if (rc != CRP_OK) {
				const std::string msg = crp_last_error();
				crp_field_destroy(f);
				throw std::runtime_error(msg);
			}
		}
		return f;
	}
	if (dynamic_cast<UniformMagneticField *>(field.get())) {
		const Vector3d b = field->getField(Vector3d(0.));
		const double B[3] = {b.x, b.y, b.z};
		check(crp_field_create_uniform(B, &f));
		return f;
	}
	throw std::runtime_error("GpuPropagation: this magnetic field has no device counterpart in the plugin");
}

GpuPropagation::GpuPropagation(ref_ptr<MagneticField> field)
    : field_(makeDeviceField(field)), prop_(nullptr), hostField_(field) {}

GpuPropagation::~GpuPropagation() {
	if (prop_)
		crp_propagator_destroy(prop_);
	if (field_)
		crp_field_destroy(field_);
}

void GpuPropagation::propagate(const std::vector<Candidate *> &c, double maxLength, long long maxSteps) const {
	const size_t n = c.size();
	if (n == 0)
		return;
	std::vector<double> pos(3 * n), dir(3 * n), E(n), q(n), cur(n), next(n), len(n), time(n), z(n), ppos(3 * n),
	    pdir(3 * n);
	std::vector<int32_t> active(n);
	for (size_t i = 0; i < n; i++) {
		const ParticleState &s = c[i]->current;
		const Vector3d x = s.getPosition(), u = s.getDirection();
		pos[3 * i] = x.x; pos[3 * i + 1] = x.y; pos[3 * i + 2] = x.z;
		dir[3 * i] = u.x; dir[3 * i + 1] = u.y; dir[3 * i + 2] = u.z;
		E[i] = s.getEnergy();
		q[i] = s.getCharge();
		cur[i] = c[i]->getCurrentStep();
		next[i] = c[i]->getNextStep();
		len[i] = c[i]->getTrajectoryLength();
		time[i] = c[i]->getTime();
		z[i] = c[i]->getRedshift();
		active[i] = c[i]->isActive() ? 1 : 0;
	}
	crp_host_state s = {};
	s.pos = pos.data(); s.dir = dir.data(); s.energy = E.data(); s.charge = q.data();
	s.current_step = cur.data(); s.next_step = next.data(); s.trajectory_length = len.data(); s.time = time.data();
	s.active = active.data(); s.prev_pos = ppos.data(); s.prev_dir = pdir.data(); s.redshift = z.data();
	std::lock_guard<std::mutex> lock(mtx_); // the device propagator's length limit is shared state
	check(crp_propagator_set_max_trajectory_length(prop_, maxLength));
	crp_batch_t *b = nullptr;
	check(crp_batch_create(n, &b));
	int rc = crp_batch_upload(b, n, &s);
	crp_run_stats st;
	if (rc == CRP_OK)
		rc = crp_propagate(prop_, b, maxSteps, &st);
	if (rc == CRP_OK)
		rc = crp_batch_download(b, &s);
	crp_batch_destroy(b);
	check(rc);
	for (size_t i = 0; i < n; i++) {
		Candidate *cand = c[i];
		if (!active[i] && !cand->isActive())
			continue;
		// the reference's setters only: setDirection normalises the device's unit vector once more, and
		// setCurrentStep adds to the length and time, which are then set to the device's values
		cand->previous = cand->current;
		cand->previous.setPosition(Vector3d(ppos[3 * i], ppos[3 * i + 1], ppos[3 * i + 2]));
		cand->previous.setDirection(Vector3d(pdir[3 * i], pdir[3 * i + 1], pdir[3 * i + 2]));
		cand->current.setPosition(Vector3d(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]));
		cand->current.setDirection(Vector3d(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]));
		cand->setCurrentStep(cur[i]);
		cand->setTrajectoryLength(len[i]);
		cand->setTime(time[i]);
		cand->setNextStep(next[i]);
		if (!active[i] && cand->isActive()) { // MaximumTrajectoryLength::process (src/module/BreakCondition.cpp:42-64)
			cand->setActive(false);
			cand->setProperty("Rejected", "MaximumTrajectoryLength");
		}
	}
}

void GpuPropagation::process(Candidate *candidate) const {
	propagate(std::vector<Candidate *>(1, candidate), 0.0, 1);
}

void GpuPropagation::run(const std::vector<ref_ptr<Candidate> > &candidates, double maxLength, long long maxSteps) const {
	std::vector<Candidate *> c;
	c.reserve(candidates.size());
	for (const ref_ptr<Candidate> &p : candidates)
		if (p->isActive())
			c.push_back(p.get());
	propagate(c, maxLength, maxSteps);
}

GpuPropagationBP::GpuPropagationBP(ref_ptr<MagneticField> field, double fixedStep) : GpuPropagation(field) {
	check(crp_propagator_create_bp(field_, 0.42, fixedStep, fixedStep, &prop_));
}

GpuPropagationBP::GpuPropagationBP(ref_ptr<MagneticField> field, double tolerance, double minStep, double maxStep)
    : GpuPropagation(field) {
	check(crp_propagator_create_bp(field_, tolerance, minStep, maxStep, &prop_));
}

std::string GpuPropagationBP::getDescription() const { return "Propagation in magnetic fields using the Boris method on the GPU"; }

GpuPropagationCK::GpuPropagationCK(ref_ptr<MagneticField> field, double tolerance, double minStep, double maxStep)
    : GpuPropagation(field) {
	check(crp_propagator_create_ck(field_, tolerance, minStep, maxStep, &prop_));
}

std::string GpuPropagationCK::getDescription() const { return "Propagation in magnetic fields using the Cash-Karp method on the GPU"; }

} // namespace crpropa_amd
