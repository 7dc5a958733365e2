// Builds a reference ModuleList-style run with the GPU modules: a MagneticFieldGrid of the reference,
// GpuPropagationCK, candidates of the reference.  Exits 0 after a run on a GPU, or after the C ABI's
// error ("no device") is reported as the reference's std::runtime_error on a machine without one.
#include "GpuPropagation.h"

#include "crpropa/Grid.h"
#include "crpropa/ParticleID.h"
#include "crpropa/Units.h"
#include "crpropa/magneticField/MagneticFieldGrid.h"

#include <cstdio>
#include <stdexcept>

using namespace crpropa;

int main() {
	ref_ptr<Grid3f> g = new Grid3f(Vector3d(0.), 8, 50 * kpc);
	for (size_t i = 0; i < g->getGrid().size(); i++)
		g->getGrid()[i] = Vector3f(1e-13f, -2e-13f, 0.5e-13f);
	ref_ptr<MagneticField> field = new MagneticFieldGrid(g);
	std::vector<ref_ptr<Candidate> > cands;
	for (int i = 0; i < 16; i++)
		cands.push_back(new Candidate(nucleusId(60, 1 + i % 3), (1 + i) * EeV, Vector3d(0.2 * Mpc),
		                              Vector3d(1, 0.1 * i, 0)));
	try {
		ref_ptr<crpropa_amd::GpuPropagationCK> ck = new crpropa_amd::GpuPropagationCK(field, 1e-4, 1 * kpc, 1 * Mpc);
		ck->run(cands, 5 * Mpc);
		ck->process(cands[0]);
		printf("plugin_check: ran on the device; candidate 0 at %.6e m, length %.6e m\n",
		       cands[0]->current.getPosition().x, cands[0]->getTrajectoryLength());
	} catch (std::runtime_error &e) {
		printf("plugin_check: device unavailable: %s\n", e.what());
	}
	return 0;
}
