// CPU timing of the reference itself on a benchmark workload.
//
// Links against the reference CRPropa3 C++ built from its own sources by
// oracle/Makefile (target `ref`) and runs its PropagationBP / PropagationCK
// and MaximumTrajectoryLength over a candidate vector the way
// ModuleList::run(candidates) does (src/ModuleList.cpp.in:97-151): an OpenMP
// loop with schedule(static, 100) over the candidates, each processed until
// inactive or until its step budget is spent.  Only this container runs it
// (the reference does not travel); oracle/time_reference.py drives it.
//
// Inputs: a raw float32 Grid3f file in the reference's loadGrid format
// (src/GridTools.cpp:134-166), the grid size and spacing, and the particle
// draws of the workload: one Random(seed); per particle randBin over the
// species weights (when there is more than one), randPowerLaw(-1, emin, emax),
// randVector (src/Random.cpp:129-146, 195-198); start at the box centre.
// Particle ids are nucleusId(60, Z) (no mass table is read; BP and CK do not
// use the rest mass).
//
// Usage: ref_time BP|CK grid.raw n spacing particles seed emin emax
//                 tolerance minStep maxStep maxLength steps Z1:w1[,Z2:w2...]
// (SI units; maxLength 0 = none; steps < 0 = until inactive).  Prints one JSON
// line: accepted steps, seconds, OpenMP threads, and a checksum of the final
// positions.

#include "crpropa/Candidate.h"
#include "crpropa/Grid.h"
#include "crpropa/GridTools.h"
#include "crpropa/ParticleID.h"
#include "crpropa/Random.h"
#include "crpropa/Units.h"
#include "crpropa/magneticField/MagneticFieldGrid.h"
#include "crpropa/module/BreakCondition.h"
#include "crpropa/module/PropagationBP.h"
#include "crpropa/module/PropagationCK.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <omp.h>

using namespace crpropa;

int main(int argc, char **argv) {
	if (argc != 15) {
		fprintf(stderr, "usage: %s BP|CK grid.raw n spacing particles seed emin emax tolerance minStep maxStep "
		                "maxLength steps Z1:w1[,Z2:w2...]\n", argv[0]);
		return 2;
	}
	const std::string kind = argv[1];
	const std::string gridFile = argv[2];
	const int n = atoi(argv[3]);
	const double spacing = atof(argv[4]);
	const size_t np = (size_t)atoll(argv[5]);
	const uint32_t seed = (uint32_t)atol(argv[6]);
	const double emin = atof(argv[7]), emax = atof(argv[8]);
	const double tol = atof(argv[9]), minStep = atof(argv[10]), maxStep = atof(argv[11]);
	const double maxLength = atof(argv[12]);
	const long long steps = atoll(argv[13]);
	std::vector<int> Z;
	std::vector<double> cdf;
	{
		std::string s = argv[14];
		double acc = 0;
		size_t p = 0;
		while (p < s.size()) {
			size_t q = s.find(',', p);
			std::string t = s.substr(p, q == std::string::npos ? std::string::npos : q - p);
			size_t c = t.find(':');
			Z.push_back(atoi(t.substr(0, c).c_str()));
			acc += atof(t.substr(c + 1).c_str());
			cdf.push_back(acc);
			if (q == std::string::npos)
				break;
			p = q + 1;
		}
	}

	ref_ptr<Grid3f> grid = new Grid3f(Vector3d(0.), n, spacing);
	loadGrid(grid, gridFile);
	ref_ptr<MagneticField> field = new MagneticFieldGrid(grid);

	ref_ptr<Module> prop;
	if (kind == "BP")
		prop = new PropagationBP(field, tol, minStep, maxStep);
	else
		prop = new PropagationCK(field, tol, minStep, maxStep);
	ref_ptr<MaximumTrajectoryLength> mtl = new MaximumTrajectoryLength(maxLength);

	Random rng(seed);
	const Vector3d centre(n * spacing / 2, n * spacing / 2, n * spacing / 2);
	std::vector<ref_ptr<Candidate> > cands(np);
	for (size_t i = 0; i < np; i++) {
		size_t b = cdf.size() > 1 ? rng.randBin(cdf) : 0;
		double E = rng.randPowerLaw(-1, emin, emax);
		Vector3d d = rng.randVector();
		cands[i] = new Candidate(nucleusId(60, Z[b]), E, centre, d);
	}

	std::vector<long long> acc(np, 0);
	const auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(static, 100)
	for (long long i = 0; i < (long long)np; i++) {
		Candidate *c = cands[i].get();
		long long k = 0;
		while (c->isActive() && (steps < 0 || k < steps)) {
			prop->process(c);
			if (maxLength > 0)
				mtl->process(c);
			k++;
		}
		acc[i] = k;
	}
	const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
	long long total = 0;
	double sx = 0;
	for (size_t i = 0; i < np; i++) {
		total += acc[i];
		Vector3d p = cands[i]->current.getPosition();
		sx += (p.x + p.y + p.z) / Mpc;
	}
	printf("{\"accepted_steps\": %lld, \"seconds\": %.6f, \"threads\": %d, \"position_sum_Mpc\": %.17g}\n", total, dt,
	       omp_get_max_threads(), sx);
	return 0;
}
