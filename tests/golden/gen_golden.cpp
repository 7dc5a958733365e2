// Golden-vector generator for the propagation hot path.
//
// Links against the reference CRPropa3 C++ (built from its own sources by
// oracle/Makefile, target `ref`) and writes small .npy fixtures under
// tests/golden/<case>/.  Only this container runs it; the fixtures (data)
// are what the tests read.
//
// What it exercises (reference file:line):
//   Grid<Vector3f>::interpolate / trilinearInterpolate   include/crpropa/Grid.h:318-335,540-582
//   PlaneWaveTurbulence ctor + getField                  src/magneticField/turbulentField/PlaneWaveTurbulence.cpp:79-240
//   PropagationBP::process (fixed and adaptive)          src/module/PropagationBP.cpp:62-124
//   PropagationCK::process                               src/module/PropagationCK.cpp:77-136
//   MaximumTrajectoryLength::process                     src/module/BreakCondition.cpp:42-64
//   Random (MT19937, randUniform, randVector, randPowerLaw, randBin) src/Random.cpp:80-205
//   PeriodicBox (current, previous, source, created shifted)     src/module/Boundary.cpp:16-27
//   Observer + ObserverTimeEvolution (detections, limitNextStep)  src/module/Observer.cpp:26-57,266-306
//   GridTools: scaleGrid, mean/rms field strength, loadGrid/dumpGrid (raw float32 and text),
//   loadGrid3fFromTxt, fromMagneticField                 src/GridTools.cpp:17-236,238-447
//
// Particle ids: every species uses nucleusId(60, Z) so that nuclearMass()
// takes its analytic branch (A > 56, src/ParticleMass.cpp:70-78) and the
// downloaded mass table is never read; the charge is Z*eplus as for any
// nucleus (src/ParticleState.cpp:50-56).  Rest mass is not used by BP
// (m = E/c^2, src/module/PropagationBP.cpp:85) or CK (q/E,
// src/module/PropagationCK.cpp:59), so trajectories do not depend on A.

#include "crpropa/Candidate.h"
#include "crpropa/Grid.h"
#include "crpropa/GridTools.h"
#include "crpropa/ParticleID.h"
#include "crpropa/Random.h"
#include "crpropa/Units.h"
#include "crpropa/magneticField/MagneticField.h"
#include "crpropa/magneticField/MagneticFieldGrid.h"
#include "crpropa/magneticField/turbulentField/PlaneWaveTurbulence.h"
#include "crpropa/module/Boundary.h"
#include "crpropa/module/BreakCondition.h"
#include "crpropa/module/Observer.h"
#include "crpropa/module/PropagationBP.h"
#include "crpropa/module/PropagationCK.h"
#include "crpropa/module/DiffusionSDE.h"
#include "crpropa/advectionField/AdvectionField.h"
#include "crpropa/magneticField/JF12Field.h"

#include <cstdio>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>
#include <sys/stat.h>
#include <gnu/libc-version.h>

using namespace crpropa;

static std::string g_root;

// ---------------------------------------------------------------- npy writer
static void write_npy(const std::string &path, const char *descr, const std::vector<size_t> &shape,
                      const void *data, size_t nbytes) {
	std::string sh = "(";
	for (size_t i = 0; i < shape.size(); i++) {
		sh += std::to_string(shape[i]);
		sh += (shape.size() == 1 || i + 1 < shape.size()) ? "," : "";
		if (i + 1 < shape.size()) sh += " ";
	}
	sh += ")";
	std::string header = std::string("{'descr': '") + descr + "', 'fortran_order': False, 'shape': " + sh + ", }";
	size_t total = 10 + header.size() + 1;
	size_t pad = (64 - total % 64) % 64;
	header.append(pad, ' ');
	header += '\n';
	FILE *f = fopen(path.c_str(), "wb");
	if (!f) { perror(path.c_str()); exit(1); }
	const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
	fwrite(magic, 1, 8, f);
	uint16_t hl = (uint16_t)header.size();
	fwrite(&hl, 2, 1, f);
	fwrite(header.data(), 1, header.size(), f);
	fwrite(data, 1, nbytes, f);
	fclose(f);
}

static std::string casedir(const std::string &name) {
	std::string d = g_root + "/" + name;
	mkdir(d.c_str(), 0755);
	return d;
}

static void save_f64(const std::string &dir, const std::string &name, const std::vector<double> &v,
                     std::vector<size_t> shape) {
	write_npy(dir + "/" + name + ".npy", "<f8", shape, v.data(), v.size() * 8);
}
static void save_f32(const std::string &dir, const std::string &name, const std::vector<float> &v,
                     std::vector<size_t> shape) {
	write_npy(dir + "/" + name + ".npy", "<f4", shape, v.data(), v.size() * 4);
}
static void save_i32(const std::string &dir, const std::string &name, const std::vector<int32_t> &v,
                     std::vector<size_t> shape) {
	write_npy(dir + "/" + name + ".npy", "<i4", shape, v.data(), v.size() * 4);
}

static void save_meta(const std::string &dir, const std::string &json) {
	FILE *f = fopen((dir + "/meta.json").c_str(), "w");
	fprintf(f, "%s\n", json.c_str());
	fclose(f);
}

static std::string fmt(double x) {
	char b[64];
	snprintf(b, sizeof b, "%.17g", x);
	return b;
}

// ---------------------------------------------------------------- inputs
// Spectrum equivalent to SimpleTurbulenceSpectrum (FFTW-gated in the
// reference, include/crpropa/magneticField/turbulentField/SimpleGridTurbulence.h:21-41):
// E(k) = k^(-s-2), lBendover = 1000 lMax, q = 0.
class SimpleSpectrum : public TurbulenceSpectrum {
public:
	SimpleSpectrum(double Brms, double lMin, double lMax, double s = 5. / 3)
	    : TurbulenceSpectrum(Brms, lMin, lMax, 1000 * lMax, s, 0) {}
	double energySpectrum(double k) const { return std::pow(k, -getSindex() - 2); }
};

// Grid filled with uniform white noise: each component randUniform(-1,1)*amp,
// drawn x,y,z per point in storage order (ix, iy, iz), z fastest.
static ref_ptr<Grid3f> make_noise_grid(int n, double spacing, Vector3d origin, uint32_t seed, double amp) {
	ref_ptr<Grid3f> g = new Grid3f(origin, n, spacing);
	Random rng(seed);
	for (int ix = 0; ix < n; ix++)
		for (int iy = 0; iy < n; iy++)
			for (int iz = 0; iz < n; iz++) {
				Vector3f &b = g->get(ix, iy, iz);
				b.x = rng.randUniform(-1, 1) * amp;
				b.y = rng.randUniform(-1, 1) * amp;
				b.z = rng.randUniform(-1, 1) * amp;
			}
	return g;
}

static void save_grid(const std::string &dir, ref_ptr<Grid3f> g) {
	size_t n = g->getNx() * g->getNy() * g->getNz();
	std::vector<float> v(n * 3);
	for (size_t i = 0; i < n; i++) {
		const Vector3f &b = g->getGrid()[i];
		v[3 * i] = b.x; v[3 * i + 1] = b.y; v[3 * i + 2] = b.z;
	}
	save_f32(dir, "grid", v, {g->getNx(), g->getNy(), g->getNz(), 3});
}

struct Particle {
	int id;
	double E;
	Vector3d pos, dir;
};

// Draw order per particle: species (if weights given), energy, direction.
static std::vector<Particle> make_particles(size_t N, uint32_t seed, double Emin, double Emax, Vector3d start,
                                            const std::vector<int> &ids, const std::vector<double> &weights) {
	Random rng(seed);
	std::vector<double> cdf;
	double acc = 0;
	for (double w : weights) { acc += w; cdf.push_back(acc); }
	std::vector<Particle> ps(N);
	for (size_t i = 0; i < N; i++) {
		int id = ids[0];
		if (ids.size() > 1) id = ids[rng.randBin(cdf)];
		ps[i].id = id;
		ps[i].E = rng.randPowerLaw(-1, Emin, Emax);
		ps[i].dir = rng.randVector();
		ps[i].pos = start;
	}
	return ps;
}

static void save_particles(const std::string &dir, const std::vector<Particle> &ps) {
	size_t N = ps.size();
	std::vector<double> pos(3 * N), d(3 * N), E(N), q(N);
	std::vector<int32_t> id(N);
	for (size_t i = 0; i < N; i++) {
		pos[3 * i] = ps[i].pos.x; pos[3 * i + 1] = ps[i].pos.y; pos[3 * i + 2] = ps[i].pos.z;
		d[3 * i] = ps[i].dir.x; d[3 * i + 1] = ps[i].dir.y; d[3 * i + 2] = ps[i].dir.z;
		E[i] = ps[i].E;
		id[i] = ps[i].id;
		ParticleState s(ps[i].id, ps[i].E, ps[i].pos, ps[i].dir);
		q[i] = s.getCharge();
	}
	save_f64(dir, "init_pos", pos, {N, 3});
	save_f64(dir, "init_dir", d, {N, 3});
	save_f64(dir, "energy", E, {N});
	save_f64(dir, "charge", q, {N});
	save_i32(dir, "id", id, {N});
}

// ---------------------------------------------------------------- trajectories
// Per-particle redshifts for the next run_traj (candidate->setRedshift), or none (z = 0)
static const std::vector<double> *g_redshift = nullptr;

// Runs `prop` (then the optional length limit) on each particle, the way
// ModuleList::run(Candidate*) loops (src/ModuleList.cpp.in:64-67), for at most
// maxSteps process() calls, recording the state after the listed step counts.
static void run_traj(const std::string &name, const Module &prop, const MaximumTrajectoryLength *maxlen,
                     const std::vector<Particle> &ps, size_t maxSteps, const std::vector<size_t> &snaps,
                     const std::string &meta) {
	std::string dir = casedir(name);
	save_particles(dir, ps);
	size_t N = ps.size(), S = snaps.size();
	std::vector<double> pos(S * N * 3), d(S * N * 3), cur(S * N), nxt(S * N), len(S * N), tim(S * N);
	std::vector<int32_t> act(S * N), nsteps(N);
	std::vector<double> prevpos(N * 3);
#pragma omp parallel for schedule(dynamic, 4)
	for (size_t i = 0; i < N; i++) {
		ParticleState st(ps[i].id, ps[i].E, ps[i].pos, ps[i].dir);
		ref_ptr<Candidate> c = new Candidate(st);
		if (g_redshift)
			c->setRedshift((*g_redshift)[i]);
		size_t si = 0, step = 0;
		while (step < maxSteps && c->isActive()) {
			prop.process(c);
			if (maxlen) maxlen->process(c);
			step++;
			while (si < S && snaps[si] == step) {
				size_t k = si * N + i;
				Vector3d p = c->current.getPosition(), u = c->current.getDirection();
				pos[3 * k] = p.x; pos[3 * k + 1] = p.y; pos[3 * k + 2] = p.z;
				d[3 * k] = u.x; d[3 * k + 1] = u.y; d[3 * k + 2] = u.z;
				cur[k] = c->getCurrentStep(); nxt[k] = c->getNextStep();
				len[k] = c->getTrajectoryLength(); tim[k] = c->getTime();
				act[k] = c->isActive() ? 1 : 0;
				si++;
			}
		}
		nsteps[i] = (int32_t)step;
		// snapshots past the particle's end hold its final state
		for (; si < S; si++) {
			size_t k = si * N + i;
			Vector3d p = c->current.getPosition(), u = c->current.getDirection();
			pos[3 * k] = p.x; pos[3 * k + 1] = p.y; pos[3 * k + 2] = p.z;
			d[3 * k] = u.x; d[3 * k + 1] = u.y; d[3 * k + 2] = u.z;
			cur[k] = c->getCurrentStep(); nxt[k] = c->getNextStep();
			len[k] = c->getTrajectoryLength(); tim[k] = c->getTime();
			act[k] = c->isActive() ? 1 : 0;
		}
		Vector3d pp = c->previous.getPosition();
		prevpos[3 * i] = pp.x; prevpos[3 * i + 1] = pp.y; prevpos[3 * i + 2] = pp.z;
	}
	std::vector<int32_t> sn(snaps.begin(), snaps.end());
	save_i32(dir, "snap_steps", sn, {S});
	save_f64(dir, "pos", pos, {S, N, 3});
	save_f64(dir, "dir", d, {S, N, 3});
	save_f64(dir, "current_step", cur, {S, N});
	save_f64(dir, "next_step", nxt, {S, N});
	save_f64(dir, "trajectory_length", len, {S, N});
	save_f64(dir, "time", tim, {S, N});
	save_i32(dir, "active", act, {S, N});
	save_i32(dir, "nsteps", nsteps, {N});
	save_f64(dir, "final_previous_pos", prevpos, {N, 3});
	if (g_redshift)
		save_f64(dir, "redshift", *g_redshift, {N});
	save_meta(dir, meta);
	printf("  %-28s N=%zu steps<=%zu\n", name.c_str(), N, maxSteps);
}

// ---------------------------------------------------------------- cases
// ---------------------------------------------------------------- observers
// Records every detection the Observer reports: particle (property "pid"),
// state at detection.  One vector per particle, so OpenMP threads never share one.
class DetectionRecorder : public Module {
public:
	struct Rec {
		Vector3d pos, dir;
		double len, time, next;
		int active, rejected; // the candidate's flags as the observer saw them
	};
	std::vector<std::vector<Rec>> *per;
	void process(Candidate *c) const {
		size_t pid = c->getProperty("pid").asUInt64();
		(*per)[pid].push_back(Rec{c->current.getPosition(), c->current.getDirection(), c->getTrajectoryLength(),
		                          c->getTime(), c->getNextStep(), c->isActive() ? 1 : 0,
		                          c->hasProperty("Rejected") ? 1 : 0});
	}
};

// ModuleList order [propagator, MaximumTrajectoryLength (optional), Observer(ObserverTimeEvolution)].
// `pre` run between the propagator and the observer, `post` after it.
static void run_observer_case(const std::string &name, const Module &prop, const std::vector<const Module *> &pre,
                              const std::vector<const Module *> &post, ObserverTimeEvolution *ote, bool deactivate,
                              const std::vector<Particle> &ps, size_t maxSteps, const std::string &meta) {
	std::string dir = casedir(name);
	save_particles(dir, ps);
	const size_t N = ps.size();
	std::vector<std::vector<DetectionRecorder::Rec>> per(N);
	ref_ptr<DetectionRecorder> rec = new DetectionRecorder();
	rec->per = &per;
	ref_ptr<Observer> obs = new Observer();
	obs->add(ote);
	obs->setDeactivateOnDetection(deactivate);
	obs->onDetection(rec);
	std::vector<double> fpos(3 * N), fdir(3 * N), flen(N), fnext(N);
	std::vector<int32_t> fact(N), nsteps(N);
#pragma omp parallel for schedule(dynamic, 4)
	for (size_t i = 0; i < N; i++) {
		ParticleState st(ps[i].id, ps[i].E, ps[i].pos, ps[i].dir);
		ref_ptr<Candidate> c = new Candidate(st);
		c->setProperty("pid", Variant::fromUInt64(i));
		size_t step = 0;
		while (step < maxSteps && c->isActive()) {
			prop.process(c);
			for (const Module *m : pre)
				m->process(c);
			obs->process(c);
			for (const Module *m : post)
				m->process(c);
			step++;
		}
		Vector3d p = c->current.getPosition(), u = c->current.getDirection();
		fpos[3 * i] = p.x; fpos[3 * i + 1] = p.y; fpos[3 * i + 2] = p.z;
		fdir[3 * i] = u.x; fdir[3 * i + 1] = u.y; fdir[3 * i + 2] = u.z;
		flen[i] = c->getTrajectoryLength();
		fnext[i] = c->getNextStep();
		fact[i] = c->isActive() ? 1 : 0;
		nsteps[i] = (int32_t)step;
	}
	const std::vector<double> &T = ote->getTimes();
	const size_t K = T.size();
	std::vector<double> dpos(K * N * 3, NAN), ddir(K * N * 3, NAN), dlen(K * N, NAN), dtim(K * N, NAN), dnext(K * N, NAN);
	std::vector<int32_t> count(N), dact(K * N, -1), drej(K * N, -1);
	for (size_t i = 0; i < N; i++) {
		count[i] = (int32_t)per[i].size();
		for (size_t k = 0; k < per[i].size() && k < K; k++) {
			const DetectionRecorder::Rec &r = per[i][k];
			size_t o = k * N + i;
			dpos[3 * o] = r.pos.x; dpos[3 * o + 1] = r.pos.y; dpos[3 * o + 2] = r.pos.z;
			ddir[3 * o] = r.dir.x; ddir[3 * o + 1] = r.dir.y; ddir[3 * o + 2] = r.dir.z;
			dlen[o] = r.len; dtim[o] = r.time; dnext[o] = r.next;
			dact[o] = r.active; drej[o] = r.rejected;
		}
	}
	save_f64(dir, "times", T, {K});
	save_f64(dir, "det_pos", dpos, {K, N, 3});
	save_f64(dir, "det_dir", ddir, {K, N, 3});
	save_f64(dir, "det_len", dlen, {K, N});
	save_f64(dir, "det_time", dtim, {K, N});
	save_f64(dir, "det_next", dnext, {K, N});
	save_i32(dir, "det_count", count, {N});
	save_i32(dir, "det_active", dact, {K, N});
	save_i32(dir, "det_rejected", drej, {K, N});
	save_f64(dir, "final_pos", fpos, {N, 3});
	save_f64(dir, "final_dir", fdir, {N, 3});
	save_f64(dir, "final_len", flen, {N});
	save_f64(dir, "final_next", fnext, {N});
	save_i32(dir, "final_active", fact, {N});
	save_i32(dir, "nsteps", nsteps, {N});
	save_meta(dir, meta);
	printf("  %-28s N=%zu steps<=%zu detections=%zu\n", name.c_str(), N, maxSteps, K);
}

// ---------------------------------------------------------------- periodic box
// ModuleList order [propagator, PeriodicBox, MaximumTrajectoryLength (optional)];
// records current/previous/source/created positions after the listed step counts.
// propagator followed by modules that move or reject the candidate (PeriodicBox, SphericalBoundary,
// MaximumTrajectoryLength), in list order, as ModuleList::process runs them
static void run_modules_case(const std::string &name, const Module &prop, const std::vector<const Module *> &mods,
                             const std::vector<Particle> &ps, size_t maxSteps, const std::vector<size_t> &snaps,
                             const std::string &meta) {
	std::string dir = casedir(name);
	save_particles(dir, ps);
	const size_t N = ps.size(), S = snaps.size();
	std::vector<double> pos(S * N * 3), prv(S * N * 3), src(S * N * 3), cre(S * N * 3), d(S * N * 3), len(S * N),
	    nxt(S * N);
	std::vector<int32_t> act(S * N);
#pragma omp parallel for schedule(dynamic, 4)
	for (size_t i = 0; i < N; i++) {
		ParticleState st(ps[i].id, ps[i].E, ps[i].pos, ps[i].dir);
		ref_ptr<Candidate> c = new Candidate(st);
		auto record = [&](size_t si) {
			size_t k = si * N + i;
			Vector3d a = c->current.getPosition(), b = c->previous.getPosition(), s0 = c->source.getPosition(),
			         s1 = c->created.getPosition(), u = c->current.getDirection();
			const Vector3d *vs[5] = {&a, &b, &s0, &s1, &u};
			std::vector<double> *out[5] = {&pos, &prv, &src, &cre, &d};
			for (int v = 0; v < 5; v++) {
				(*out[v])[3 * k] = vs[v]->x;
				(*out[v])[3 * k + 1] = vs[v]->y;
				(*out[v])[3 * k + 2] = vs[v]->z;
			}
			len[k] = c->getTrajectoryLength();
			nxt[k] = c->getNextStep();
			act[k] = c->isActive() ? 1 : 0;
		};
		size_t si = 0, step = 0;
		while (step < maxSteps && c->isActive()) {
			prop.process(c);
			for (const Module *m : mods)
				m->process(c);
			step++;
			while (si < S && snaps[si] == step)
				record(si++);
		}
		for (; si < S; si++) // snapshots past the particle's end hold its final state
			record(si);
	}
	std::vector<double> sn(snaps.begin(), snaps.end());
	save_f64(dir, "snap_steps", sn, {S});
	save_f64(dir, "pos", pos, {S, N, 3});
	save_f64(dir, "previous_pos", prv, {S, N, 3});
	save_f64(dir, "source_pos", src, {S, N, 3});
	save_f64(dir, "created_pos", cre, {S, N, 3});
	save_f64(dir, "dir", d, {S, N, 3});
	save_f64(dir, "trajectory_length", len, {S, N});
	save_f64(dir, "next_step", nxt, {S, N});
	save_i32(dir, "active", act, {S, N});
	save_meta(dir, meta);
	printf("  %-28s N=%zu steps<=%zu\n", name.c_str(), N, maxSteps);
}

static void gen_trilinear() {
	// 16^3 grid with a non-trivial origin; spacing 50 kpc; amplitude 1 nG.
	const int n = 16;
	const double sp = 50 * kpc;
	Vector3d origin(-3.1 * kpc, 7.25 * kpc, 0.0);
	ref_ptr<Grid3f> g = make_noise_grid(n, sp, origin, 1234, nG);
	Random rng(99);
	const size_t M = 4096;
	std::vector<double> pos(M * 3);
	double L = n * sp;
	for (size_t i = 0; i < M; i++) {
		for (int a = 0; a < 3; a++) {
			double o = origin.data[a];
			double x;
			switch (i % 8) {
			case 0: x = o + rng.randUniform(0, L); break;                          // inside
			case 1: x = o + rng.randUniform(-3 * L, 4 * L); break;                 // wrapped
			case 2: x = o + sp / 2 + sp * (int)rng.randUniform(-n, 2 * n); break;  // on grid points
			case 3: x = o + sp * (int)rng.randUniform(-n, 2 * n); break;           // on cell faces
			case 4: x = o + rng.randUniform(-sp, sp); break;                       // near the lower edge
			case 5: x = o + L + rng.randUniform(-sp, sp); break;                   // near the upper edge
			case 6: x = o + rng.randUniform(-1e4 * L, 1e4 * L); break;             // far away
			default: x = rng.randUniform(-1, 1) * sp; break;                       // around 0
			}
			pos[3 * i + a] = x;
		}
	}
	std::string dir = casedir("trilinear16");
	save_grid(dir, g);
	save_f64(dir, "pos", pos, {M, 3});
	const char *names[3] = {"B_periodic", "B_reflective", "B_clip"};
	for (int mode = 0; mode < 3; mode++) {
		g->setReflective(mode == 1);
		g->setClipVolume(mode == 2);
		std::vector<double> B(M * 3);
		for (size_t i = 0; i < M; i++) {
			Vector3d b = g->interpolate(Vector3d(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]));
			B[3 * i] = b.x; B[3 * i + 1] = b.y; B[3 * i + 2] = b.z;
		}
		save_f64(dir, names[mode], B, {M, 3});
	}
	save_meta(dir, "{\"n\": 16, \"spacing\": " + fmt(sp) + ", \"origin\": [" + fmt(origin.x) + ", " + fmt(origin.y) +
	                   ", " + fmt(origin.z) + "], \"grid_seed\": 1234, \"amp\": " + fmt(nG) + "}");
	printf("  trilinear16 M=%zu\n", M);
}

// ModulatedMagneticFieldGrid (src/magneticField/MagneticFieldGrid.cpp:20-55): a Grid3f times a
// Grid1f of other origin, spacing and size; the constructor makes the Grid3f periodic and the
// Grid1f reflective
static ref_ptr<Grid1f> make_mod_grid(int nx, int ny, int nz, double spacing, Vector3d origin, uint32_t seed) {
	ref_ptr<Grid1f> m = new Grid1f(origin, nx, ny, nz, spacing);
	Random rng(seed);
	for (int ix = 0; ix < nx; ix++)
		for (int iy = 0; iy < ny; iy++)
			for (int iz = 0; iz < nz; iz++)
				m->get(ix, iy, iz) = (float)rng.randUniform(0.2, 1.8);
	return m;
}

static void gen_modulated() {
	const int n = 16;
	const double sp = 50 * kpc;
	Vector3d origin(-3.1 * kpc, 7.25 * kpc, 0.0);
	ref_ptr<Grid3f> g = make_noise_grid(n, sp, origin, 1234, nG);
	Vector3d mo(11 * kpc, -40 * kpc, 5 * kpc);
	const double msp = 70 * kpc;
	ref_ptr<Grid1f> m = make_mod_grid(12, 10, 14, msp, mo, 4321);
	Random rng(98);
	const size_t M = 4096;
	const double L = n * sp;
	std::vector<double> pos(M * 3);
	for (size_t i = 0; i < M; i++)
		for (int a = 0; a < 3; a++) {
			double x;
			switch (i % 4) {
			case 0: x = origin.data[a] + rng.randUniform(0, L); break;
			case 1: x = origin.data[a] + rng.randUniform(-3 * L, 4 * L); break;
			case 2: x = mo.data[a] + msp * (int)rng.randUniform(-10, 30); break; // on modulation cell faces
			default: x = rng.randUniform(-1e3, 1e3) * L; break;
			}
			pos[3 * i + a] = x;
		}
	std::string dir = casedir("modulated16");
	save_grid(dir, g);
	std::vector<float> mv(12 * 10 * 14);
	for (int ix = 0; ix < 12; ix++)
		for (int iy = 0; iy < 10; iy++)
			for (int iz = 0; iz < 14; iz++)
				mv[(ix * 10 + iy) * 14 + iz] = m->get(ix, iy, iz);
	save_f32(dir, "mod", mv, {12, 10, 14});
	save_f64(dir, "pos", pos, {M, 3});
	// as constructed (grid periodic, modulation reflective); both periodic; modulation clipped
	ModulatedMagneticFieldGrid field(g, m);
	const char *names[3] = {"B_default", "B_periodic", "B_modclip"};
	for (int mode = 0; mode < 3; mode++) {
		if (mode == 1)
			field.setReflective(false, false);
		if (mode == 2) {
			field.setReflective(false, true);
			m->setClipVolume(true);
		}
		std::vector<double> B(M * 3);
		for (size_t i = 0; i < M; i++) {
			Vector3d b = field.getField(Vector3d(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]));
			B[3 * i] = b.x; B[3 * i + 1] = b.y; B[3 * i + 2] = b.z;
		}
		save_f64(dir, names[mode], B, {M, 3});
	}
	m->setClipVolume(false);
	save_meta(dir, "{\"n\": 16, \"spacing\": " + fmt(sp) + ", \"origin\": [" + fmt(origin.x) + ", " + fmt(origin.y) +
	                   ", " + fmt(origin.z) + "], \"mod_shape\": [12, 10, 14], \"mod_spacing\": " + fmt(msp) +
	                   ", \"mod_origin\": [" + fmt(mo.x) + ", " + fmt(mo.y) + ", " + fmt(mo.z) + "]}");
	printf("  modulated16 M=%zu\n", M);
}

// ---------------------------------------------------------------- JF12
// JF12Field samples (regular parts, striated from randomStriated(42), a 16^3 noise grid as the
// turbulent grid at 4 pc, all together, PlanckJF12bField) and one Cash-Karp trajectory case of
// charged nuclei from Earth through the full model.
static void gen_jf12() {
	Random rng(77);
	const size_t M = 4096;
	std::vector<double> pos(M * 3);
	for (size_t i = 0; i < M; i++) {
		double x, y, z;
		switch (i % 8) {
		case 0: case 1: case 2:
			x = rng.randUniform(-22, 22) * kpc; y = rng.randUniform(-22, 22) * kpc; z = rng.randUniform(-22, 22) * kpc;
			break;
		case 3: case 4: // the disk
			x = rng.randUniform(-20, 20) * kpc; y = rng.randUniform(-20, 20) * kpc; z = rng.randUniform(-1, 1) * kpc;
			break;
		case 5: { // the ring 3 < r < 5 kpc and inside 3 kpc
			double r = rng.randUniform(0, 5.5) * kpc, ph = rng.randUniform(-M_PI, M_PI);
			x = r * cos(ph); y = r * sin(ph); z = rng.randUniform(-2, 2) * kpc;
			break;
		}
		case 6: // z = 0 exactly, and on the z axis
			x = rng.randUniform(-20, 20) * kpc; y = rng.randUniform(-20, 20) * kpc; z = 0;
			if (i % 16 == 14) { x = 0; y = 0; z = rng.randUniform(-15, 15) * kpc; }
			break;
		default: // the X-field's inner region r < rc, near the plane
			x = rng.randUniform(-6, 6) * kpc; y = rng.randUniform(-6, 6) * kpc; z = rng.randUniform(-3, 3) * kpc;
			break;
		}
		pos[3 * i] = x; pos[3 * i + 1] = y; pos[3 * i + 2] = z;
	}
	std::string dir = casedir("jf12");
	save_f64(dir, "pos", pos, {M, 3});
	ref_ptr<Grid3f> turb = make_noise_grid(16, 4 * pc, Vector3d(0.), 555, 1.0);
	save_grid(dir, turb);
	auto sample = [&](const MagneticField &f, const char *name) {
		std::vector<double> B(M * 3);
		for (size_t i = 0; i < M; i++) {
			Vector3d b = f.getField(Vector3d(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]));
			B[3 * i] = b.x; B[3 * i + 1] = b.y; B[3 * i + 2] = b.z;
		}
		save_f64(dir, name, B, {M, 3});
	};
	JF12Field reg;
	sample(reg, "B_regular");
	const char *parts[3] = {"B_disk", "B_toroidal", "B_x"};
	for (int k = 0; k < 3; k++) {
		JF12Field f;
		f.setUseDiskField(k == 0);
		f.setUseToroidalHaloField(k == 1);
		f.setUseXField(k == 2);
		sample(f, parts[k]);
	}
	JF12Field str;
	str.randomStriated(42);
	sample(str, "B_striated");
	JF12Field tur;
	tur.setTurbulentGrid(turb);
	tur.setUseRegularField(false);
	sample(tur, "B_turbulent");
	JF12Field full;
	full.randomStriated(42);
	full.setTurbulentGrid(turb);
	sample(full, "B_full");
	PlanckJF12bField planck;
	planck.randomStriated(42);
	planck.setTurbulentGrid(turb);
	sample(planck, "B_planck");
	save_meta(dir, "{\"striated_seed\": 42, \"turb_n\": 16, \"turb_spacing\": " + fmt(4 * pc) + "}");
	printf("  jf12 M=%zu\n", M);

	// nuclei (A = 60, Z = 1: no mass table needed) from Earth, 1-10 EeV, through the full model
	std::vector<Particle> ps = make_particles(64, 31, 1 * EeV, 10 * EeV, Vector3d(-8.5 * kpc, 0, 0), {nucleusId(60, 1)}, {1.0});
	ref_ptr<JF12Field> fm = new JF12Field();
	fm->randomStriated(42);
	fm->setTurbulentGrid(turb);
	PropagationCK ck(fm, 1e-4, 0.1 * pc, 1 * kpc);
	MaximumTrajectoryLength ml(25 * kpc);
	run_traj("ck_jf12", ck, &ml, ps, 100000, {10, 50, 200, 100000},
	         "{\"propagator\": \"CK\", \"tolerance\": 1e-4, \"minStep\": " + fmt(0.1 * pc) + ", \"maxStep\": " +
	             fmt(1 * kpc) + ", \"maxTrajectoryLength\": " + fmt(25 * kpc) +
	             ", \"field\": \"jf12_full\", \"striated_seed\": 42}");
}

// ---------------------------------------------------------------- DiffusionSDE
// Serial run with Random::instance() seeded once (omp thread 0): particle i takes all its steps
// before particle i+1, [DiffusionSDE, MaximumTrajectoryLength] per step.
static void run_sde_case(const std::string &name, const DiffusionSDE &sde, double maxLen,
                         const std::vector<Particle> &ps, size_t maxSteps, uint32_t seed, const std::string &meta) {
	std::string dir = casedir(name);
	save_particles(dir, ps);
	const size_t N = ps.size();
	MaximumTrajectoryLength ml(maxLen);
	Random::instance().seed(seed);
	std::vector<double> pos(N * 3), d(N * 3), len(N), nxt(N), tim(N);
	std::vector<int32_t> act(N), nsteps(N);
	for (size_t i = 0; i < N; i++) {
		ParticleState st(ps[i].id, ps[i].E, ps[i].pos, ps[i].dir);
		ref_ptr<Candidate> c = new Candidate(st);
		size_t step = 0;
		while (step < maxSteps && c->isActive()) {
			sde.process(c);
			if (maxLen > 0)
				ml.process(c);
			step++;
		}
		Vector3d p = c->current.getPosition(), u = c->current.getDirection();
		pos[3 * i] = p.x; pos[3 * i + 1] = p.y; pos[3 * i + 2] = p.z;
		d[3 * i] = u.x; d[3 * i + 1] = u.y; d[3 * i + 2] = u.z;
		len[i] = c->getTrajectoryLength();
		nxt[i] = c->getNextStep();
		tim[i] = c->getTime();
		act[i] = c->isActive() ? 1 : 0;
		nsteps[i] = (int32_t)step;
	}
	save_f64(dir, "final_pos", pos, {N, 3});
	save_f64(dir, "final_dir", d, {N, 3});
	save_f64(dir, "final_len", len, {N});
	save_f64(dir, "final_next", nxt, {N});
	save_f64(dir, "final_time", tim, {N});
	save_i32(dir, "final_active", act, {N});
	save_i32(dir, "nsteps", nsteps, {N});
	save_meta(dir, meta);
	printf("  %-28s N=%zu steps<=%zu\n", name.c_str(), N, maxSteps);
}

static void gen_sde() {
	std::vector<int> species = {nucleusId(60, 1), nucleusId(60, 2), nucleusId(60, 26), nucleusId(60, 0)};
	std::vector<double> wts = {0.5, 0.2, 0.2, 0.1};
	{
		// uniform 1 nG field along z, epsilon 0.1
		ref_ptr<MagneticField> f = new UniformMagneticField(Vector3d(0, 0, 1 * nG));
		DiffusionSDE sde(f, 1e-4, 1 * pc, 10 * kpc, 0.1);
		std::vector<Particle> ps = make_particles(96, 70, 1 * TeV, 1 * PeV, Vector3d(0.), species, wts);
		run_sde_case("sde_uniform", sde, 50 * kpc, ps, 400, 4711,
		             "{\"field\": {\"type\": \"uniform\", \"B\": [0, 0, " + fmt(nG) + "]}, \"tolerance\": 1e-4, "
		             "\"minStep\": " + fmt(pc) + ", \"maxStep\": " + fmt(10 * kpc) + ", \"epsilon\": 0.1, \"alpha\": " +
		                 fmt(1. / 3.) + ", \"scale\": 1, \"maxTrajectoryLength\": " + fmt(50 * kpc) + ", \"seed\": 4711}");
	}
	{
		// 32^3 noise grid with 20 pc spacing: the field line curves within a step, the step search halves
		const double sp = 20 * pc;
		ref_ptr<Grid3f> g = make_noise_grid(32, sp, Vector3d(0.), 909, 1 * muG);
		std::string dir = casedir("sde_grid32");
		save_grid(dir, g);
		ref_ptr<MagneticField> f = new MagneticFieldGrid(g);
		DiffusionSDE sde(f, 1e-6, 0.1 * pc, 100 * pc, 0.3);
		sde.setAlpha(0.5);
		sde.setScale(0.2);
		Vector3d c(16 * sp);
		std::vector<Particle> ps = make_particles(64, 71, 1 * TeV, 10 * PeV, c, species, wts);
		run_sde_case("sde_grid32", sde, 2 * kpc, ps, 600, 4712,
		             "{\"grid_spacing\": " + fmt(sp) + ", \"tolerance\": 1e-6, \"minStep\": " + fmt(0.1 * pc) +
		                 ", \"maxStep\": " + fmt(100 * pc) + ", \"epsilon\": 0.3, \"alpha\": 0.5, \"scale\": 0.2, "
		                 "\"maxTrajectoryLength\": " + fmt(2 * kpc) + ", \"seed\": 4712}");
	}
	{
		// advection with a field, and neutral particles drifting
		ref_ptr<MagneticField> f = new UniformMagneticField(Vector3d(0.3 * nG, 0, 1 * nG));
		ref_ptr<AdvectionField> adv = new UniformAdvectionField(Vector3d(1e5, -2e5, 3e4));
		DiffusionSDE sde(f, adv, 1e-4, 1 * pc, 10 * kpc, 0.2);
		std::vector<Particle> ps = make_particles(64, 72, 1 * TeV, 1 * PeV, Vector3d(0.), species, wts);
		run_sde_case("sde_advection", sde, 20 * kpc, ps, 300, 4713,
		             "{\"field\": {\"type\": \"uniform\", \"B\": [" + fmt(0.3 * nG) + ", 0, " + fmt(nG) +
		                 "]}, \"advection\": [100000.0, -200000.0, 30000.0], \"tolerance\": 1e-4, \"minStep\": " +
		                 fmt(pc) + ", \"maxStep\": " + fmt(10 * kpc) + ", \"epsilon\": 0.2, \"alpha\": " + fmt(1. / 3.) +
		                 ", \"scale\": 1, \"maxTrajectoryLength\": " + fmt(20 * kpc) + ", \"seed\": 4713}");
	}
	{
		// galactic winds: ConstantSphericalAdvectionField and SphericalAdvectionField around an offset origin
		ref_ptr<MagneticField> f = new UniformMagneticField(Vector3d(0.3 * nG, 0, 1 * nG));
		const Vector3d o(1 * kpc, -0.5 * kpc, 0.2 * kpc);
		ref_ptr<AdvectionField> a1 = new ConstantSphericalAdvectionField(o, 3e5);
		ref_ptr<AdvectionField> a2 = new SphericalAdvectionField(o, 6 * kpc, 4e5, 2 * kpc, 1.5);
		const char *names[2] = {"sde_advection_cspherical", "sde_advection_spherical"};
		const std::string par[2] = {"[300000.0]", "[" + fmt(6 * kpc) + ", 400000.0, " + fmt(2 * kpc) + ", 1.5]"};
		for (int k = 0; k < 2; k++) {
			DiffusionSDE sde(f, k == 0 ? a1 : a2, 1e-4, 1 * pc, 10 * kpc, 0.2);
			std::vector<Particle> ps = make_particles(32, 81 + k, 1 * TeV, 1 * PeV, Vector3d(0.), species, wts);
			run_sde_case(names[k], sde, 20 * kpc, ps, 300, 4714 + k,
			             "{\"field\": {\"type\": \"uniform\", \"B\": [" + fmt(0.3 * nG) + ", 0, " + fmt(nG) +
			                 "]}, \"advection_kind\": " + std::to_string(k + 2) + ", \"advection_origin\": [" + fmt(o.x) +
			                 ", " + fmt(o.y) + ", " + fmt(o.z) + "], \"advection_par\": " + par[k] +
			                 ", \"tolerance\": 1e-4, \"minStep\": " + fmt(pc) + ", \"maxStep\": " + fmt(10 * kpc) +
			                 ", \"epsilon\": 0.2, \"alpha\": " + fmt(1. / 3.) + ", \"scale\": 1, \"maxTrajectoryLength\": " +
			                 fmt(20 * kpc) + ", \"seed\": " + std::to_string(4714 + k) + "}");
		}
	}
}

// Advection fields of src/advectionField/AdvectionField.cpp beyond the uniform and spherical winds: field
// samples of each class and DiffusionSDE runs in each (the parameters as crp_propagator_add_sde_advection
// orders them, in "advection_terms": [kind, [parameters]] per summed term).
static std::string adv_terms_json(const std::vector<std::pair<int, std::vector<double>>> &terms) {
	std::string s = "[";
	for (size_t t = 0; t < terms.size(); t++) {
		s += (t ? ", [" : "[") + std::to_string(terms[t].first) + ", [";
		for (size_t i = 0; i < terms[t].second.size(); i++)
			s += (i ? ", " : "") + fmt(terms[t].second[i]);
		s += "]]";
	}
	return s + "]";
}

static void gen_advection() {
	typedef std::vector<std::pair<int, std::vector<double>>> Terms;
	const Vector3d o(1 * kpc, -0.5 * kpc, 0.2 * kpc);
	ref_ptr<AdvectionField> cart = new OneDimensionalCartesianShock(4., 1e6, 0.5 * kpc);
	ref_ptr<AdvectionField> sph1c = new OneDimensionalSphericalShock(3 * kpc, 1e6, 4., 0.5 * kpc, true);
	ref_ptr<AdvectionField> sph1 = new OneDimensionalSphericalShock(3 * kpc, 1e6, 4., 0.5 * kpc, false);
	ref_ptr<AdvectionField> obl = new ObliqueAdvectionShock(3., 8e5, 2e5, 0.3 * kpc);
	ref_ptr<SphericalAdvectionShock> sas = new SphericalAdvectionShock(o, 2 * kpc, 6e5, 0.3 * kpc);
	sas->setRRot(3 * kpc);
	sas->setAzimuthalSpeed(2e5);
	ref_ptr<AdvectionFieldList> lst = new AdvectionFieldList();
	lst->addField(new UniformAdvectionField(Vector3d(1e5, -2e5, 3e4)));
	lst->addField(sph1);
	lst->addField(new ConstantSphericalAdvectionField(o, 3e5));
	const Terms tCart = {{4, {4., 1e6, 0.5 * kpc}}};
	const Terms tSph1c = {{5, {3 * kpc, 1e6, 4., 0.5 * kpc, 1}}};
	const Terms tSph1 = {{5, {3 * kpc, 1e6, 4., 0.5 * kpc, 0}}};
	const Terms tObl = {{6, {3., 8e5, 2e5, 0.3 * kpc}}};
	const Terms tSas = {{7, {o.x, o.y, o.z, 2 * kpc, 6e5, 0.3 * kpc, 3 * kpc, 2e5}}};
	const Terms tLst = {{1, {1e5, -2e5, 3e4}}, tSph1[0], {2, {o.x, o.y, o.z, 3e5}}};
	const char *names[6] = {"cartesian_shock", "spherical_shock_1d_cool", "spherical_shock_1d", "oblique_shock",
	                        "spherical_advection_shock", "list"};
	const AdvectionField *fields[6] = {cart.get(), sph1c.get(), sph1.get(), obl.get(), sas.get(), lst.get()};
	const Terms *terms[6] = {&tCart, &tSph1c, &tSph1, &tObl, &tSas, &tLst};
	{
		Random rng(811);
		const size_t M = 2048;
		std::vector<double> pos(M * 3);
		for (size_t i = 0; i < M; i++) {
			Vector3d p;
			switch (i % 4) {
			case 0: p = rng.randVector() * rng.randUniform(0, 6 * kpc); break;           // around 0
			case 1: p = o + rng.randVector() * rng.randUniform(1.5 * kpc, 2.5 * kpc); break; // near r0
			case 2: p = rng.randVector() * rng.randUniform(2.5 * kpc, 3.5 * kpc); break;     // near rShock
			default: p = Vector3d(rng.randUniform(-1, 1) * kpc, rng.randUniform(-5, 5) * kpc, 0); break;
			}
			pos[3 * i] = p.x; pos[3 * i + 1] = p.y; pos[3 * i + 2] = p.z;
		}
		std::string dir = casedir("advection_fields");
		save_f64(dir, "pos", pos, {M, 3});
		std::string meta = "{";
		for (int k = 0; k < 6; k++) {
			std::vector<double> v(M * 3);
			for (size_t i = 0; i < M; i++) {
				const Vector3d w = fields[k]->getField(Vector3d(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]));
				v[3 * i] = w.x; v[3 * i + 1] = w.y; v[3 * i + 2] = w.z;
			}
			save_f64(dir, std::string("v_") + names[k], v, {M, 3});
			meta += std::string(k ? ", " : "") + "\"" + names[k] + "\": " + adv_terms_json(*terms[k]);
		}
		save_meta(dir, meta + "}");
		printf("  advection_fields M=%zu\n", M);
	}
	std::vector<int> species = {nucleusId(60, 1), nucleusId(60, 2), nucleusId(60, 26), nucleusId(60, 0)};
	std::vector<double> wts = {0.5, 0.2, 0.2, 0.1};
	ref_ptr<MagneticField> f = new UniformMagneticField(Vector3d(0.3 * nG, 0, 1 * nG));
	const Vector3d starts[6] = {Vector3d(0.2 * kpc, 0, 0), Vector3d(2.8 * kpc, 0.5 * kpc, 0),
	                            Vector3d(2.8 * kpc, 0.5 * kpc, 0), Vector3d(0.1 * kpc, 0, 0),
	                            o + Vector3d(1.9 * kpc, 0, 0), Vector3d(2.8 * kpc, 0.5 * kpc, 0)};
	const ref_ptr<AdvectionField> refs[6] = {cart, sph1c, sph1, obl, sas, lst};
	for (int k = 0; k < 6; k++) {
		DiffusionSDE sde(f, refs[k], 1e-4, 1 * pc, 10 * kpc, 0.2);
		std::vector<Particle> ps = make_particles(32, 90 + k, 1 * TeV, 1 * PeV, starts[k], species, wts);
		const std::string name = std::string("sde_adv_") + names[k];
		run_sde_case(name, sde, 20 * kpc, ps, 300, 4720 + k,
		             "{\"field\": {\"type\": \"uniform\", \"B\": [" + fmt(0.3 * nG) + ", 0, " + fmt(nG) +
		                 "]}, \"advection_terms\": " + adv_terms_json(*terms[k]) + ", \"tolerance\": 1e-4, \"minStep\": " +
		                 fmt(pc) + ", \"maxStep\": " + fmt(10 * kpc) + ", \"epsilon\": 0.2, \"alpha\": " + fmt(1. / 3.) +
		                 ", \"scale\": 1, \"maxTrajectoryLength\": " + fmt(20 * kpc) + ", \"seed\": " +
		                 std::to_string(4720 + k) + "}");
	}
}

static void gen_planewave_case(const std::string &name, const TurbulenceSpectrum &spec, int Nm, int seed,
                               double posScale, const std::string &meta) {
	PlaneWaveTurbulence pw(spec, Nm, seed);
	Random rng(7);
	const size_t M = 2048;
	std::vector<double> pos(M * 3), B(M * 3);
	for (size_t i = 0; i < M; i++) {
		Vector3d p = rng.randVector() * rng.randUniform(0, posScale);
		Vector3d b = pw.getField(p);
		pos[3 * i] = p.x; pos[3 * i + 1] = p.y; pos[3 * i + 2] = p.z;
		B[3 * i] = b.x; B[3 * i + 1] = b.y; B[3 * i + 2] = b.z;
	}
	std::string dir = casedir(name);
	save_f64(dir, "pos", pos, {M, 3});
	save_f64(dir, "B", B, {M, 3});
	save_meta(dir, meta);
	printf("  %-28s Nm=%d\n", name.c_str(), Nm);
}

static std::string grid_meta(const std::string &extra, int n, double sp, uint32_t seed, double amp) {
	return "{\"grid\": {\"n\": " + std::to_string(n) + ", \"spacing\": " + fmt(sp) +
	       ", \"origin\": [0, 0, 0], \"seed\": " + std::to_string(seed) + ", \"amp\": " + fmt(amp) + "}, " + extra +
	       "}";
}

// Grid tools on a non-cubic 6x5x4 grid with non-uniform spacing: raw and text
// dumps (files kept as fixtures), loads with a conversion factor, the field
// statistics and a resampling through MagneticFieldGrid (fromMagneticField).
static void gen_gridtools() {
	std::string d = casedir("gridtools");
	ref_ptr<Grid3f> g = new Grid3f(Vector3d(-1 * Mpc, 0.5 * Mpc, 2 * Mpc), 6, 5, 4, Vector3d(50 * kpc, 40 * kpc, 30 * kpc));
	Random rng(77);
	for (int ix = 0; ix < 6; ix++)
		for (int iy = 0; iy < 5; iy++)
			for (int iz = 0; iz < 4; iz++) {
				Vector3f &b = g->get(ix, iy, iz);
				b.x = rng.randUniform(-1, 1) * nG;
				b.y = rng.randUniform(-1, 1) * nG + 0.3 * nG;
				b.z = rng.randUniform(-1, 1) * nG;
			}
	save_grid(d, g);
	dumpGrid(g, d + "/dump_c1.raw", 1);
	dumpGrid(g, d + "/dump_c3.raw", 3.5);
	dumpGridToTxt(g, d + "/dump_props.txt", 1 / nG, true);
	dumpGridToTxt(g, d + "/dump_plain.txt", 1 / nG, false);
	{
		ref_ptr<Grid3f> h = new Grid3f(g->getOrigin(), 6, 5, 4, g->getSpacing());
		loadGrid(h, d + "/dump_c1.raw", 0.25);
		save_grid(d, h);
		rename((d + "/grid.npy").c_str(), (d + "/loaded_raw_c025.npy").c_str());
	}
	{
		ref_ptr<Grid3f> h = loadGrid3fFromTxt(d + "/dump_props.txt", nG);
		save_grid(d, h);
		rename((d + "/grid.npy").c_str(), (d + "/loaded_txt.npy").c_str());
	}
	{
		ref_ptr<Grid3f> h = new Grid3f(*g);
		scaleGrid(h, 2.5);
		save_grid(d, h);
		rename((d + "/grid.npy").c_str(), (d + "/scaled_25.npy").c_str());
	}
	{
		// resample the (periodic) grid on a finer, shifted 7x6x5 grid
		ref_ptr<Grid3f> r = new Grid3f(Vector3d(-0.9 * Mpc, 0.52 * Mpc, 2.01 * Mpc), 7, 6, 5, Vector3d(33 * kpc));
		fromMagneticField(r, new MagneticFieldGrid(g));
		save_grid(d, r);
		rename((d + "/grid.npy").c_str(), (d + "/resampled.npy").c_str());
	}
	save_grid(d, g);
	Vector3f mv = meanFieldVector(g);
	std::array<float, 3> pa = rmsFieldStrengthPerAxis(g);
	save_f32(d, "mean_vector", {mv.x, mv.y, mv.z}, {3});
	save_f32(d, "rms_per_axis", {pa[0], pa[1], pa[2]}, {3});
	save_f64(d, "strengths", {meanFieldStrength(g), rmsFieldStrength(g)}, {2});
	save_meta(d, "{\"origin\": [" + fmt(-1 * Mpc) + ", " + fmt(0.5 * Mpc) + ", " + fmt(2 * Mpc) +
	                 "], \"n\": [6, 5, 4], \"spacing\": [" + fmt(50 * kpc) + ", " + fmt(40 * kpc) + ", " +
	                 fmt(30 * kpc) + "], \"seed\": 77, \"dump_c3\": 3.5, \"txt_conversion\": " + fmt(1 / nG) +
	                 ", \"load_raw_conversion\": 0.25, \"scale\": 2.5, \"resampled\": {\"origin\": [" +
	                 fmt(-0.9 * Mpc) + ", " + fmt(0.52 * Mpc) + ", " + fmt(2.01 * Mpc) + "], \"n\": [7, 6, 5], \"spacing\": " +
	                 fmt(33 * kpc) + "}}");
}

// ---------------------------------------------------------------- field decorators
// PeriodicMagneticField (periodic and reflective), MagneticFieldList and MagneticFieldEvolution
// (src/magneticField/MagneticField.cpp:41-79) around the 32^3 noise grid of the trajectory cases:
// field samples at 4096 positions (getField(position), and getField(position, z) with a redshift
// per sample), and trajectories through them.
static void gen_decorators() {
	const int n = 32;
	const double sp = 50 * kpc, L = n * sp;
	ref_ptr<Grid3f> g32 = make_noise_grid(n, sp, Vector3d(0.), 2024, nG);
	ref_ptr<MagneticField> fgrid = new MagneticFieldGrid(g32);
	ref_ptr<MagneticField> funi = new UniformMagneticField(Vector3d(0.3 * nG, -0.2 * nG, 0.5 * nG));
	SimpleSpectrum spec2(3 * nG, 10 * kpc, 2 * Mpc, 11. / 3);
	ref_ptr<MagneticField> fpw = new PlaneWaveTurbulence(spec2, 2, 5); // the planewave_simple2 table
	const Vector3d ext(1.3 * L, 0.7 * L, L), org(-0.2 * L, 0.1 * L, 0.);
	ref_ptr<MagneticField> fper = new PeriodicMagneticField(fgrid, ext, org, false);
	ref_ptr<MagneticField> fref = new PeriodicMagneticField(fgrid, ext, org, true);
	ref_ptr<MagneticFieldList> flist = new MagneticFieldList();
	flist->addField(fgrid);
	flist->addField(funi);
	flist->addField(fpw);
	ref_ptr<MagneticField> fevo = new MagneticFieldEvolution(fgrid, 1.7);
	// nested: a list of a reflective periodic grid, an evolution (its one-argument getField is the base
	// class's zero) and a periodic list
	ref_ptr<MagneticFieldList> inner = new MagneticFieldList();
	inner->addField(fgrid);
	inner->addField(funi);
	ref_ptr<MagneticFieldList> nested = new MagneticFieldList();
	nested->addField(fref);
	nested->addField(new MagneticFieldEvolution(fgrid, 2.0));
	nested->addField(new PeriodicMagneticField(inner, Vector3d(0.8 * L, 1.1 * L, 0.6 * L)));
	ref_ptr<MagneticField> fevolist = new MagneticFieldEvolution(inner, -0.5);
	// the propagators call getField(position, z) through MagneticField (the list's own getField hides it)
	ref_ptr<MagneticField> fnested = nested;

	Random rng(99);
	const size_t M = 4096;
	std::vector<double> pos(3 * M), z(M);
	std::vector<double> Bper(3 * M), Bref(3 * M), Blist(3 * M), Bevo(3 * M), Bevo1(3 * M), Bnest(3 * M), Bevol(3 * M);
	for (size_t i = 0; i < M; i++) {
		Vector3d p(rng.randUniform(-2 * L, 3 * L), rng.randUniform(-2 * L, 3 * L), rng.randUniform(-2 * L, 3 * L));
		if (i < 64) // on period boundaries
			p = org + Vector3d(ext.x * (double)((int)i % 4 - 1), ext.y * (double)((int)i / 4 % 4 - 1), ext.z * (double)((int)i / 16 - 1));
		z[i] = rng.randUniform(0, 3);
		pos[3 * i] = p.x; pos[3 * i + 1] = p.y; pos[3 * i + 2] = p.z;
		const Vector3d b[7] = {fper->getField(p), fref->getField(p), flist->getField(p), fevo->getField(p, z[i]),
		                       fevo->getField(p), fnested->getField(p, z[i]), fevolist->getField(p, z[i])};
		std::vector<double> *out[7] = {&Bper, &Bref, &Blist, &Bevo, &Bevo1, &Bnest, &Bevol};
		for (int k = 0; k < 7; k++) {
			(*out[k])[3 * i] = b[k].x;
			(*out[k])[3 * i + 1] = b[k].y;
			(*out[k])[3 * i + 2] = b[k].z;
		}
	}
	{
		std::string d = casedir("decorators32");
		save_f64(d, "pos", pos, {M, 3});
		save_f64(d, "z", z, {M});
		save_f64(d, "B_periodic", Bper, {M, 3});
		save_f64(d, "B_reflective", Bref, {M, 3});
		save_f64(d, "B_list", Blist, {M, 3});
		save_f64(d, "B_evolution", Bevo, {M, 3});
		save_f64(d, "B_evolution_noz", Bevo1, {M, 3});
		save_f64(d, "B_nested", Bnest, {M, 3});
		save_f64(d, "B_evolution_list", Bevol, {M, 3});
		save_meta(d, "{\"grid\": \"grid32\", \"uniform\": [" + fmt(0.3 * nG) + ", " + fmt(-0.2 * nG) + ", " +
		                 fmt(0.5 * nG) + "], \"planewave\": \"planewave_simple2\", \"extends\": [" + fmt(ext.x) +
		                 ", " + fmt(ext.y) + ", " + fmt(ext.z) + "], \"origin\": [" + fmt(org.x) + ", " + fmt(org.y) +
		                 ", " + fmt(org.z) + "], \"evolution_m\": 1.7, \"nested_evolution_m\": 2.0, "
		                 "\"nested_extends\": [" + fmt(0.8 * L) + ", " + fmt(1.1 * L) + ", " + fmt(0.6 * L) +
		                 "], \"evolution_list_m\": -0.5}");
		printf("  %-28s M=%zu\n", "decorators32", M);
	}

	const Vector3d centre(L / 2, L / 2, L / 2);
	std::vector<int> proton = {nucleusId(60, 1)};
	std::vector<int> species = {nucleusId(60, 1), nucleusId(60, 2), nucleusId(60, 7), nucleusId(60, 14), nucleusId(60, 26)};
	std::vector<double> one = {1.0}, wts = {0.5, 0.2, 0.15, 0.1, 0.05};
	{
		ref_ptr<MagneticFieldList> gl = new MagneticFieldList();
		gl->addField(fgrid);
		gl->addField(new UniformMagneticField(Vector3d(1 * nG, 0, 0)));
		std::vector<Particle> ps = make_particles(128, 60, 1 * EeV, 100 * EeV, centre, proton, one);
		PropagationBP bp(gl, 50 * kpc);
		run_traj("bp_fixed_list32", bp, nullptr, ps, 400, {1, 10, 100, 400},
		         "{\"propagator\": \"BP\", \"fixedStep\": " + fmt(50 * kpc) +
		             ", \"field\": \"list(grid32, uniform [1 nG, 0, 0])\", \"particle_seed\": 60}");
	}
	{
		ref_ptr<MagneticField> pr = new PeriodicMagneticField(fgrid, Vector3d(1.2 * L, 0.9 * L, 1.1 * L),
		                                                      Vector3d(-0.1 * L, 0.05 * L, 0.), true);
		std::vector<Particle> ps = make_particles(128, 61, 1 * EeV, 300 * EeV, centre, species, wts);
		PropagationCK ck(pr, 1e-4, 1 * kpc, 1 * Mpc);
		MaximumTrajectoryLength ml(6 * Mpc);
		run_traj("ck_reflective32", ck, &ml, ps, 3000, {1, 10, 100, 1000, 3000},
		         "{\"propagator\": \"CK\", \"tolerance\": 1e-4, \"minStep\": " + fmt(kpc) + ", \"maxStep\": " +
		             fmt(Mpc) + ", \"maxTrajectoryLength\": " + fmt(6 * Mpc) +
		             ", \"field\": \"periodic(grid32) reflective\", \"extends\": [" + fmt(1.2 * L) + ", " +
		             fmt(0.9 * L) + ", " + fmt(1.1 * L) + "], \"origin\": [" + fmt(-0.1 * L) + ", " + fmt(0.05 * L) +
		             ", 0], \"particle_seed\": 61}");
	}
	{
		ref_ptr<MagneticField> ev = new MagneticFieldEvolution(fgrid, 1.5);
		std::vector<Particle> ps = make_particles(128, 62, 1 * EeV, 100 * EeV, centre, proton, one);
		Random rz(63);
		std::vector<double> zs(ps.size());
		for (size_t i = 0; i < zs.size(); i++)
			zs[i] = rz.randUniform(0, 2);
		g_redshift = &zs;
		PropagationBP bp(ev, 50 * kpc);
		run_traj("bp_fixed_evolution32", bp, nullptr, ps, 400, {1, 10, 100, 400},
		         "{\"propagator\": \"BP\", \"fixedStep\": " + fmt(50 * kpc) +
		             ", \"field\": \"evolution(grid32, m=1.5)\", \"evolution_m\": 1.5, \"particle_seed\": 62}");
		g_redshift = nullptr;
	}
}

// ---------------------------------------------------------------- tricubic / nearest neighbour
#ifdef HAVE_SIMD
// Grid::setInterpolationType (include/crpropa/Grid.h:255-264) with TRICUBIC (Grid3f: Grid.h:408-492,
// Grid1f: :499-536) and NEAREST_NEIGHBOUR (closestValue, :383-405).  Only a build with HAVE_SIMD
// (SIMD_EXTENSIONS "avx": -DHAVE_SIMD -msse ... -mavx, no -mfma) evaluates tricubic Grid3f; the
// target ref-simd of oracle/Makefile builds this generator that way and runs it with "interp".
static void fill_noise(Grid3f &g, uint32_t seed, double amp) {
	Random rng(seed);
	for (size_t ix = 0; ix < g.getNx(); ix++)
		for (size_t iy = 0; iy < g.getNy(); iy++)
			for (size_t iz = 0; iz < g.getNz(); iz++) {
				Vector3f &b = g.get(ix, iy, iz);
				b.x = rng.randUniform(-1, 1) * amp;
				b.y = rng.randUniform(-1, 1) * amp;
				b.z = rng.randUniform(-1, 1) * amp;
			}
}

static void gen_interp() {
	// field samples: a 9 x 8 x 7 Grid3f and a 6 x 5 x 7 Grid1f, both with spacing vectors
	{
		const size_t nx = 9, ny = 8, nz = 7, mx = 6, my = 5, mz = 7;
		const Vector3d sp(0.7 * kpc, 0.9 * kpc, 1.1 * kpc), origin(-1.5 * kpc, 0.25 * kpc, 2 * kpc);
		const Vector3d msp(1.3 * kpc, 0.8 * kpc, 1.0 * kpc), mo(0.4 * kpc, -2 * kpc, 1.5 * kpc);
		ref_ptr<Grid3f> g = new Grid3f(origin, nx, ny, nz, sp);
		fill_noise(*g, 808, nG);
		ref_ptr<Grid1f> m = new Grid1f(mo, mx, my, mz, msp);
		{
			Random rm(809);
			for (size_t ix = 0; ix < mx; ix++)
				for (size_t iy = 0; iy < my; iy++)
					for (size_t iz = 0; iz < mz; iz++)
						m->get(ix, iy, iz) = (float)rm.randUniform(0.2, 1.8);
		}
		Random rng(810);
		const size_t M = 4096;
		std::vector<double> pos(M * 3);
		const Vector3d L(nx * sp.x, ny * sp.y, nz * sp.z);
		for (size_t i = 0; i < M; i++)
			for (int a = 0; a < 3; a++) {
				const double o = origin.data[a], l = L.data[a], h = sp.data[a];
				const int n = (int)(l / h + 0.5);
				double x;
				switch (i % 8) {
				case 0: x = o + rng.randUniform(0, l); break;                          // inside
				case 1: x = o + rng.randUniform(-3 * l, 4 * l); break;                 // a few periods out
				case 2: x = o + h / 2 + h * (int)rng.randUniform(-n, 2 * n); break;    // on grid points
				case 3: x = o + h * (int)rng.randUniform(-n, 2 * n); break;            // half-way between points
				case 4: x = o + rng.randUniform(-h, h); break;                         // near the lower edge
				case 5: x = o + l + rng.randUniform(-h, h); break;                     // near the upper edge
				case 6: x = o + rng.randUniform(-300 * l, 300 * l); break;             // far away
				default: x = mo.data[a] + msp.data[a] * (0.5 + (int)rng.randUniform(-8, 16)); break; // Grid1f points
				}
				pos[3 * i + a] = x;
			}
		std::string dir = casedir("interp_grids");
		save_grid(dir, g);
		std::vector<float> mv(mx * my * mz);
		for (size_t i = 0; i < mv.size(); i++)
			mv[i] = m->getGrid()[i];
		save_f32(dir, "mod", mv, {mx, my, mz});
		save_f64(dir, "pos", pos, {M, 3});
		struct Mode {
			const char *name;
			interpolationType t;
			bool reflective, clip;
		} modes[5] = {{"tricubic_periodic", TRICUBIC, false, false},
		              {"tricubic_reflective", TRICUBIC, true, false},
		              {"tricubic_clip", TRICUBIC, false, true},
		              {"nearest_periodic", NEAREST_NEIGHBOUR, false, false},
		              {"nearest_reflective", NEAREST_NEIGHBOUR, true, false}};
		for (const Mode &md : modes) {
			g->setInterpolationType(md.t);
			g->setReflective(md.reflective);
			g->setClipVolume(md.clip);
			m->setInterpolationType(md.t);
			m->setReflective(md.reflective);
			m->setClipVolume(md.clip);
			MagneticFieldGrid field(g);
			std::vector<double> B(M * 3);
			std::vector<float> S(M);
			for (size_t i = 0; i < M; i++) {
				const Vector3d x(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
				const Vector3d b = field.getField(x);
				B[3 * i] = b.x; B[3 * i + 1] = b.y; B[3 * i + 2] = b.z;
				S[i] = m->interpolate(x);
			}
			save_f64(dir, std::string("B_") + md.name, B, {M, 3});
			save_f32(dir, std::string("m_") + md.name, S, {M});
		}
		save_meta(dir, "{\"shape\": [9, 8, 7], \"spacing\": [" + fmt(sp.x) + ", " + fmt(sp.y) + ", " + fmt(sp.z) +
		                   "], \"origin\": [" + fmt(origin.x) + ", " + fmt(origin.y) + ", " + fmt(origin.z) +
		                   "], \"mod_shape\": [6, 5, 7], \"mod_spacing\": [" + fmt(msp.x) + ", " + fmt(msp.y) + ", " +
		                   fmt(msp.z) + "], \"mod_origin\": [" + fmt(mo.x) + ", " + fmt(mo.y) + ", " + fmt(mo.z) +
		                   "], \"grid_seed\": 808, \"mod_seed\": 809, \"pos_seed\": 810, \"amp\": " + fmt(nG) +
		                   ", \"build\": \"HAVE_SIMD, -mavx, no FMA\"}");
		printf("  interp_grids M=%zu\n", M);
	}

	// trajectories through the shared 32^3 grid (tests/golden/grid32) read tricubic or nearest
	const int n = 32;
	const double sp = 50 * kpc;
	ref_ptr<Grid3f> g32 = make_noise_grid(n, sp, Vector3d(0.), 2024, nG);
	Vector3d centre(n * sp / 2, n * sp / 2, n * sp / 2);
	std::vector<int> proton = {nucleusId(60, 1)};
	std::vector<int> species = {nucleusId(60, 1), nucleusId(60, 2), nucleusId(60, 7), nucleusId(60, 14), nucleusId(60, 26)};
	std::vector<double> one = {1.0}, wts = {0.5, 0.2, 0.15, 0.1, 0.05};
	{
		ref_ptr<Grid3f> gt = new Grid3f(*g32);
		gt->setInterpolationType(TRICUBIC);
		ref_ptr<MagneticField> f = new MagneticFieldGrid(gt);
		std::vector<Particle> ps = make_particles(128, 70, 1 * EeV, 300 * EeV, centre, species, wts);
		PropagationCK ck(f, 1e-4, 1 * kpc, 1 * Mpc);
		MaximumTrajectoryLength ml(10 * Mpc);
		run_traj("ck_grid32_tricubic", ck, &ml, ps, 3000, {1, 10, 100, 1000, 3000},
		         "{\"propagator\": \"CK\", \"tolerance\": 1e-4, \"minStep\": " + fmt(kpc) + ", \"maxStep\": " +
		             fmt(Mpc) + ", \"maxTrajectoryLength\": " + fmt(10 * Mpc) +
		             ", \"field\": \"grid32\", \"interpolation\": \"TRICUBIC\", \"particle_seed\": 70}");
	}
	{
		ref_ptr<Grid3f> gt = new Grid3f(*g32);
		gt->setInterpolationType(TRICUBIC);
		gt->setReflective(true);
		ref_ptr<MagneticField> f = new MagneticFieldGrid(gt);
		std::vector<Particle> ps = make_particles(128, 71, 1 * EeV, 100 * EeV, centre, proton, one);
		PropagationBP bp(f, 0.1 * Mpc);
		run_traj("bp_fixed_grid32_tricubic_reflective", bp, nullptr, ps, 600, {1, 10, 100, 600},
		         "{\"propagator\": \"BP\", \"fixedStep\": " + fmt(0.1 * Mpc) +
		             ", \"field\": \"grid32\", \"interpolation\": \"TRICUBIC\", \"reflective\": true, "
		             "\"particle_seed\": 71}");
	}
	{
		ref_ptr<Grid3f> gn = new Grid3f(*g32);
		gn->setInterpolationType(NEAREST_NEIGHBOUR);
		gn->setReflective(true);
		ref_ptr<MagneticField> f = new MagneticFieldGrid(gn);
		std::vector<Particle> ps = make_particles(128, 72, 1 * EeV, 100 * EeV, centre, proton, one);
		PropagationBP bp(f, 0.1 * Mpc);
		run_traj("bp_fixed_grid32_nearest_reflective", bp, nullptr, ps, 600, {1, 10, 100, 600},
		         "{\"propagator\": \"BP\", \"fixedStep\": " + fmt(0.1 * Mpc) +
		             ", \"field\": \"grid32\", \"interpolation\": \"NEAREST_NEIGHBOUR\", \"reflective\": true, "
		             "\"particle_seed\": 72}");
	}
	{
		// ModulatedMagneticFieldGrid over grid32 (periodic, tricubic) and a reflective 9x11x10 Grid1f
		// (tricubic), as bp_fixed_grid32_modulated builds them otherwise
		Vector3d mo(-0.3 * Mpc, 0.1 * Mpc, 0.25 * Mpc);
		const double msp = 0.2 * Mpc;
		ref_ptr<Grid1f> mg = make_mod_grid(9, 11, 10, msp, mo, 777);
		ref_ptr<Grid3f> gt = new Grid3f(*g32);
		gt->setInterpolationType(TRICUBIC);
		mg->setInterpolationType(TRICUBIC);
		ref_ptr<MagneticField> f = new ModulatedMagneticFieldGrid(gt, mg);
		std::vector<Particle> ps = make_particles(128, 73, 1 * EeV, 300 * EeV, centre, species, wts);
		PropagationCK ck(f, 1e-4, 1 * kpc, 1 * Mpc);
		MaximumTrajectoryLength ml(8 * Mpc);
		run_traj("ck_grid32_modulated_tricubic", ck, &ml, ps, 3000, {1, 10, 100, 1000, 3000},
		         "{\"propagator\": \"CK\", \"tolerance\": 1e-4, \"minStep\": " + fmt(kpc) + ", \"maxStep\": " +
		             fmt(Mpc) + ", \"maxTrajectoryLength\": " + fmt(8 * Mpc) +
		             ", \"field\": \"grid32_modulated\", \"interpolation\": \"TRICUBIC\", \"mod_shape\": [9, 11, 10], "
		             "\"mod_reflective\": true, \"mod_spacing\": " + fmt(msp) + ", \"mod_origin\": [" + fmt(mo.x) + ", " + fmt(mo.y) + ", " +
		             fmt(mo.z) + "], \"mod_seed\": 777, \"particle_seed\": 73}");
		std::vector<float> mv(9 * 11 * 10);
		for (size_t i = 0; i < mv.size(); i++)
			mv[i] = mg->getGrid()[i];
		save_f32(casedir("ck_grid32_modulated_tricubic"), "mod", mv, {9, 11, 10});
	}
}
#endif

int main(int argc, char **argv) {
	g_root = argc > 1 ? argv[1] : "tests/golden";
	mkdir(g_root.c_str(), 0755);
	if (argc > 2 && std::string(argv[2]) == "decorators") { // only the field-decorator cases
		gen_decorators();
		return 0;
	}
	if (argc > 2 && std::string(argv[2]) == "advection") { // only the advection-field cases
		gen_advection();
		return 0;
	}
	if (argc > 2 && std::string(argv[2]) == "interp") { // only the tricubic / nearest-neighbour cases
#ifdef HAVE_SIMD
		gen_interp();
		FILE *f = fopen((g_root + "/PROVENANCE.txt").c_str(), "a");
		fprintf(f, "interp_grids, *_tricubic*, *_nearest*: the same generator built with the reference's SIMD_EXTENSIONS=avx "
		           "flags (-DHAVE_SIMD -msse -msse2 -msse3 -msse4.1 -msse4.2 -mavx, no -mfma; oracle/Makefile target "
		           "ref-simd), g++ %s\n",
		        __VERSION__);
		fclose(f);
		return 0;
#else
		fprintf(stderr, "gen_golden: the interp cases need a build with HAVE_SIMD (oracle/Makefile target ref-simd)\n");
		return 1;
#endif
	}
	printf("gen_golden: glibc %s, g++ %s\n", gnu_get_libc_version(), __VERSION__);
	{
		FILE *f = fopen((g_root + "/PROVENANCE.txt").c_str(), "w");
		fprintf(f, "Generated by tests/golden/gen_golden.cpp linked against the reference C++ (oracle/Makefile target ref)\n");
		fprintf(f, "compiler: g++ %s\nflags: -O3 -DNDEBUG -std=c++11 -fopenmp (no -march: no FMA)\nglibc: %s\n",
		        __VERSION__, gnu_get_libc_version());
		fclose(f);
	}

	gen_trilinear();
	gen_modulated();
	gen_sde();
	gen_advection();
	gen_jf12();
	gen_gridtools();
	gen_decorators();

	// Plane waves: the C3 mode table (Simple spectrum), the reference test's
	// general spectrum (test/testPropagation.cpp:492), and a 2-mode table.
	SimpleSpectrum specC3(1 * muG, 1 * pc, 100 * pc);
	gen_planewave_case("planewave_c3", specC3, 256, 11, 500 * pc,
	                   "{\"spectrum\": \"simple\", \"Brms\": " + fmt(muG) + ", \"lMin\": " + fmt(pc) +
	                       ", \"lMax\": " + fmt(100 * pc) + ", \"sIndex\": " + fmt(5. / 3) +
	                       ", \"Nm\": 256, \"seed\": 11}");
	TurbulenceSpectrum specGen(gauss, pc, 100 * pc);
	gen_planewave_case("planewave_general10", specGen, 10, 1, 50 * pc,
	                   "{\"spectrum\": \"general\", \"Brms\": " + fmt(gauss) + ", \"lMin\": " + fmt(pc) +
	                       ", \"lMax\": " + fmt(100 * pc) + ", \"lBendover\": 1, \"sIndex\": " + fmt(5. / 3) +
	                       ", \"qIndex\": 4, \"Nm\": 10, \"seed\": 1}");
	SimpleSpectrum spec2(3 * nG, 10 * kpc, 2 * Mpc, 11. / 3);
	gen_planewave_case("planewave_simple2", spec2, 2, 5, 5 * Mpc,
	                   "{\"spectrum\": \"simple\", \"Brms\": " + fmt(3 * nG) + ", \"lMin\": " + fmt(10 * kpc) +
	                       ", \"lMax\": " + fmt(2 * Mpc) + ", \"sIndex\": " + fmt(11. / 3) +
	                       ", \"Nm\": 2, \"seed\": 5}");

	// ---- grid trajectories (32^3, spacing 50 kpc, white noise 1 nG, periodic)
	const int n = 32;
	const double sp = 50 * kpc;
	const uint32_t gseed = 2024;
	ref_ptr<Grid3f> g32 = make_noise_grid(n, sp, Vector3d(0.), gseed, nG);
	{
		std::string d = casedir("grid32");
		save_grid(d, g32);
		save_meta(d, grid_meta("\"note\": \"shared by the *_grid32 trajectory cases\"", n, sp, gseed, nG));
	}
	ref_ptr<MagneticField> fgrid = new MagneticFieldGrid(g32);
	Vector3d centre(n * sp / 2, n * sp / 2, n * sp / 2);
	std::vector<int> proton = {nucleusId(60, 1)};
	std::vector<int> species = {nucleusId(60, 1), nucleusId(60, 2), nucleusId(60, 7), nucleusId(60, 14), nucleusId(60, 26)};
	std::vector<double> one = {1.0};
	std::vector<double> wts = {0.5, 0.2, 0.15, 0.1, 0.05};

	// ---- ModulatedMagneticFieldGrid: grid32 times a reflective 9x11x10 modulation grid
	{
		Vector3d mo(-0.3 * Mpc, 0.1 * Mpc, 0.25 * Mpc);
		const double msp = 0.2 * Mpc;
		ref_ptr<Grid1f> mg = make_mod_grid(9, 11, 10, msp, mo, 777);
		ref_ptr<MagneticField> fmod = new ModulatedMagneticFieldGrid(g32, mg);
		std::vector<float> mv(9 * 11 * 10);
		for (int ix = 0; ix < 9; ix++)
			for (int iy = 0; iy < 11; iy++)
				for (int iz = 0; iz < 10; iz++)
					mv[(ix * 11 + iy) * 10 + iz] = mg->get(ix, iy, iz);
		const std::string modmeta = "\"field\": \"grid32_modulated\", \"mod_shape\": [9, 11, 10], \"mod_spacing\": " +
		                            fmt(msp) + ", \"mod_origin\": [" + fmt(mo.x) + ", " + fmt(mo.y) + ", " + fmt(mo.z) +
		                            "], \"mod_reflective\": true";
		{
			std::vector<Particle> ps = make_particles(128, 58, 1 * EeV, 100 * EeV, centre, proton, one);
			PropagationBP bp(fmod, 0.1 * Mpc);
			run_traj("bp_fixed_grid32_modulated", bp, nullptr, ps, 300, {1, 10, 100, 300},
			         "{\"propagator\": \"BP\", \"fixedStep\": " + fmt(0.1 * Mpc) + ", " + modmeta +
			             ", \"particle_seed\": 58}");
			save_f32(casedir("bp_fixed_grid32_modulated"), "mod", mv, {9, 11, 10});
		}
		{
			std::vector<Particle> ps = make_particles(128, 59, 1 * EeV, 300 * EeV, centre, species, wts);
			PropagationCK ck(fmod, 1e-4, 1 * kpc, 1 * Mpc);
			MaximumTrajectoryLength ml(12 * Mpc);
			run_traj("ck_grid32_modulated", ck, &ml, ps, 3000, {1, 10, 100, 1000, 3000},
			         "{\"propagator\": \"CK\", \"tolerance\": 1e-4, \"minStep\": " + fmt(kpc) +
			             ", \"maxStep\": " + fmt(Mpc) + ", \"maxTrajectoryLength\": " + fmt(12 * Mpc) + ", " + modmeta +
			             ", \"particle_seed\": 59}");
			save_f32(casedir("ck_grid32_modulated"), "mod", mv, {9, 11, 10});
		}
	}

	{
		std::vector<Particle> ps = make_particles(256, 42, 1 * EeV, 100 * EeV, centre, proton, one);
		PropagationBP bp(fgrid, 0.1 * Mpc);
		run_traj("bp_fixed_grid32", bp, nullptr, ps, 1000, {1, 2, 10, 100, 1000},
		         "{\"propagator\": \"BP\", \"fixedStep\": " + fmt(0.1 * Mpc) +
		             ", \"field\": \"grid32\", \"particle_seed\": 42, \"Emin\": " + fmt(EeV) +
		             ", \"Emax\": " + fmt(100 * EeV) + "}");
	}
	{
		std::vector<Particle> ps = make_particles(256, 43, 1 * EeV, 100 * EeV, centre, proton, one);
		PropagationBP bp(fgrid, 1e-3, 1 * kpc, 1 * Mpc);
		run_traj("bp_adaptive_grid32", bp, nullptr, ps, 1000, {1, 2, 10, 100, 1000},
		         "{\"propagator\": \"BP\", \"tolerance\": 1e-3, \"minStep\": " + fmt(kpc) +
		             ", \"maxStep\": " + fmt(Mpc) + ", \"field\": \"grid32\", \"particle_seed\": 43, \"Emin\": " +
		             fmt(EeV) + ", \"Emax\": " + fmt(100 * EeV) + "}");
	}
	{
		std::vector<Particle> ps = make_particles(256, 44, 1 * EeV, 300 * EeV, centre, species, wts);
		PropagationCK ck(fgrid, 1e-4, 1 * kpc, 1 * Mpc);
		MaximumTrajectoryLength ml(20 * Mpc);
		run_traj("ck_grid32_maxlen", ck, &ml, ps, 4000, {1, 2, 10, 100, 1000, 4000},
		         "{\"propagator\": \"CK\", \"tolerance\": 1e-4, \"minStep\": " + fmt(kpc) +
		             ", \"maxStep\": " + fmt(Mpc) + ", \"maxTrajectoryLength\": " + fmt(20 * Mpc) +
		             ", \"field\": \"grid32\", \"particle_seed\": 44, \"Emin\": " + fmt(EeV) + ", \"Emax\": " +
		             fmt(300 * EeV) + ", \"species_weights\": [0.5, 0.2, 0.15, 0.1, 0.05]}");
	}
	{
		std::vector<Particle> ps = make_particles(128, 45, 1 * EeV, 100 * EeV, centre, proton, one);
		PropagationCK ck(fgrid, 1e-4, 10 * kpc, 10 * kpc);
		run_traj("ck_fixed_grid32", ck, nullptr, ps, 500, {1, 10, 100, 500},
		         "{\"propagator\": \"CK\", \"tolerance\": 1e-4, \"minStep\": " + fmt(10 * kpc) +
		             ", \"maxStep\": " + fmt(10 * kpc) + ", \"field\": \"grid32\", \"particle_seed\": 45, \"Emin\": " +
		             fmt(EeV) + ", \"Emax\": " + fmt(100 * EeV) + "}");
	}

	// ---- the same grid read with reflective continuation and with clipVolume:
	// particles leave the 1.6 Mpc box within the run (Grid.h:49-59, 320-327)
	{
		ref_ptr<Grid3f> gr = new Grid3f(*g32);
		gr->setReflective(true);
		ref_ptr<MagneticField> fr = new MagneticFieldGrid(gr);
		std::vector<Particle> ps = make_particles(128, 46, 1 * EeV, 100 * EeV, centre, proton, one);
		PropagationBP bp(fr, 0.1 * Mpc);
		run_traj("bp_fixed_grid32_reflective", bp, nullptr, ps, 600, {1, 10, 100, 600},
		         "{\"propagator\": \"BP\", \"fixedStep\": " + fmt(0.1 * Mpc) +
		             ", \"field\": \"grid32\", \"reflective\": true, \"particle_seed\": 46, \"Emin\": " + fmt(EeV) +
		             ", \"Emax\": " + fmt(100 * EeV) + "}");
	}
	{
		ref_ptr<Grid3f> gc = new Grid3f(*g32);
		gc->setClipVolume(true);
		ref_ptr<MagneticField> fc = new MagneticFieldGrid(gc);
		std::vector<Particle> ps = make_particles(128, 47, 1 * EeV, 100 * EeV, centre, proton, one);
		PropagationCK ck(fc, 1e-4, 1 * kpc, 1 * Mpc);
		MaximumTrajectoryLength ml(10 * Mpc);
		run_traj("ck_grid32_clip", ck, &ml, ps, 3000, {1, 10, 100, 1000, 3000},
		         "{\"propagator\": \"CK\", \"tolerance\": 1e-4, \"minStep\": " + fmt(kpc) +
		             ", \"maxStep\": " + fmt(Mpc) + ", \"maxTrajectoryLength\": " + fmt(10 * Mpc) +
		             ", \"field\": \"grid32\", \"clip\": true, \"particle_seed\": 47, \"Emin\": " + fmt(EeV) +
		             ", \"Emax\": " + fmt(100 * EeV) + "}");
	}

	// ---- PeriodicBox around the 32^3 grid (origin 0, 1.6 Mpc): positions are wrapped back,
	// source and created positions move the other way
	{
		PeriodicBox box(Vector3d(0.), Vector3d(n * sp));
		std::vector<Particle> ps = make_particles(128, 50, 1 * EeV, 100 * EeV, centre, proton, one);
		PropagationBP bp(fgrid, 0.1 * Mpc);
		run_modules_case("bp_fixed_grid32_box", bp, {&box}, ps, 500, {1, 10, 100, 500},
		             "{\"propagator\": \"BP\", \"fixedStep\": " + fmt(0.1 * Mpc) +
		                 ", \"field\": \"grid32\", \"box\": {\"origin\": [0, 0, 0], \"size\": [" + fmt(n * sp) +
		                 ", " + fmt(n * sp) + ", " + fmt(n * sp) + "]}, \"particle_seed\": 50}");
	}
	{
		PeriodicBox box(Vector3d(-0.2 * Mpc, 0.1 * Mpc, 0.), Vector3d(1.1 * Mpc, 1.3 * Mpc, 0.9 * Mpc));
		std::vector<Particle> ps = make_particles(128, 51, 1 * EeV, 300 * EeV, centre, species, wts);
		PropagationCK ck(fgrid, 1e-4, 1 * kpc, 1 * Mpc);
		MaximumTrajectoryLength ml(15 * Mpc);
		run_modules_case("ck_grid32_box", ck, {&box, &ml}, ps, 4000, {1, 10, 100, 1000, 4000},
		             "{\"propagator\": \"CK\", \"tolerance\": 1e-4, \"minStep\": " + fmt(kpc) +
		                 ", \"maxStep\": " + fmt(Mpc) + ", \"maxTrajectoryLength\": " + fmt(15 * Mpc) +
		                 ", \"field\": \"grid32\", \"box\": {\"origin\": [" + fmt(-0.2 * Mpc) + ", " + fmt(0.1 * Mpc) +
		                 ", 0], \"size\": [" + fmt(1.1 * Mpc) + ", " + fmt(1.3 * Mpc) + ", " + fmt(0.9 * Mpc) +
		                 "]}, \"particle_seed\": 51}");
	}

	// ---- SphericalBoundary (reject outside, limitStep with margin): fixed-step BP alone, adaptive CK
	// with a length limit
	{
		std::vector<Particle> ps = make_particles(128, 54, 1 * EeV, 100 * EeV, centre, proton, one);
		PropagationBP bp(fgrid, 0.1 * Mpc);
		Vector3d sc = centre + Vector3d(0.1 * Mpc, -0.05 * Mpc, 0.);
		SphericalBoundary sph(sc, 1.2 * Mpc);
		run_modules_case("bp_fixed_grid32_sphere", bp, {&sph}, ps, 300, {1, 10, 100, 300},
		                 "{\"propagator\": \"BP\", \"fixedStep\": " + fmt(0.1 * Mpc) +
		                     ", \"field\": \"grid32\", \"sphere\": {\"center\": [" + fmt(sc.x) + ", " + fmt(sc.y) +
		                     ", " + fmt(sc.z) + "], \"radius\": " + fmt(1.2 * Mpc) + ", \"margin\": " + fmt(0.1 * kpc) +
		                     ", \"limitStep\": true}, \"particle_seed\": 54}");
	}
	{
		std::vector<Particle> ps = make_particles(128, 55, 1 * EeV, 300 * EeV, centre, species, wts);
		PropagationCK ck(fgrid, 1e-4, 1 * kpc, 1 * Mpc);
		Vector3d sc = centre + Vector3d(-0.1 * Mpc, 0., 0.2 * Mpc);
		SphericalBoundary sph(sc, 2 * Mpc);
		sph.setMargin(0.05 * Mpc);
		MaximumTrajectoryLength ml(15 * Mpc);
		run_modules_case("ck_grid32_sphere", ck, {&sph, &ml}, ps, 4000, {1, 10, 100, 1000, 4000},
		                 "{\"propagator\": \"CK\", \"tolerance\": 1e-4, \"minStep\": " + fmt(kpc) +
		                     ", \"maxStep\": " + fmt(Mpc) + ", \"maxTrajectoryLength\": " + fmt(15 * Mpc) +
		                     ", \"field\": \"grid32\", \"sphere\": {\"center\": [" + fmt(sc.x) + ", " + fmt(sc.y) +
		                     ", " + fmt(sc.z) + "], \"radius\": " + fmt(2 * Mpc) + ", \"margin\": " + fmt(0.05 * Mpc) +
		                     ", \"limitStep\": true}, \"particle_seed\": 55}");
	}

	// ---- ObserverTimeEvolution: log-spaced times with a length limit (CK), and an
	// explicit list with deactivation on detection (fixed-step BP)
	{
		std::vector<Particle> ps = make_particles(256, 48, 1 * EeV, 300 * EeV, centre, species, wts);
		PropagationCK ck(fgrid, 1e-4, 1 * kpc, 1 * Mpc);
		MaximumTrajectoryLength ml(12 * Mpc);
		run_observer_case("ck_grid32_observer", ck, {&ml}, {}, new ObserverTimeEvolution(0.5 * Mpc, 10 * Mpc, 12, true), false,
		                  ps, 100000,
		                  "{\"propagator\": \"CK\", \"tolerance\": 1e-4, \"minStep\": " + fmt(kpc) +
		                      ", \"maxStep\": " + fmt(Mpc) + ", \"maxTrajectoryLength\": " + fmt(12 * Mpc) +
		                      ", \"field\": \"grid32\", \"observer\": {\"min\": " + fmt(0.5 * Mpc) + ", \"max\": " +
		                      fmt(10 * Mpc) + ", \"n\": 12, \"log\": true, \"deactivate\": false}, \"particle_seed\": 48}");
	}
	{
		std::vector<Particle> ps = make_particles(128, 49, 1 * EeV, 100 * EeV, centre, proton, one);
		PropagationBP bp(fgrid, 0.1 * Mpc);
		std::vector<double> tl = {0.35 * Mpc, 1.0 * Mpc, 1.05 * Mpc, 2.5 * Mpc, 7.77 * Mpc};
		run_observer_case("bp_fixed_grid32_observer", bp, {}, {}, new ObserverTimeEvolution(tl), false, ps, 100,
		                  "{\"propagator\": \"BP\", \"fixedStep\": " + fmt(0.1 * Mpc) +
		                      ", \"field\": \"grid32\", \"observer\": {\"list\": true, \"deactivate\": false}, "
		                      "\"particle_seed\": 49}");
	}

	// ---- module order around the observer: observer before the boundary and the length limit
	// (the detection sees nextStep before their limits), and a boundary before the observer
	// (detections after the rejection see an inactive, flagged candidate)
	{
		std::vector<Particle> ps = make_particles(128, 56, 1 * EeV, 300 * EeV, centre, species, wts);
		PropagationCK ck(fgrid, 1e-4, 1 * kpc, 1 * Mpc);
		Vector3d sc = centre + Vector3d(0.05 * Mpc, 0., -0.1 * Mpc);
		SphericalBoundary sph(sc, 3 * Mpc);
		sph.setMargin(0.02 * Mpc);
		MaximumTrajectoryLength ml(6 * Mpc);
		std::vector<double> tl = {0.2 * Mpc, 0.9 * Mpc, 1.7 * Mpc, 2.6 * Mpc, 3.0 * Mpc, 4.4 * Mpc, 5.9 * Mpc, 6.0 * Mpc};
		run_observer_case("ck_grid32_observer_first", ck, {}, {&sph, &ml}, new ObserverTimeEvolution(tl), false, ps,
		                  100000,
		                  "{\"propagator\": \"CK\", \"tolerance\": 1e-4, \"minStep\": " + fmt(kpc) +
		                      ", \"maxStep\": " + fmt(Mpc) + ", \"maxTrajectoryLength\": " + fmt(6 * Mpc) +
		                      ", \"field\": \"grid32\", \"order\": [\"observer\", \"sphere\", \"maxlen\"], "
		                      "\"sphere\": {\"center\": [" + fmt(sc.x) + ", " + fmt(sc.y) + ", " + fmt(sc.z) +
		                      "], \"radius\": " + fmt(3 * Mpc) + ", \"margin\": " + fmt(0.02 * Mpc) +
		                      ", \"limitStep\": true}, \"observer\": {\"list\": true, \"deactivate\": false}, "
		                      "\"particle_seed\": 56}");
	}
	{
		std::vector<Particle> ps = make_particles(128, 57, 1 * EeV, 100 * EeV, centre, proton, one);
		PropagationBP bp(fgrid, 0.1 * Mpc);
		SphericalBoundary sph(centre, 0.75 * Mpc);
		std::vector<double> tl = {0.3 * Mpc, 0.7 * Mpc, 0.8 * Mpc, 0.9 * Mpc, 1.0 * Mpc};
		run_observer_case("bp_fixed_grid32_sphere_observer", bp, {&sph}, {}, new ObserverTimeEvolution(tl), false, ps,
		                  100,
		                  "{\"propagator\": \"BP\", \"fixedStep\": " + fmt(0.1 * Mpc) +
		                      ", \"field\": \"grid32\", \"order\": [\"sphere\", \"observer\"], "
		                      "\"sphere\": {\"center\": [" + fmt(centre.x) + ", " + fmt(centre.y) + ", " +
		                      fmt(centre.z) + "], \"radius\": " + fmt(0.75 * Mpc) + ", \"margin\": " + fmt(0.1 * kpc) +
		                      ", \"limitStep\": true}, \"observer\": {\"list\": true, \"deactivate\": false}, "
		                      "\"particle_seed\": 57}");
	}

	// ---- C1-shaped: 64^3 grid (50 kpc), 4096 protons, BP fixed 0.1 Mpc x 2000
	{
		const int n64 = 64;
		ref_ptr<Grid3f> g64 = make_noise_grid(n64, sp, Vector3d(0.), 4242, nG);
		ref_ptr<MagneticField> f64 = new MagneticFieldGrid(g64);
		Vector3d c64(n64 * sp / 2, n64 * sp / 2, n64 * sp / 2);
		std::vector<Particle> ps = make_particles(4096, 42, 1 * EeV, 100 * EeV, c64, proton, one);
		PropagationBP bp(f64, 0.1 * Mpc);
		run_traj("bp_fixed_grid64_c1", bp, nullptr, ps, 2000, {1, 1000, 2000},
		         "{\"propagator\": \"BP\", \"fixedStep\": " + fmt(0.1 * Mpc) +
		             ", \"grid\": {\"n\": 64, \"spacing\": " + fmt(sp) + ", \"origin\": [0, 0, 0], \"seed\": 4242, \"amp\": " +
		             fmt(nG) + ", \"generated\": \"noise grid from the same recipe; not stored\"}, \"particle_seed\": 42, \"Emin\": " +
		             fmt(EeV) + ", \"Emax\": " + fmt(100 * EeV) + "}");
	}

	// ---- plane-wave trajectories
	{
		SimpleSpectrum spec(1 * muG, 1 * pc, 100 * pc);
		ref_ptr<MagneticField> pw = new PlaneWaveTurbulence(spec, 32, 11);
		std::vector<Particle> ps = make_particles(128, 46, 1e15 * eV, 1e18 * eV, Vector3d(0.), proton, one);
		PropagationBP bp(pw, 0.05 * pc);
		run_traj("bp_fixed_planewave32", bp, nullptr, ps, 1000, {1, 10, 100, 1000},
		         "{\"propagator\": \"BP\", \"fixedStep\": " + fmt(0.05 * pc) +
		             ", \"field\": {\"type\": \"planewave\", \"spectrum\": \"simple\", \"Brms\": " + fmt(muG) +
		             ", \"lMin\": " + fmt(pc) + ", \"lMax\": " + fmt(100 * pc) + ", \"sIndex\": " + fmt(5. / 3) +
		             ", \"Nm\": 32, \"seed\": 11}, \"particle_seed\": 46, \"Emin\": " + fmt(1e15 * eV) +
		             ", \"Emax\": " + fmt(1e18 * eV) + "}");
		PropagationCK ck(pw, 1e-4, 0.001 * pc, 1 * pc);
		std::vector<Particle> ps2 = make_particles(128, 47, 1e15 * eV, 1e18 * eV, Vector3d(0.), proton, one);
		run_traj("ck_planewave32", ck, nullptr, ps2, 300, {1, 10, 100, 300},
		         "{\"propagator\": \"CK\", \"tolerance\": 1e-4, \"minStep\": " + fmt(0.001 * pc) +
		             ", \"maxStep\": " + fmt(pc) +
		             ", \"field\": {\"type\": \"planewave\", \"spectrum\": \"simple\", \"Brms\": " + fmt(muG) +
		             ", \"lMin\": " + fmt(pc) + ", \"lMax\": " + fmt(100 * pc) + ", \"sIndex\": " + fmt(5. / 3) +
		             ", \"Nm\": 32, \"seed\": 11}, \"particle_seed\": 47, \"Emin\": " + fmt(1e15 * eV) +
		             ", \"Emax\": " + fmt(1e18 * eV) + "}");
	}

	// ---- uniform field: the reference tests' gyration setup, many directions
	{
		ref_ptr<MagneticField> uf = new UniformMagneticField(Vector3d(0, 0, 1 * nG));
		std::vector<Particle> ps = make_particles(64, 48, 1 * EeV, 100 * EeV, Vector3d(0.), species, wts);
		PropagationBP bp(uf, 10 * Mpc);
		run_traj("bp_fixed_uniform", bp, nullptr, ps, 100, {1, 10, 100},
		         "{\"propagator\": \"BP\", \"fixedStep\": " + fmt(10 * Mpc) +
		             ", \"field\": {\"type\": \"uniform\", \"B\": [0, 0, " + fmt(nG) +
		             "]}, \"particle_seed\": 48, \"Emin\": " + fmt(EeV) + ", \"Emax\": " + fmt(100 * EeV) + "}");
		PropagationCK ck(uf, 1e-4, 0.1 * kpc, 1 * Gpc);
		run_traj("ck_uniform", ck, nullptr, ps, 100, {1, 10, 100},
		         "{\"propagator\": \"CK\", \"tolerance\": 1e-4, \"minStep\": " + fmt(0.1 * kpc) +
		             ", \"maxStep\": " + fmt(Gpc) + ", \"field\": {\"type\": \"uniform\", \"B\": [0, 0, " + fmt(nG) +
		             "]}, \"particle_seed\": 48, \"Emin\": " + fmt(EeV) + ", \"Emax\": " + fmt(100 * EeV) + "}");
	}
	printf("done\n");
	return 0;
}
