// Host side of the fields: Grid3f storage, the reference's field classes, and
// their device descriptors (include/crpropa3_amd.h).
#include "crpropa/Grid.h"
#include "crpropa/Random.h"
#include "crpropa/magneticField/MagneticField.h"
#include "crpropa/magneticField/MagneticFieldGrid.h"
#include "crpropa/magneticField/JF12Field.h"
#include "crpropa/magneticField/turbulentField/GridTurbulence.h"
#include "crpropa/Units.h"
#include "crpropa/magneticField/turbulentField/PlaneWaveTurbulence.h"

#include "crpropa3_amd.h"

#include <cmath>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

namespace crpropa {

static void throwIf(int rc) {
	if (rc != CRP_OK)
		throw std::runtime_error(crp_last_error());
}

void evaluateOnDevice(const MagneticField &f, const double *pos, size_t n, double *B) {
	crp_field *d = f.deviceField();
	if (!d)
		throw std::runtime_error("crpropa3_amd: this magnetic field has no device implementation");
	throwIf(crp_field_eval(d, pos, n, B));
}

// ------------------------------------------------------------------ Grid3f
Grid3f::Grid3f(Vector3d origin, size_t N, double spacing)
    : Nx(0), Ny(0), Nz(0), origin(origin), spacing(Vector3d(spacing)), clipVolume(false), reflective(false),
      ipolType(TRILINEAR), dev(nullptr), dirty(true) {
	setGridSize(N, N, N);
}
Grid3f::Grid3f(Vector3d origin, size_t nx, size_t ny, size_t nz, double spacing)
    : Nx(0), Ny(0), Nz(0), origin(origin), spacing(Vector3d(spacing)), clipVolume(false), reflective(false),
      ipolType(TRILINEAR), dev(nullptr), dirty(true) {
	setGridSize(nx, ny, nz);
}
Grid3f::Grid3f(Vector3d origin, size_t nx, size_t ny, size_t nz, Vector3d spacing)
    : Nx(0), Ny(0), Nz(0), origin(origin), spacing(spacing), clipVolume(false), reflective(false),
      ipolType(TRILINEAR), dev(nullptr), dirty(true) {
	setGridSize(nx, ny, nz);
}
Grid3f::Grid3f(const GridProperties &p)
    : Nx(0), Ny(0), Nz(0), origin(p.origin), spacing(p.spacing), clipVolume(p.clipVolume),
      reflective(p.reflective), ipolType(TRILINEAR), dev(nullptr), dirty(true) {
	setGridSize(p.Nx, p.Ny, p.Nz);
	setInterpolationType(p.ipol);
}

Grid3f::~Grid3f() {
	if (dev)
		crp_field_destroy(dev);
}

void Grid3f::setGridSize(size_t nx, size_t ny, size_t nz) {
	Nx = nx;
	Ny = ny;
	Nz = nz;
	grid.resize(nx * ny * nz);
	changed();
}

void Grid3f::setInterpolationType(interpolationType t) {
	if (t != TRILINEAR && t != TRICUBIC && t != NEAREST_NEIGHBOUR)
		throw std::runtime_error("InterpolationType: unknown interpolation type");
	ipolType = t;
	changed();
}

// the device field's interpolation types as the host grids hold them; destroys f on failure
static void setDeviceInterpolation(crp_field *f, int which, interpolationType t) {
	if (t == TRILINEAR)
		return;
	const int rc = crp_field_set_grid_interpolation(f, which, (int)t);
	if (rc != CRP_OK) {
		const std::string msg = crp_last_error();
		crp_field_destroy(f);
		throw std::runtime_error(msg);
	}
}

Vector3d Grid3f::positionFromIndex(int index) const {
	int ix = index / (Ny * Nz);
	int iy = (index / Nz) % Ny;
	int iz = index % Nz;
	return Vector3d(ix, iy, iz) * spacing + (origin + spacing / 2.);
}

crp_field *Grid3f::deviceField() const {
	std::lock_guard<std::mutex> lock(mtx);
	if (dev && !dirty)
		return dev;
	if (dev) {
		crp_field_destroy(dev);
		dev = nullptr;
	}
	const double o[3] = {origin.x, origin.y, origin.z};
	const double s[3] = {spacing.x, spacing.y, spacing.z};
	crp_field *f = nullptr;
	throwIf(crp_field_create_grid3f(reinterpret_cast<const float *>(grid.data()), (int)Nx, (int)Ny, (int)Nz, o, s,
	                                reflective ? 1 : 0, clipVolume ? 1 : 0, &f));
	setDeviceInterpolation(f, CRP_GRID_VALUES, ipolType);
	dev = f;
	dirty = false;
	return dev;
}

Vector3f Grid3f::interpolate(const Vector3d &position) const {
	const double p[3] = {position.x, position.y, position.z};
	double B[3];
	throwIf(crp_field_eval(deviceField(), p, 1, B));
	return Vector3f((float)B[0], (float)B[1], (float)B[2]);
}

static_assert(sizeof(Vector3f) == 3 * sizeof(float), "Vector3f must be three packed floats");

// ------------------------------------------------------------------ simple fields
UniformMagneticField::~UniformMagneticField() {
	if (dev)
		crp_field_destroy(dev);
}

crp_field *UniformMagneticField::deviceField() const {
	std::lock_guard<std::mutex> lock(mtx);
	if (!dev) {
		const double b[3] = {value.x, value.y, value.z};
		throwIf(crp_field_create_uniform(b, &dev));
	}
	return dev;
}

Vector3d MagneticFieldGrid::getField(const Vector3d &position) const {
	if (!grid.valid())
		return Vector3d(0.);
	const double p[3] = {position.x, position.y, position.z};
	double B[3];
	throwIf(crp_field_eval(grid->deviceField(), p, 1, B));
	return Vector3d(B[0], B[1], B[2]);
}

// ------------------------------------------------------------------ decorators
DecoratorField::~DecoratorField() {
	if (dev)
		crp_field_destroy(dev);
}

void DecoratorField::invalidate() {
	std::lock_guard<std::mutex> lock(mtx);
	if (dev)
		crp_field_destroy(dev);
	dev = nullptr;
	devChildren.clear();
}

crp_field *DecoratorField::deviceField() const {
	std::vector<crp_field *> ch;
	for (const ref_ptr<MagneticField> &f : childFields()) {
		crp_field *d = f.valid() ? f->deviceField() : nullptr;
		if (f.valid() && !d)
			return nullptr; // a child the device cannot evaluate
		ch.push_back(d);
	}
	std::lock_guard<std::mutex> lock(mtx);
	if (dev && ch == devChildren)
		return dev;
	if (dev)
		crp_field_destroy(dev);
	dev = nullptr;
	throwIf(create(ch, &dev));
	devChildren = ch;
	return dev;
}

PeriodicMagneticField::PeriodicMagneticField(ref_ptr<MagneticField> field, const Vector3d &extends)
    : field(field), origin(0, 0, 0), extends(extends), reflective(false) {}

PeriodicMagneticField::PeriodicMagneticField(ref_ptr<MagneticField> field, const Vector3d &extends,
                                             const Vector3d &origin, bool reflective)
    : field(field), origin(origin), extends(extends), reflective(reflective) {}

Vector3d &PeriodicMagneticField::getOrigin() {
	invalidate(); // the reference returns a reference that may be written through
	return origin;
}
void PeriodicMagneticField::setOrigin(const Vector3d &o) {
	origin = o;
	invalidate();
}
Vector3d &PeriodicMagneticField::getExtends() {
	invalidate();
	return extends;
}
void PeriodicMagneticField::setExtends(const Vector3d &e) {
	extends = e;
	invalidate();
}
bool PeriodicMagneticField::isReflective() { return reflective; }
void PeriodicMagneticField::setReflective(bool r) {
	reflective = r;
	invalidate();
}

int PeriodicMagneticField::create(const std::vector<crp_field *> &ch, crp_field **out) const {
	const double e[3] = {extends.x, extends.y, extends.z}, o[3] = {origin.x, origin.y, origin.z};
	return crp_field_create_periodic(ch[0], e, o, reflective ? 1 : 0, out);
}

Vector3d PeriodicMagneticField::getField(const Vector3d &position) const {
	double B[3];
	const double p[3] = {position.x, position.y, position.z};
	evaluateOnDevice(*this, p, 1, B);
	return Vector3d(B[0], B[1], B[2]);
}

void MagneticFieldList::addField(ref_ptr<MagneticField> field) {
	fields.push_back(field);
	invalidate();
}

int MagneticFieldList::create(const std::vector<crp_field *> &ch, crp_field **out) const {
	int rc = crp_field_create_list(out);
	for (size_t i = 0; rc == CRP_OK && i < ch.size(); i++)
		rc = crp_field_list_add(*out, ch[i]);
	if (rc != CRP_OK && *out) {
		crp_field_destroy(*out);
		*out = nullptr;
	}
	return rc;
}

Vector3d MagneticFieldList::getField(const Vector3d &position) const {
	double B[3];
	const double p[3] = {position.x, position.y, position.z};
	evaluateOnDevice(*this, p, 1, B);
	return Vector3d(B[0], B[1], B[2]);
}

MagneticFieldEvolution::MagneticFieldEvolution(ref_ptr<MagneticField> field, double m) : field(field), m(m) {}

int MagneticFieldEvolution::create(const std::vector<crp_field *> &ch, crp_field **out) const {
	return crp_field_create_evolution(ch[0], m, out);
}

Vector3d MagneticFieldEvolution::getField(const Vector3d &position, double z) const {
	crp_field *d = deviceField();
	if (!d)
		throw std::runtime_error("crpropa3_amd: this magnetic field has no device implementation");
	double B[3];
	const double p[3] = {position.x, position.y, position.z};
	throwIf(crp_field_eval_z(d, p, &z, 1, B));
	return Vector3d(B[0], B[1], B[2]);
}

// ------------------------------------------------------------------ plane waves
// Mode table exactly as PlaneWaveTurbulence::PlaneWaveTurbulence builds it
// (src/magneticField/turbulentField/PlaneWaveTurbulence.cpp:99-178), with the
// host's libm, so the table equals the reference's bit for bit.
PlaneWaveTurbulence::PlaneWaveTurbulence(const TurbulenceSpectrum &spectrum, int Nm, int seed)
    : TurbulentField(spectrum), Nm(Nm), dev(nullptr) {
	if (Nm <= 1)
		throw std::runtime_error("PlaneWaveTurbulence: Nm <= 1. Specify at least two wavemodes in "
		                         "order to generate the k distribution properly.");
	Random random;
	if (seed != 0)
		random.seed((uint32_t)seed);
	double kmax = 2 * M_PI / spectrum.getLmin();
	double kmin = 2 * M_PI / spectrum.getLmax();
	xi.assign(Nm, Vector3d(0.));
	kappa.assign(Nm, Vector3d(0.));
	phi.assign(Nm, 0.);
	costheta.assign(Nm, 0.);
	beta.assign(Nm, 0.);
	Ak.assign(Nm, 0.);
	k.assign(Nm, 0.);
	double delta = std::log10(kmax / kmin);
	for (int i = 0; i < Nm; i++)
		k[i] = std::pow(10, std::log10(kmin) + ((double)i) / ((double)(Nm - 1)) * delta);
	double delta_k0 = (k[1] - k[0]) / k[1];
	double Ak2_sum = 0;
	for (int i = 0; i < Nm; i++) {
		double kk = k[i];
		double kHat = kk * spectrum.getLbendover();
		double Gk = spectrum.energySpectrum(kk) * (1 + kHat * kHat);
		Ak[i] = Gk * delta_k0 * kk;
		Ak2_sum += Ak[i];
		double ph = random.randUniform(-M_PI, M_PI);
		double ct = random.randUniform(-1., 1.);
		double st = std::sqrt(1 - ct * ct);
		double alpha = random.randUniform(0, 2 * M_PI);
		double b = random.randUniform(0, 2 * M_PI);
		kappa[i] = Vector3d(st * std::cos(ph), st * std::sin(ph), ct);
		xi[i] = Vector3d(ct * std::cos(ph) * std::cos(alpha) + std::sin(ph) * std::sin(alpha),
		                 ct * std::sin(ph) * std::cos(alpha) - std::cos(ph) * std::sin(alpha), -st * std::cos(alpha));
		phi[i] = ph;
		costheta[i] = ct;
		beta[i] = b;
	}
	for (int i = 0; i < Nm; i++)
		Ak[i] = std::sqrt(2 * Ak[i] / Ak2_sum) * spectrum.getBrms();
}

PlaneWaveTurbulence::~PlaneWaveTurbulence() {
	if (dev)
		crp_field_destroy(dev);
}

crp_field *PlaneWaveTurbulence::deviceField() const {
	std::lock_guard<std::mutex> lock(mtx);
	if (!dev) {
		std::vector<double> x(3 * Nm), kp(3 * Nm);
		for (int i = 0; i < Nm; i++) {
			x[3 * i] = xi[i].x; x[3 * i + 1] = xi[i].y; x[3 * i + 2] = xi[i].z;
			kp[3 * i] = kappa[i].x; kp[3 * i + 1] = kappa[i].y; kp[3 * i + 2] = kappa[i].z;
		}
		throwIf(crp_field_create_planewave(Nm, k.data(), Ak.data(), x.data(), kp.data(), beta.data(), &dev));
	}
	return dev;
}

Vector3d PlaneWaveTurbulence::getField(const Vector3d &pos) const {
	const double p[3] = {pos.x, pos.y, pos.z};
	double B[3];
	throwIf(crp_field_eval(deviceField(), p, 1, B));
	return Vector3d(B[0], B[1], B[2]);
}

} // namespace crpropa

namespace crpropa {

// ------------------------------------------------------------------ ModulatedMagneticFieldGrid
ModulatedMagneticFieldGrid::ModulatedMagneticFieldGrid(ref_ptr<Grid3f> g, ref_ptr<Grid1f> m) {
	g->setReflective(false);
	m->setReflective(true);
	setGrid(g);
	setModulationGrid(m);
}

ModulatedMagneticFieldGrid::~ModulatedMagneticFieldGrid() {
	if (dev)
		crp_field_destroy(dev);
}

void ModulatedMagneticFieldGrid::setReflective(bool gridReflective, bool modGridReflective) {
	grid->setReflective(gridReflective);
	modGrid->setReflective(modGridReflective);
}

crp_field *ModulatedMagneticFieldGrid::deviceField() const {
	if (!grid.valid() || !modGrid.valid())
		return nullptr;
	std::lock_guard<std::mutex> lock(mtx);
	if (dev && devGrid == grid.get() && devMod == modGrid.get() && devGridVersion == grid->getVersion() &&
	    devModVersion == modGrid->getVersion())
		return dev;
	if (dev) {
		crp_field_destroy(dev);
		dev = nullptr;
	}
	const Grid3f &g = *grid;
	const Grid1f &m = *modGrid;
	const Vector3d go = g.getOrigin(), gs = g.getSpacing(), mo = m.getOrigin(), ms = m.getSpacing();
	const double o3[3] = {go.x, go.y, go.z}, s3[3] = {gs.x, gs.y, gs.z};
	const double o1[3] = {mo.x, mo.y, mo.z}, s1[3] = {ms.x, ms.y, ms.z};
	crp_field *f = nullptr;
	throwIf(crp_field_create_modulated_grid3f(
	    reinterpret_cast<const float *>(g.getGrid().data()), (int)g.getNx(), (int)g.getNy(), (int)g.getNz(), o3, s3,
	    g.isReflective() ? 1 : 0, g.getClipVolume() ? 1 : 0, m.getGrid().data(), (int)m.getNx(), (int)m.getNy(),
	    (int)m.getNz(), o1, s1, m.isReflective() ? 1 : 0, m.getClipVolume() ? 1 : 0, &f));
	setDeviceInterpolation(f, CRP_GRID_VALUES, g.getInterpolationType());
	setDeviceInterpolation(f, CRP_GRID_MODULATION, m.getInterpolationType());
	dev = f;
	devGrid = grid.get();
	devMod = modGrid.get();
	devGridVersion = grid->getVersion();
	devModVersion = modGrid->getVersion();
	return dev;
}

Vector3d ModulatedMagneticFieldGrid::getField(const Vector3d &position) const {
	const double p[3] = {position.x, position.y, position.z};
	double B[3];
	crp_field *d = deviceField();
	if (!d)
		throw std::runtime_error("ModulatedMagneticFieldGrid: grid or modulation grid not set");
	throwIf(crp_field_eval(d, p, 1, B));
	return Vector3d(B[0], B[1], B[2]);
}

} // namespace crpropa

// ------------------------------------------------------------------ JF12Field
namespace crpropa {

JF12Field::JF12Field() : JF12Field(0) {}

JF12Field::JF12Field(int planck) {
	if (crp_jf12_default_params(planck, &p) != CRP_OK)
		throw std::runtime_error(crp_last_error());
}

JF12Field::~JF12Field() { releaseDevice(); }

void JF12Field::releaseDevice() const {
	for (crp_field *&v : partDev) {
		if (v)
			crp_field_destroy(v);
		v = nullptr;
	}
	if (dev)
		crp_field_destroy(dev);
	dev = nullptr;
}

void JF12Field::randomStriated(int seed) {
	// JF12Field.cpp:81-94: 100^3 points at 0.1 kpc, round(rand()) * 2 - 1 in storage order
	const int N = 100;
	ref_ptr<Grid1f> g = new Grid1f(Vector3d(0.), N, 0.1 * kpc);
	Random random;
	if (seed != 0)
		random.seed(seed);
	for (int ix = 0; ix < N; ix++)
		for (int iy = 0; iy < N; iy++)
			for (int iz = 0; iz < N; iz++)
				g->get(ix, iy, iz) = (float)(std::round(random.rand()) * 2 - 1);
	setStriatedGrid(g);
}

void JF12Field::randomTurbulent(int seed) {
	// JF12Field.cpp:97-108: SimpleGridTurbulence(Brms 1, lMin 8 pc, lMax 272 pc) on 256^3 points at 4 pc
	ref_ptr<Grid3f> g = new Grid3f(Vector3d(0.), 256, 4 * pc);
	initTurbulence(g, 1, 8 * pc, 272 * pc, -11. / 3., seed);
	setTurbulentGrid(g);
}

void JF12Field::setStriatedGrid(ref_ptr<Grid1f> grid) {
	p.use_striated = 1;
	striatedGrid = grid;
}

void JF12Field::setTurbulentGrid(ref_ptr<Grid3f> grid) {
	p.use_turbulent = 1;
	turbulentGrid = grid;
}

void JF12Field::setUseRegularField(bool use) { p.use_regular = use; }
void JF12Field::setUseDiskField(bool use) { p.use_disk = use; }
void JF12Field::setUseToroidalHaloField(bool use) { p.use_toroidal_halo = use; }
void JF12Field::setUseXField(bool use) { p.use_x = use; }

void JF12Field::setUseStriatedField(bool use) {
	if (use && !striatedGrid.valid()) {
		std::cerr << "JF12Field: No striated field set: ignored. Run e.g. randomStriated()." << std::endl;
		return;
	}
	p.use_striated = use;
}

void JF12Field::setUseTurbulentField(bool use) {
	if (use && !turbulentGrid.valid()) {
		std::cerr << "JF12Field: No turbulent field set: ignored. Run e.g. randomTurbulent()." << std::endl;
		return;
	}
	p.use_turbulent = use;
}

crp_field *JF12Field::deviceField() const {
	std::lock_guard<std::mutex> lock(mtx);
	return deviceFieldLocked();
}

crp_field *JF12Field::deviceFieldLocked() const {
	// the grids that are set go to the device whether or not their switches are on, so that the part
	// getters can read them through views
	const Grid1f *s = striatedGrid.valid() ? striatedGrid.get() : nullptr;
	const Grid3f *t = turbulentGrid.valid() ? turbulentGrid.get() : nullptr;
	if (dev && std::memcmp(&devParams, &p, sizeof(p)) == 0 && devStr == s && devTurb == t &&
	    (!s || devStrVersion == s->getVersion()) && (!t || devTurbVersion == t->getVersion()))
		return dev;
	releaseDevice();
	int sn[3] = {0, 0, 0}, tn[3] = {0, 0, 0};
	double so[3] = {0, 0, 0}, ss[3] = {1, 1, 1}, to[3] = {0, 0, 0}, ts[3] = {1, 1, 1};
	if (s) {
		const Vector3d o = s->getOrigin(), sp = s->getSpacing();
		sn[0] = (int)s->getNx(); sn[1] = (int)s->getNy(); sn[2] = (int)s->getNz();
		so[0] = o.x; so[1] = o.y; so[2] = o.z;
		ss[0] = sp.x; ss[1] = sp.y; ss[2] = sp.z;
	}
	if (t) {
		const Vector3d o = t->getOrigin(), sp = t->getSpacing();
		tn[0] = (int)t->getNx(); tn[1] = (int)t->getNy(); tn[2] = (int)t->getNz();
		to[0] = o.x; to[1] = o.y; to[2] = o.z;
		ts[0] = sp.x; ts[1] = sp.y; ts[2] = sp.z;
	}
	crp_field *f = nullptr;
	throwIf(crp_field_create_jf12(&p, s ? s->getGrid().data() : nullptr, s ? sn : nullptr, so, ss,
	                              s && s->isReflective() ? 1 : 0,
	                              t ? reinterpret_cast<const float *>(t->getGrid().data()) : nullptr, t ? tn : nullptr,
	                              to, ts, t && t->isReflective() ? 1 : 0, t && t->getClipVolume() ? 1 : 0, &f));
	if (t)
		setDeviceInterpolation(f, CRP_GRID_VALUES, t->getInterpolationType());
	dev = f;
	devParams = p;
	devStr = s;
	devTurb = t;
	devStrVersion = s ? s->getVersion() : 0;
	devTurbVersion = t ? t->getVersion() : 0;
	return dev;
}

Vector3d JF12Field::evalWith(const Vector3d &pos, int regular, int striated, int turbulent) const {
	const int k = turbulent ? 2 : (striated ? 1 : 0);
	// the device field and its part views are shared by all calls on this field: one call at a time
	std::lock_guard<std::mutex> lock(mtx);
	crp_field *base = deviceFieldLocked();
	crp_jf12_params q = p;
	q.use_regular = regular;
	q.use_striated = striated && striatedGrid.valid();
	q.use_turbulent = turbulent && turbulentGrid.valid();
	if (!partDev[k] || std::memcmp(&partParams[k], &q, sizeof(q)) != 0) {
		if (partDev[k])
			crp_field_destroy(partDev[k]);
		partDev[k] = nullptr;
		throwIf(crp_field_create_jf12_view(base, &q, &partDev[k]));
		partParams[k] = q;
	}
	const double x[3] = {pos.x, pos.y, pos.z};
	double B[3];
	throwIf(crp_field_eval(partDev[k], x, 1, B));
	return Vector3d(B[0], B[1], B[2]);
}

Vector3d JF12Field::getRegularField(const Vector3d &pos) const { return evalWith(pos, 1, 0, 0); }
Vector3d JF12Field::getStriatedField(const Vector3d &pos) const { return evalWith(pos, 1, 1, 0); }
Vector3d JF12Field::getTurbulentField(const Vector3d &pos) const { return evalWith(pos, 0, 0, 1); }

Vector3d JF12Field::getField(const Vector3d &pos) const {
	const double x[3] = {pos.x, pos.y, pos.z};
	double B[3];
	evaluateOnDevice(*this, x, 1, B);
	return Vector3d(B[0], B[1], B[2]);
}

} // namespace crpropa
