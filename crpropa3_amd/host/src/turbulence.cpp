// GridTurbulence / SimpleGridTurbulence (src/magneticField/turbulentField/GridTurbulence.cpp,
// SimpleGridTurbulence.cpp): MT19937 draws on the host in the reference's order, the k-space
// field and the inverse FFT on the GPU (crp_turbulence_generate), scaleGrid to Brms with the
// reference's rmsFieldStrength.
#include "crpropa/magneticField/turbulentField/GridTurbulence.h"
#include "crpropa/GridTools.h"
#include "crpropa/Random.h"

#include "crpropa3_amd.h"

#include <stdexcept>
#include <vector>

namespace crpropa {

static void throwIf(int rc) {
	if (rc != CRP_OK)
		throw std::runtime_error(crp_last_error());
}

// fills grid with the unnormalised field of the given spectrum kind; the draws come from Random
// seeded as the reference seeds it (seed 0: /dev/urandom)
static void generate(ref_ptr<Grid3f> grid, int kind, double lMin, double lMax, double alpha, double lbo, double s,
                     double q, unsigned int seed) {
	crp_turbulence_desc d;
	d.n = (int)grid->getNx();
	d.spacing = grid->getSpacing().x;
	d.lmin = lMin;
	d.lmax = lMax;
	d.kind = kind;
	d.alpha = alpha;
	d.lbendover = lbo;
	d.sindex = s;
	d.qindex = q;
	uint64_t modes = 0;
	throwIf(crp_turbulence_mode_count(&d, &modes));
	const uint64_t nd = modes * (kind == CRP_TURBULENCE_SIMPLE ? 4 : 2);
	std::vector<uint32_t> draws(nd);
	Random random;
	if (seed != 0)
		random.seed(seed);
	for (uint64_t i = 0; i < nd; i++)
		draws[i] = random.randInt();
	throwIf(crp_turbulence_generate(&d, draws.data(), nd, reinterpret_cast<float *>(grid->getGrid().data())));
}

void GridTurbulence::checkGridRequirements(ref_ptr<Grid3f> grid, double lMin, double lMax) {
	const size_t Nx = grid->getNx(), Ny = grid->getNy(), Nz = grid->getNz();
	const Vector3d spacing = grid->getSpacing();
	if ((Nx != Ny) or (Ny != Nz))
		throw std::runtime_error("turbulentField: only cubic grid supported");
	if ((spacing.x != spacing.y) or (spacing.y != spacing.z))
		throw std::runtime_error("turbulentField: only equal spacing suported");
	if (lMin < 2 * spacing.x)
		throw std::runtime_error("turbulentField: lMin < 2 * spacing");
	if (lMax > Nx * spacing.x)
		throw std::runtime_error("turbulentField: lMax > size");
	if (lMax < lMin)
		throw std::runtime_error("lMax < lMin");
}

GridTurbulence::GridTurbulence(const TurbulenceSpectrum &spectrum, const GridProperties &gridProp,
                               unsigned int seed)
    : GridTurbulence(spectrum, gridProp, seed, true) {}

GridTurbulence::GridTurbulence(const TurbulenceSpectrum &spectrum, const GridProperties &gridProp,
                               unsigned int seed, bool fill)
    : TurbulentField(spectrum), seed(seed) {
	gridPtr = new Grid3f(gridProp);
	checkGridRequirements(gridPtr, spectrum.getLmin(), spectrum.getLmax());
	if (!fill)
		return;
	// GridTurbulence::initTurbulence (GridTurbulence.cpp:29-124): the spectrum's own E(k)
	const bool powerlaw = dynamic_cast<const SimpleTurbulenceSpectrum *>(&spectrum) != nullptr;
	generate(gridPtr, powerlaw ? CRP_TURBULENCE_GENERAL_POWERLAW : CRP_TURBULENCE_GENERAL, spectrum.getLmin(),
	         spectrum.getLmax(), 0, spectrum.getLbendover(), spectrum.getSindex(), spectrum.getQindex(), seed);
	scaleGrid(gridPtr, spectrum.getBrms() / rmsFieldStrength(gridPtr));
}

Vector3d GridTurbulence::getField(const Vector3d &pos) const { return gridPtr->interpolate(pos); }

Vector3f GridTurbulence::getMeanFieldVector() const { return meanFieldVector(gridPtr); }
double GridTurbulence::getMeanFieldStrength() const { return meanFieldStrength(gridPtr); }
double GridTurbulence::getRmsFieldStrength() const { return rmsFieldStrength(gridPtr); }
std::array<float, 3> GridTurbulence::getRmsFieldStrengthPerAxis() const { return rmsFieldStrengthPerAxis(gridPtr); }
void GridTurbulence::dumpToFile(std::string filename) const { dumpGrid(gridPtr, filename); }

// The reference's constructor first runs the general algorithm with the power-law spectrum and then
// overwrites the grid with initTurbulence (SimpleGridTurbulence.cpp:10-16); only the second pass
// determines the grid, so it is the only one run here.
SimpleGridTurbulence::SimpleGridTurbulence(const SimpleTurbulenceSpectrum &spectrum, const GridProperties &gridProp,
                                           unsigned int seed)
    : GridTurbulence(spectrum, gridProp, seed, false) {
	initTurbulence(gridPtr, spectrum.getBrms(), spectrum.getLmin(), spectrum.getLmax(), -spectrum.getSindex() - 2,
	               seed);
}

void SimpleGridTurbulence::initTurbulence(ref_ptr<Grid3f> grid, double Brms, double lMin, double lMax, double alpha,
                                          int seed) {
	checkGridRequirements(grid, lMin, lMax);
	generate(grid, CRP_TURBULENCE_SIMPLE, lMin, lMax, alpha, 0, 0, 0, (unsigned int)seed);
	scaleGrid(grid, Brms / rmsFieldStrength(grid)); // normalize to Brms
}

double turbulentCorrelationLength(double lMin, double lMax, double alpha) {
	return SimpleTurbulenceSpectrum::turbulentCorrelationLength(lMin, lMax, -alpha - 2);
}

void initTurbulence(ref_ptr<Grid3f> grid, double Brms, double lMin, double lMax, double alpha, int seed) {
	SimpleGridTurbulence::initTurbulence(grid, Brms, lMin, lMax, alpha, seed);
}

} // namespace crpropa
