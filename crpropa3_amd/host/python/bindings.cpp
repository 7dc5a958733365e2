// Python steering for the propagation path, with the reference's class names
// and call patterns (the SWIG module of python/2_headers.i; SWIG is replaced by
// pybind11).  Python subclasses of Module and MagneticField work as in the
// reference's directors: Python modules run on the host between GPU steps;
// Python fields cannot run on the GPU and are rejected by the propagators.
#include <pybind11/numpy.h>
#include <pybind11/operators.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "crpropa/Candidate.h"
#include "crpropa/Grid.h"
#include "crpropa/GridTools.h"
#include "crpropa/Module.h"
#include "crpropa/ModuleList.h"
#include "crpropa/ParticleState.h"
#include "crpropa/Random.h"
#include "crpropa/Units.h"
#include "crpropa/Vector3.h"
#include "crpropa/magneticField/MagneticField.h"
#include "crpropa/magneticField/MagneticFieldGrid.h"
#include "crpropa/magneticField/turbulentField/GridTurbulence.h"
#include "crpropa/magneticField/turbulentField/PlaneWaveTurbulence.h"
#include "crpropa/module/Boundary.h"
#include "crpropa/module/BreakCondition.h"
#include "crpropa/module/Observer.h"
#include "crpropa/module/Output.h"
#include "crpropa/module/PropagationBP.h"
#include "crpropa/module/PropagationCK.h"
#include "crpropa/module/DiffusionSDE.h"
#include "crpropa/magneticField/JF12Field.h"
#include "crpropa/advectionField/AdvectionField.h"

namespace py = pybind11;
using namespace crpropa;

PYBIND11_DECLARE_HOLDER_TYPE(T, crpropa::ref_ptr<T>, true);

namespace {

class PyModule : public Module {
public:
	void process(Candidate *c) const override { PYBIND11_OVERRIDE_PURE(void, Module, process, c); }
	std::string getDescription() const override { PYBIND11_OVERRIDE(std::string, Module, getDescription); }
};

class PyMagneticField : public MagneticField {
public:
	Vector3d getField(const Vector3d &p) const override { PYBIND11_OVERRIDE(Vector3d, MagneticField, getField, p); }
	Vector3d getField(const Vector3d &p, double z) const override {
		PYBIND11_OVERRIDE(Vector3d, MagneticField, getField, p, z);
	}
};

template <typename T>
void bind_vector3(py::module_ &m, const char *name) {
	typedef Vector3<T> V;
	py::class_<V>(m, name)
	    .def(py::init<>())
	    .def(py::init<T>())
	    .def(py::init<T, T, T>())
	    .def(py::init([](py::sequence s) {
		    if (py::len(s) != 3)
			    throw std::runtime_error("Expected a sequence of length 3");
		    return V(s[0].cast<T>(), s[1].cast<T>(), s[2].cast<T>());
	    }))
	    .def_readwrite("x", &V::x)
	    .def_readwrite("y", &V::y)
	    .def_readwrite("z", &V::z)
	    .def("getX", &V::getX)
	    .def("getY", &V::getY)
	    .def("getZ", &V::getZ)
	    .def("getR", &V::getR)
	    .def("getR2", &V::getR2)
	    .def("getUnitVector", &V::getUnitVector)
	    .def("getPhi", &V::getPhi)
	    .def("getTheta", &V::getTheta)
	    .def("getUnitVectorPhi", &V::getUnitVectorPhi)
	    .def("getUnitVectorTheta", &V::getUnitVectorTheta)
	    .def("getDistanceTo", &V::getDistanceTo)
	    .def("getAngleTo", &V::getAngleTo)
	    .def("dot", &V::dot)
	    .def("cross", &V::cross)
	    .def(py::self + py::self)
	    .def(py::self - py::self)
	    .def(py::self * py::self)
	    .def(py::self / py::self)
	    .def(py::self * T())
	    .def(T() * py::self)
	    .def(py::self / T())
	    .def(py::self == py::self)
	    .def(-py::self)
	    .def("__len__", [](const V &) { return 3; })
	    .def("__getitem__",
	         [](const V &v, int i) {
		         if (i < 0 || i > 2)
			         throw py::index_error();
		         return i == 0 ? v.x : (i == 1 ? v.y : v.z);
	         })
	    .def("__array__", [](const V &v, py::args, py::kwargs) {
		    py::array_t<T> a(3);
		    auto r = a.template mutable_unchecked<1>();
		    r(0) = v.x;
		    r(1) = v.y;
		    r(2) = v.z;
		    return a;
	    })
	    .def("__repr__", [name](const V &v) {
		    char buf[160];
		    snprintf(buf, sizeof buf, "Vector(%.6G, %.6G, %.6G)", (double)v.x, (double)v.y, (double)v.z);
		    return std::string(buf);
	    });
}

} // namespace

PYBIND11_MODULE(_crpropa, m) {
	m.doc() = "crpropa3_amd: GPU (MI355X) drop-in for CRPropa3's PropagationBP/CK hot path";

	// ---- units (include/crpropa/Units.h)
#define UNIT(n) m.attr(#n) = n
	UNIT(meter); UNIT(second); UNIT(kilogram); UNIT(ampere); UNIT(newton); UNIT(joule); UNIT(tesla); UNIT(volt);
	UNIT(eplus); UNIT(c_light); UNIT(c_squared); UNIT(amu); UNIT(mass_proton); UNIT(mass_electron);
	UNIT(gauss); UNIT(microgauss); UNIT(nanogauss); UNIT(muG); UNIT(nG);
	UNIT(electronvolt); UNIT(eV); UNIT(keV); UNIT(MeV); UNIT(GeV); UNIT(TeV); UNIT(PeV); UNIT(EeV);
	UNIT(au); UNIT(parsec); UNIT(kiloparsec); UNIT(megaparsec); UNIT(gigaparsec); UNIT(pc); UNIT(kpc); UNIT(Mpc);
	UNIT(Gpc); UNIT(kilometer); UNIT(km);
	UNIT(sec); UNIT(minute); UNIT(hour); UNIT(day); UNIT(year); UNIT(kiloyear); UNIT(Megayear); UNIT(Gigayear);
	UNIT(yr); UNIT(kyr); UNIT(Myr); UNIT(Gyr);
#undef UNIT

	bind_vector3<double>(m, "Vector3d");
	bind_vector3<float>(m, "Vector3f");
	py::implicitly_convertible<py::sequence, Vector3d>();

	m.def("nucleusId", &nucleusId);
	m.def("chargeNumber", &chargeNumber);
	m.def("massNumber", &massNumber);
	m.def("isNucleus", &isNucleus);

	py::class_<ParticleState>(m, "ParticleState")
	    .def(py::init<int, double, Vector3d, Vector3d>(), py::arg("id") = 0, py::arg("energy") = 0.,
	         py::arg("position") = Vector3d(0, 0, 0), py::arg("direction") = Vector3d(-1, 0, 0))
	    .def("setPosition", &ParticleState::setPosition)
	    .def("getPosition", &ParticleState::getPosition)
	    .def("setDirection", &ParticleState::setDirection)
	    .def("getDirection", &ParticleState::getDirection)
	    .def("setEnergy", &ParticleState::setEnergy)
	    .def("getEnergy", &ParticleState::getEnergy)
	    .def("getRigidity", &ParticleState::getRigidity)
	    .def("setId", &ParticleState::setId)
	    .def("getId", &ParticleState::getId)
	    .def("getCharge", &ParticleState::getCharge)
	    .def("setCharge", &ParticleState::setCharge)
	    .def("getMass", &ParticleState::getMass)
	    .def("getVelocity", &ParticleState::getVelocity)
	    .def("getMomentum", &ParticleState::getMomentum)
	    .def("getDescription", &ParticleState::getDescription)
	    .def("__repr__", &ParticleState::getDescription);

	py::class_<Candidate, ref_ptr<Candidate>>(m, "Candidate")
	    .def(py::init<int, double, Vector3d, Vector3d, double, double, std::string>(), py::arg("id") = 0,
	         py::arg("energy") = 0., py::arg("position") = Vector3d(0, 0, 0),
	         py::arg("direction") = Vector3d(-1, 0, 0), py::arg("z") = 0., py::arg("weight") = 1.,
	         py::arg("tagOrigin") = "PRIM")
	    .def(py::init<const ParticleState &>())
	    .def_readwrite("source", &Candidate::source)
	    .def_readwrite("created", &Candidate::created)
	    .def_readwrite("current", &Candidate::current)
	    .def_readwrite("previous", &Candidate::previous)
	    .def("isActive", &Candidate::isActive)
	    .def("setActive", &Candidate::setActive)
	    .def("getTrajectoryLength", &Candidate::getTrajectoryLength)
	    .def("setTrajectoryLength", &Candidate::setTrajectoryLength)
	    .def("getRedshift", &Candidate::getRedshift)
	    .def("setRedshift", &Candidate::setRedshift)
	    .def("getWeight", &Candidate::getWeight)
	    .def("setWeight", &Candidate::setWeight)
	    .def("getCurrentStep", &Candidate::getCurrentStep)
	    .def("setCurrentStep", &Candidate::setCurrentStep)
	    .def("getNextStep", &Candidate::getNextStep)
	    .def("setNextStep", &Candidate::setNextStep)
	    .def("limitNextStep", &Candidate::limitNextStep)
	    .def("getTime", &Candidate::getTime)
	    .def("setTime", &Candidate::setTime)
	    .def("getVelocity", &Candidate::getVelocity)
	    .def("getTagOrigin", &Candidate::getTagOrigin)
	    .def("setTagOrigin", &Candidate::setTagOrigin)
	    .def("setProperty", &Candidate::setProperty)
	    .def("getProperty", &Candidate::getProperty)
	    .def("hasProperty", &Candidate::hasProperty)
	    .def("removeProperty", &Candidate::removeProperty)
	    .def("getSerialNumber", &Candidate::getSerialNumber)
	    .def("getDiffusionSteps", &Candidate::getDiffusionSteps)
	    .def("setDiffusionSteps", &Candidate::setDiffusionSteps)
	    .def("setSerialNumber", &Candidate::setSerialNumber)
	    .def("getSourceSerialNumber", &Candidate::getSourceSerialNumber)
	    .def("getCreatedSerialNumber", &Candidate::getCreatedSerialNumber)
	    .def_static("setNextSerialNumber", &Candidate::setNextSerialNumber)
	    .def_static("getNextSerialNumber", &Candidate::getNextSerialNumber)
	    .def("restart", &Candidate::restart)
	    .def("getDescription", &Candidate::getDescription);

	py::class_<Module, PyModule, ref_ptr<Module>>(m, "Module")
	    .def(py::init<>())
	    .def("process", (void (Module::*)(Candidate *) const) & Module::process)
	    .def("getDescription", &Module::getDescription)
	    .def("setDescription", &Module::setDescription);

	py::class_<AbstractCondition, Module, ref_ptr<AbstractCondition>>(m, "AbstractCondition")
	    .def("setMakeRejectedInactive", &AbstractCondition::setMakeRejectedInactive)
	    .def("setRejectFlag", &AbstractCondition::setRejectFlag)
	    .def("onReject", &AbstractCondition::onReject);

	py::class_<MagneticField, PyMagneticField, ref_ptr<MagneticField>>(m, "MagneticField")
	    .def(py::init<>())
	    .def("getField", (Vector3d(MagneticField::*)(const Vector3d &) const) & MagneticField::getField)
	    .def("getField", (Vector3d(MagneticField::*)(const Vector3d &, double) const) & MagneticField::getField);

	py::class_<UniformMagneticField, MagneticField, ref_ptr<UniformMagneticField>>(m, "UniformMagneticField")
	    .def(py::init<const Vector3d &>());

	py::class_<DecoratorField, MagneticField, ref_ptr<DecoratorField>>(m, "DecoratorField");
	py::class_<PeriodicMagneticField, DecoratorField, ref_ptr<PeriodicMagneticField>>(m, "PeriodicMagneticField")
	    .def(py::init<ref_ptr<MagneticField>, const Vector3d &>(), py::arg("field"), py::arg("extends"))
	    .def(py::init<ref_ptr<MagneticField>, const Vector3d &, const Vector3d &, bool>(), py::arg("field"),
	         py::arg("extends"), py::arg("origin"), py::arg("reflective"))
	    .def("getOrigin", &PeriodicMagneticField::getOrigin)
	    .def("setOrigin", &PeriodicMagneticField::setOrigin)
	    .def("getExtends", &PeriodicMagneticField::getExtends)
	    .def("setExtends", &PeriodicMagneticField::setExtends)
	    .def("isReflective", &PeriodicMagneticField::isReflective)
	    .def("setReflective", &PeriodicMagneticField::setReflective);
	py::class_<MagneticFieldList, DecoratorField, ref_ptr<MagneticFieldList>>(m, "MagneticFieldList")
	    .def(py::init<>())
	    .def("addField", &MagneticFieldList::addField);
	py::class_<MagneticFieldEvolution, DecoratorField, ref_ptr<MagneticFieldEvolution>>(m, "MagneticFieldEvolution")
	    .def(py::init<ref_ptr<MagneticField>, double>(), py::arg("field"), py::arg("m"))
	    .def("getField", &MagneticFieldEvolution::getField, py::arg("position"), py::arg("z") = 0.0);

	py::enum_<interpolationType>(m, "interpolationType")
	    .value("TRILINEAR", TRILINEAR)
	    .value("TRICUBIC", TRICUBIC)
	    .value("NEAREST_NEIGHBOUR", NEAREST_NEIGHBOUR)
	    .export_values();

	py::class_<GridProperties, ref_ptr<GridProperties>>(m, "GridProperties")
	    .def(py::init<Vector3d, size_t, double>())
	    .def(py::init<Vector3d, size_t, size_t, size_t, double>())
	    .def(py::init<Vector3d, size_t, size_t, size_t, Vector3d>())
	    .def("setReflective", &GridProperties::setReflective)
	    .def("setClipVolume", &GridProperties::setClipVolume)
	    .def("setInterpolationType", &GridProperties::setInterpolationType)
	    .def_readwrite("Nx", &GridProperties::Nx)
	    .def_readwrite("Ny", &GridProperties::Ny)
	    .def_readwrite("Nz", &GridProperties::Nz)
	    .def_readwrite("origin", &GridProperties::origin)
	    .def_readwrite("spacing", &GridProperties::spacing);

	py::class_<Grid3f, ref_ptr<Grid3f>>(m, "Grid3f", py::buffer_protocol())
	    .def(py::init<Vector3d, size_t, double>())
	    .def(py::init<Vector3d, size_t, size_t, size_t, double>())
	    .def(py::init<Vector3d, size_t, size_t, size_t, Vector3d>())
	    .def(py::init<const GridProperties &>())
	    .def("getNx", &Grid3f::getNx)
	    .def("getNy", &Grid3f::getNy)
	    .def("getNz", &Grid3f::getNz)
	    .def("getOrigin", &Grid3f::getOrigin)
	    .def("getSpacing", &Grid3f::getSpacing)
	    .def("setReflective", &Grid3f::setReflective)
	    .def("isReflective", &Grid3f::isReflective)
	    .def("setClipVolume", &Grid3f::setClipVolume)
	    .def("getClipVolume", &Grid3f::getClipVolume)
	    .def("setInterpolationType", &Grid3f::setInterpolationType)
	    .def("getInterpolationType", &Grid3f::getInterpolationType)
	    .def("getValue", &Grid3f::getValue)
	    .def("setValue", &Grid3f::setValue)
	    .def("interpolate", &Grid3f::interpolate)
	    .def("positionFromIndex", &Grid3f::positionFromIndex)
	    .def("getSizeOf", &Grid3f::getSizeOf)
	    .def("changed", &Grid3f::changed)
	    .def_buffer([](Grid3f &g) -> py::buffer_info {
		    std::vector<Vector3f> &v = g.getGrid();
		    return py::buffer_info(v.data(), sizeof(float), py::format_descriptor<float>::format(), 4,
		                           {g.getNx(), g.getNy(), g.getNz(), (size_t)3},
		                           {sizeof(float) * 3 * g.getNy() * g.getNz(), sizeof(float) * 3 * g.getNz(),
		                            sizeof(float) * 3, sizeof(float)});
	    })
	    .def("array",
	         [](py::object self) {
		         Grid3f &g = self.cast<Grid3f &>();
		         g.changed();
		         return py::array_t<float>(py::buffer_info(
		                                       g.getGrid().data(), sizeof(float),
		                                       py::format_descriptor<float>::format(), 4,
		                                       {g.getNx(), g.getNy(), g.getNz(), (size_t)3},
		                                       {sizeof(float) * 3 * g.getNy() * g.getNz(),
		                                        sizeof(float) * 3 * g.getNz(), sizeof(float) * 3, sizeof(float)}),
		                                   self);
	         },
	         "Writable numpy view (Nx, Ny, Nz, 3) of the grid values.");

	py::class_<MagneticFieldGrid, MagneticField, ref_ptr<MagneticFieldGrid>>(m, "MagneticFieldGrid")
	    .def(py::init<ref_ptr<Grid3f>>())
	    .def("setGrid", &MagneticFieldGrid::setGrid)
	    .def("getGrid", &MagneticFieldGrid::getGrid);

	py::class_<Grid1f, ref_ptr<Grid1f>>(m, "Grid1f")
	    .def(py::init<Vector3d, size_t, double>(), py::arg("origin"), py::arg("N"), py::arg("spacing"))
	    .def(py::init<Vector3d, size_t, size_t, size_t, double>())
	    .def(py::init<Vector3d, size_t, size_t, size_t, Vector3d>())
	    .def(py::init<const GridProperties &>())
	    .def("setOrigin", &Grid1f::setOrigin)
	    .def("setSpacing", &Grid1f::setSpacing)
	    .def("setReflective", &Grid1f::setReflective)
	    .def("setClipVolume", &Grid1f::setClipVolume)
	    .def("setInterpolationType", &Grid1f::setInterpolationType)
	    .def("getInterpolationType", &Grid1f::getInterpolationType)
	    .def("getOrigin", &Grid1f::getOrigin)
	    .def("getSpacing", &Grid1f::getSpacing)
	    .def("getNx", &Grid1f::getNx)
	    .def("getNy", &Grid1f::getNy)
	    .def("getNz", &Grid1f::getNz)
	    .def("isReflective", &Grid1f::isReflective)
	    .def("getClipVolume", &Grid1f::getClipVolume)
	    .def("getValue", &Grid1f::getValue)
	    .def("setValue", &Grid1f::setValue)
	    .def("array",
	         [](py::object self) {
		         Grid1f &g = self.cast<Grid1f &>();
		         g.changed();
		         return py::array_t<float>(py::buffer_info(g.getGrid().data(), sizeof(float),
		                                                   py::format_descriptor<float>::format(), 3,
		                                                   {g.getNx(), g.getNy(), g.getNz()},
		                                                   {sizeof(float) * g.getNy() * g.getNz(),
		                                                    sizeof(float) * g.getNz(), sizeof(float)}),
		                                   self);
	         },
	         "Writable numpy view (Nx, Ny, Nz) of the grid values.");

	py::class_<ModulatedMagneticFieldGrid, MagneticField, ref_ptr<ModulatedMagneticFieldGrid>>(
	    m, "ModulatedMagneticFieldGrid")
	    .def(py::init<>())
	    .def(py::init<ref_ptr<Grid3f>, ref_ptr<Grid1f>>())
	    .def("setGrid", &ModulatedMagneticFieldGrid::setGrid)
	    .def("setModulationGrid", &ModulatedMagneticFieldGrid::setModulationGrid)
	    .def("getGrid", &ModulatedMagneticFieldGrid::getGrid)
	    .def("getModulationGrid", &ModulatedMagneticFieldGrid::getModulationGrid)
	    .def("setReflective", &ModulatedMagneticFieldGrid::setReflective);

	// grid tools (include/crpropa/GridTools.h)
	m.def("scaleGrid", &scaleGrid, py::arg("grid"), py::arg("a"));
	m.def("meanFieldVector", &meanFieldVector);
	m.def("meanFieldStrength", &meanFieldStrength);
	m.def("rmsFieldStrength", &rmsFieldStrength);
	m.def("rmsFieldStrengthPerAxis", &rmsFieldStrengthPerAxis);
	m.def("fromMagneticField", &fromMagneticField, py::arg("grid"), py::arg("field"));
	m.def("loadGrid", &loadGrid, py::arg("grid"), py::arg("filename"), py::arg("conversion") = 1.);
	m.def("dumpGrid", &dumpGrid, py::arg("grid"), py::arg("filename"), py::arg("conversion") = 1.);
	m.def("loadGridFromTxt", &loadGridFromTxt, py::arg("grid"), py::arg("filename"), py::arg("conversion") = 1.);
	m.def("loadGrid3fFromTxt", &loadGrid3fFromTxt, py::arg("filename"), py::arg("conversion") = 1.);
	m.def("dumpGridToTxt", &dumpGridToTxt, py::arg("grid"), py::arg("filename"), py::arg("conversion") = 1.,
	      py::arg("saveProp") = false);

	py::class_<TurbulenceSpectrum, ref_ptr<TurbulenceSpectrum>>(m, "TurbulenceSpectrum")
	    .def(py::init<double, double, double, double, double, double>(), py::arg("Brms"), py::arg("lMin"),
	         py::arg("lMax"), py::arg("lBendover") = 1., py::arg("sIndex") = 5. / 3., py::arg("qIndex") = 4.)
	    .def("getBrms", &TurbulenceSpectrum::getBrms)
	    .def("getLmin", &TurbulenceSpectrum::getLmin)
	    .def("getLmax", &TurbulenceSpectrum::getLmax)
	    .def("getLbendover", &TurbulenceSpectrum::getLbendover)
	    .def("getSindex", &TurbulenceSpectrum::getSindex)
	    .def("getQindex", &TurbulenceSpectrum::getQindex)
	    .def("energySpectrum", &TurbulenceSpectrum::energySpectrum)
	    .def("getCorrelationLength", &TurbulenceSpectrum::getCorrelationLength);

	py::class_<SimpleTurbulenceSpectrum, TurbulenceSpectrum, ref_ptr<SimpleTurbulenceSpectrum>>(
	    m, "SimpleTurbulenceSpectrum")
	    .def(py::init<double, double, double, double>(), py::arg("Brms"), py::arg("lMin"), py::arg("lMax"),
	         py::arg("sIndex") = 5. / 3.)
	    .def_static("turbulentCorrelationLength", &SimpleTurbulenceSpectrum::turbulentCorrelationLength);

	py::class_<TurbulentField, MagneticField, ref_ptr<TurbulentField>>(m, "TurbulentField")
	    .def("getBrms", &TurbulentField::getBrms)
	    .def("getCorrelationLength", &TurbulentField::getCorrelationLength);

	py::class_<GridTurbulence, TurbulentField, ref_ptr<GridTurbulence>>(m, "GridTurbulence")
	    .def(py::init<const TurbulenceSpectrum &, const GridProperties &, unsigned int>(), py::arg("spectrum"),
	         py::arg("gridProp"), py::arg("seed") = 0, py::call_guard<py::gil_scoped_release>())
	    .def("getGrid", &GridTurbulence::getGrid)
	    .def("getMeanFieldVector", &GridTurbulence::getMeanFieldVector)
	    .def("getMeanFieldStrength", &GridTurbulence::getMeanFieldStrength)
	    .def("getRmsFieldStrength", &GridTurbulence::getRmsFieldStrength)
	    .def("getRmsFieldStrengthPerAxis", &GridTurbulence::getRmsFieldStrengthPerAxis)
	    .def("dumpToFile", &GridTurbulence::dumpToFile)
	    .def_static("checkGridRequirements", &GridTurbulence::checkGridRequirements);

	py::class_<SimpleGridTurbulence, GridTurbulence, ref_ptr<SimpleGridTurbulence>>(m, "SimpleGridTurbulence")
	    .def(py::init<const SimpleTurbulenceSpectrum &, const GridProperties &, unsigned int>(), py::arg("spectrum"),
	         py::arg("gridProp"), py::arg("seed") = 0, py::call_guard<py::gil_scoped_release>())
	    .def_static("initTurbulence", &SimpleGridTurbulence::initTurbulence, py::arg("grid"), py::arg("Brms"),
	                py::arg("lMin"), py::arg("lMax"), py::arg("alpha"), py::arg("seed"),
	                py::call_guard<py::gil_scoped_release>());
	m.def("initTurbulence", &initTurbulence, py::arg("grid"), py::arg("Brms"), py::arg("lMin"), py::arg("lMax"),
	      py::arg("alpha") = -11 / 3., py::arg("seed") = 0, py::call_guard<py::gil_scoped_release>());
	m.def("turbulentCorrelationLength", &turbulentCorrelationLength, py::arg("lMin"), py::arg("lMax"),
	      py::arg("alpha") = -11 / 3.);

	py::class_<PlaneWaveTurbulence, TurbulentField, ref_ptr<PlaneWaveTurbulence>>(m, "PlaneWaveTurbulence")
	    .def(py::init<const TurbulenceSpectrum &, int, int>(), py::arg("spectrum"), py::arg("Nm") = 64,
	         py::arg("seed") = 0)
	    .def("getNm", &PlaneWaveTurbulence::getNm)
	    .def("getK", &PlaneWaveTurbulence::getK)
	    .def("getAk", &PlaneWaveTurbulence::getAk)
	    .def("getBeta", &PlaneWaveTurbulence::getBeta)
	    .def("getXi", &PlaneWaveTurbulence::getXi)
	    .def("getKappa", &PlaneWaveTurbulence::getKappa);

	py::class_<DevicePropagator, Module, ref_ptr<DevicePropagator>> dp(m, "DevicePropagator");
	py::enum_<DevicePropagator::Precision>(dp, "Precision")
	    .value("FP64", DevicePropagator::FP64)
	    .value("MIXED", DevicePropagator::MIXED)
	    .export_values();
	dp.def("setField", &DevicePropagator::setField)
	    .def("getField", &DevicePropagator::getField)
	    .def("setTolerance", &DevicePropagator::setTolerance)
	    .def("setMinimumStep", &DevicePropagator::setMinimumStep)
	    .def("setMaximumStep", &DevicePropagator::setMaximumStep)
	    .def("getTolerance", &DevicePropagator::getTolerance)
	    .def("getMinimumStep", &DevicePropagator::getMinimumStep)
	    .def("getMaximumStep", &DevicePropagator::getMaximumStep)
	    .def("setPrecision", &DevicePropagator::setPrecision)
	    .def("getPrecision", &DevicePropagator::getPrecision)
	    .def("setLaunchConfig", &DevicePropagator::setLaunchConfig, py::arg("block"), py::arg("blocksPerCU") = 0,
	         py::arg("stepsPerLaunch") = 0, py::arg("useGraph") = false)
	    .def("getLaunchConfig", &DevicePropagator::getLaunchConfig)
	    .def("setAutotune", &DevicePropagator::setAutotune, py::arg("on"), py::arg("probeSteps") = 32)
	    .def("getAutotune", &DevicePropagator::getAutotune);

	py::class_<PhasePoint>(m, "PhasePoint")
	    .def(py::init<>())
	    .def(py::init<const Vector3d &, const Vector3d &>())
	    .def(py::init<double>())
	    .def_readwrite("x", &PhasePoint::x)
	    .def_readwrite("u", &PhasePoint::u)
	    .def("__mul__", &PhasePoint::operator*)
	    .def("__iadd__", &PhasePoint::operator+=, py::return_value_policy::reference_internal);

	auto bp = py::class_<PropagationBP, DevicePropagator, ref_ptr<PropagationBP>>(m, "PropagationBP")
	    .def(py::init<ref_ptr<MagneticField>, double>(), py::arg("field") = ref_ptr<MagneticField>(),
	         py::arg("fixedStep") = 1. * kpc)
	    .def(py::init<ref_ptr<MagneticField>, double, double, double>(), py::arg("field"), py::arg("tolerance"),
	         py::arg("minStep"), py::arg("maxStep"))
	    .def("dY", &PropagationBP::dY, py::arg("pos"), py::arg("dir"), py::arg("step"), py::arg("z"), py::arg("q"),
	         py::arg("m"))
	    .def("errorEstimation", &PropagationBP::errorEstimation)
	    .def("tryStep", &PropagationBP::tryStep)
	    .def("getFieldAtPosition", &PropagationBP::getFieldAtPosition);
	bp.attr("Y") = m.attr("PhasePoint");

	auto ck = py::class_<PropagationCK, DevicePropagator, ref_ptr<PropagationCK>>(m, "PropagationCK")
	    .def(py::init<ref_ptr<MagneticField>, double, double, double>(), py::arg("field") = ref_ptr<MagneticField>(),
	         py::arg("tolerance") = 1e-4, py::arg("minStep") = 0.1 * kpc, py::arg("maxStep") = 1 * Gpc)
	    .def("dYdt", &PropagationCK::dYdt)
	    .def("tryStep", &PropagationCK::tryStep)
	    .def("getFieldAtPosition", &PropagationCK::getFieldAtPosition);
	ck.attr("Y") = m.attr("PhasePoint");

	// ---- JF12 (src/magneticField/JF12Field.cpp)
	py::class_<JF12Field, MagneticField, ref_ptr<JF12Field>>(m, "JF12Field")
	    .def(py::init<>())
	    .def("randomStriated", &JF12Field::randomStriated, py::arg("seed") = 0)
	    .def("randomTurbulent", &JF12Field::randomTurbulent, py::arg("seed") = 0)
	    .def("setStriatedGrid", &JF12Field::setStriatedGrid)
	    .def("setTurbulentGrid", &JF12Field::setTurbulentGrid)
	    .def("getStriatedGrid", &JF12Field::getStriatedGrid)
	    .def("getTurbulentGrid", &JF12Field::getTurbulentGrid)
	    .def("setUseRegularField", &JF12Field::setUseRegularField)
	    .def("setUseDiskField", &JF12Field::setUseDiskField)
	    .def("setUseToroidalHaloField", &JF12Field::setUseToroidalHaloField)
	    .def("setUseXField", &JF12Field::setUseXField)
	    .def("setUseStriatedField", &JF12Field::setUseStriatedField)
	    .def("setUseTurbulentField", &JF12Field::setUseTurbulentField)
	    .def("isUsingRegularField", &JF12Field::isUsingRegularField)
	    .def("isUsingDiskField", &JF12Field::isUsingDiskField)
	    .def("isUsingToroidalHaloField", &JF12Field::isUsingToroidalHaloField)
	    .def("isUsingXField", &JF12Field::isUsingXField)
	    .def("isUsingStriatedField", &JF12Field::isUsingStriatedField)
	    .def("isUsingTurbulentField", &JF12Field::isUsingTurbulentField)
	    .def("getRegularField", &JF12Field::getRegularField)
	    .def("getStriatedField", &JF12Field::getStriatedField)
	    .def("getTurbulentField", &JF12Field::getTurbulentField);
	py::class_<PlanckJF12bField, JF12Field, ref_ptr<PlanckJF12bField>>(m, "PlanckJF12bField").def(py::init<>());

	// ---- advection fields and DiffusionSDE
	py::class_<AdvectionField, ref_ptr<AdvectionField>>(m, "AdvectionField")
	    .def("getField", &AdvectionField::getField, py::arg("position"), py::arg("time") = 0.)
	    .def("getDivergence", &AdvectionField::getDivergence, py::arg("position"), py::arg("time") = 0.);
	py::class_<UniformAdvectionField, AdvectionField, ref_ptr<UniformAdvectionField>>(m, "UniformAdvectionField")
	    .def(py::init<const Vector3d &>())
	    .def("getDescription", &UniformAdvectionField::getDescription);

	py::class_<ConstantSphericalAdvectionField, AdvectionField, ref_ptr<ConstantSphericalAdvectionField>>(
	    m, "ConstantSphericalAdvectionField")
	    .def(py::init<const Vector3d, double>())
	    .def("setOrigin", &ConstantSphericalAdvectionField::setOrigin)
	    .def("setVWind", &ConstantSphericalAdvectionField::setVWind)
	    .def("getOrigin", &ConstantSphericalAdvectionField::getOrigin)
	    .def("getVWind", &ConstantSphericalAdvectionField::getVWind)
	    .def("getDescription", &ConstantSphericalAdvectionField::getDescription);
	py::class_<SphericalAdvectionField, AdvectionField, ref_ptr<SphericalAdvectionField>>(m, "SphericalAdvectionField")
	    .def(py::init<const Vector3d, double, double, double, double>())
	    .def("getV", &SphericalAdvectionField::getV)
	    .def("setOrigin", &SphericalAdvectionField::setOrigin)
	    .def("setRadius", &SphericalAdvectionField::setRadius)
	    .def("setVMax", &SphericalAdvectionField::setVMax)
	    .def("setTau", &SphericalAdvectionField::setTau)
	    .def("setAlpha", &SphericalAdvectionField::setAlpha)
	    .def("getOrigin", &SphericalAdvectionField::getOrigin)
	    .def("getRadius", &SphericalAdvectionField::getRadius)
	    .def("getVMax", &SphericalAdvectionField::getVMax)
	    .def("getTau", &SphericalAdvectionField::getTau)
	    .def("getAlpha", &SphericalAdvectionField::getAlpha)
	    .def("getDescription", &SphericalAdvectionField::getDescription);
	py::class_<AdvectionFieldList, AdvectionField, ref_ptr<AdvectionFieldList>>(m, "AdvectionFieldList")
	    .def(py::init<>())
	    .def("addField", &AdvectionFieldList::addField);
	py::class_<OneDimensionalCartesianShock, AdvectionField, ref_ptr<OneDimensionalCartesianShock>>(
	    m, "OneDimensionalCartesianShock")
	    .def(py::init<double, double, double>(), py::arg("compressionRatio"), py::arg("vUp"), py::arg("lShock"))
	    .def("setComp", &OneDimensionalCartesianShock::setComp)
	    .def("setVup", &OneDimensionalCartesianShock::setVup)
	    .def("setShockwidth", &OneDimensionalCartesianShock::setShockwidth)
	    .def("getComp", &OneDimensionalCartesianShock::getComp)
	    .def("getVup", &OneDimensionalCartesianShock::getVup)
	    .def("getShockwidth", &OneDimensionalCartesianShock::getShockwidth)
	    .def("getDescription", &OneDimensionalCartesianShock::getDescription);
	py::class_<OneDimensionalSphericalShock, AdvectionField, ref_ptr<OneDimensionalSphericalShock>>(
	    m, "OneDimensionalSphericalShock")
	    .def(py::init<double, double, double, double, bool>(), py::arg("rShock"), py::arg("vUp"),
	         py::arg("compressionRatio"), py::arg("lShock"), py::arg("coolUpstream"))
	    .def("setComp", &OneDimensionalSphericalShock::setComp)
	    .def("setVup", &OneDimensionalSphericalShock::setVup)
	    .def("setShockwidth", &OneDimensionalSphericalShock::setShockwidth)
	    .def("setShockRadius", &OneDimensionalSphericalShock::setShockRadius)
	    .def("setCooling", &OneDimensionalSphericalShock::setCooling)
	    .def("getComp", &OneDimensionalSphericalShock::getComp)
	    .def("getVup", &OneDimensionalSphericalShock::getVup)
	    .def("getShockwidth", &OneDimensionalSphericalShock::getShockwidth)
	    .def("getShockRadius", &OneDimensionalSphericalShock::getShockRadius)
	    .def("getCooling", &OneDimensionalSphericalShock::getCooling)
	    .def("getDescription", &OneDimensionalSphericalShock::getDescription);
	py::class_<ObliqueAdvectionShock, AdvectionField, ref_ptr<ObliqueAdvectionShock>>(m, "ObliqueAdvectionShock")
	    .def(py::init<double, double, double, double>(), py::arg("compressionRatio"), py::arg("vXUp"), py::arg("vY"),
	         py::arg("lShock"))
	    .def("setComp", &ObliqueAdvectionShock::setComp)
	    .def("setVup", &ObliqueAdvectionShock::setVup)
	    .def("setVy", &ObliqueAdvectionShock::setVy)
	    .def("setShockwidth", &ObliqueAdvectionShock::setShockwidth)
	    .def("getComp", &ObliqueAdvectionShock::getComp)
	    .def("getVup", &ObliqueAdvectionShock::getVup)
	    .def("getVy", &ObliqueAdvectionShock::getVy)
	    .def("getShockwidth", &ObliqueAdvectionShock::getShockwidth)
	    .def("getDescription", &ObliqueAdvectionShock::getDescription);
	py::class_<SphericalAdvectionShock, AdvectionField, ref_ptr<SphericalAdvectionShock>>(m, "SphericalAdvectionShock")
	    .def(py::init<const Vector3d, double, double, double>(), py::arg("origin"), py::arg("r_0"), py::arg("v_0"),
	         py::arg("l"))
	    .def("setOrigin", &SphericalAdvectionShock::setOrigin)
	    .def("setR0", &SphericalAdvectionShock::setR0)
	    .def("setV0", &SphericalAdvectionShock::setV0)
	    .def("setLambda", &SphericalAdvectionShock::setLambda)
	    .def("setRRot", &SphericalAdvectionShock::setRRot)
	    .def("setAzimuthalSpeed", &SphericalAdvectionShock::setAzimuthalSpeed)
	    .def("getOrigin", &SphericalAdvectionShock::getOrigin)
	    .def("getR0", &SphericalAdvectionShock::getR0)
	    .def("getV0", &SphericalAdvectionShock::getV0)
	    .def("getLambda", &SphericalAdvectionShock::getLambda)
	    .def("getRRot", &SphericalAdvectionShock::getRRot)
	    .def("getAzimuthalSpeed", &SphericalAdvectionShock::getAzimuthalSpeed)
	    .def("getDescription", &SphericalAdvectionShock::getDescription);

	py::class_<DiffusionSDE, DevicePropagator, ref_ptr<DiffusionSDE>>(m, "DiffusionSDE")
	    .def(py::init<ref_ptr<MagneticField>, double, double, double, double>(), py::arg("magneticField"),
	         py::arg("tolerance") = 1e-4, py::arg("minStep") = 10 * pc, py::arg("maxStep") = 1 * kpc,
	         py::arg("epsilon") = 0.1)
	    .def(py::init<ref_ptr<MagneticField>, ref_ptr<AdvectionField>, double, double, double, double>(),
	         py::arg("magneticField"), py::arg("advectionField"), py::arg("tolerance") = 1e-4,
	         py::arg("minStep") = 10 * pc, py::arg("maxStep") = 1 * kpc, py::arg("epsilon") = 0.1)
	    .def("setEpsilon", &DiffusionSDE::setEpsilon)
	    .def("setAlpha", &DiffusionSDE::setAlpha)
	    .def("setScale", &DiffusionSDE::setScale)
	    .def("setMagneticField", &DiffusionSDE::setMagneticField)
	    .def("setAdvectionField", &DiffusionSDE::setAdvectionField)
	    .def("getEpsilon", &DiffusionSDE::getEpsilon)
	    .def("getAlpha", &DiffusionSDE::getAlpha)
	    .def("getScale", &DiffusionSDE::getScale)
	    .def("getMagneticField", &DiffusionSDE::getMagneticField)
	    .def("getAdvectionField", &DiffusionSDE::getAdvectionField)
	    .def("getMagneticFieldAtPosition", &DiffusionSDE::getMagneticFieldAtPosition)
	    .def("getAdvectionFieldAtPosition", &DiffusionSDE::getAdvectionFieldAtPosition)
	    .def("setSeed", &DiffusionSDE::setSeed)
	    .def("getSeed", &DiffusionSDE::getSeed)
	    .def("getDescription", &DiffusionSDE::getDescription);

	py::class_<MaximumTrajectoryLength, AbstractCondition, ref_ptr<MaximumTrajectoryLength>>(
	    m, "MaximumTrajectoryLength")
	    .def(py::init<double>(), py::arg("maxLength") = 0.)
	    .def("setMaximumTrajectoryLength", &MaximumTrajectoryLength::setMaximumTrajectoryLength)
	    .def("getMaximumTrajectoryLength", &MaximumTrajectoryLength::getMaximumTrajectoryLength)
	    .def("addObserverPosition", &MaximumTrajectoryLength::addObserverPosition)
	    .def("getObserverPositions", &MaximumTrajectoryLength::getObserverPositions);

	py::class_<PeriodicBox, Module, ref_ptr<PeriodicBox>>(m, "PeriodicBox")
	    .def(py::init<>())
	    .def(py::init<Vector3d, Vector3d>())
	    .def("setOrigin", &PeriodicBox::setOrigin)
	    .def("setSize", &PeriodicBox::setSize)
	    .def("process", [](const PeriodicBox &b, Candidate *c) { b.process(c); });

	py::class_<SphericalBoundary, AbstractCondition, ref_ptr<SphericalBoundary>>(m, "SphericalBoundary")
	    .def(py::init<>())
	    .def(py::init<Vector3d, double>())
	    .def("setCenter", &SphericalBoundary::setCenter)
	    .def("setRadius", &SphericalBoundary::setRadius)
	    .def("setMargin", &SphericalBoundary::setMargin)
	    .def("setLimitStep", &SphericalBoundary::setLimitStep)
	    .def("process", [](const SphericalBoundary &b, Candidate *c) { b.process(c); });

	// observers (include/crpropa/module/Observer.h)
	py::enum_<DetectionState>(m, "DetectionState")
	    .value("DETECTED", DETECTED)
	    .value("VETO", VETO)
	    .value("NOTHING", NOTHING)
	    .export_values();
	py::class_<ObserverFeature, ref_ptr<ObserverFeature>>(m, "ObserverFeature")
	    .def("getDescription", &ObserverFeature::getDescription);
	py::class_<ObserverTimeEvolution, ObserverFeature, ref_ptr<ObserverTimeEvolution>>(m, "ObserverTimeEvolution")
	    .def(py::init<>())
	    .def(py::init<double, double, double>(), py::arg("min"), py::arg("dist"), py::arg("numb"))
	    .def(py::init<double, double, double, bool>(), py::arg("min"), py::arg("max"), py::arg("numb"),
	         py::arg("log"))
	    .def(py::init<const std::vector<double> &>())
	    .def("checkDetection", &ObserverTimeEvolution::checkDetection)
	    .def("addTime", &ObserverTimeEvolution::addTime)
	    .def("addTimeRange", &ObserverTimeEvolution::addTimeRange, py::arg("min"), py::arg("max"), py::arg("numb"),
	         py::arg("log") = false)
	    .def("setTimes", &ObserverTimeEvolution::setTimes)
	    .def("getTimes", &ObserverTimeEvolution::getTimes)
	    .def("getTime", &ObserverTimeEvolution::getTime)
	    .def("clear", &ObserverTimeEvolution::clear)
	    .def("empty", &ObserverTimeEvolution::empty)
	    .def("constructDetListIfEmpty", &ObserverTimeEvolution::constructDetListIfEmpty)
	    .def("setMinimum", &ObserverTimeEvolution::setMinimum)
	    .def("setMaximum", &ObserverTimeEvolution::setMaximum)
	    .def("setNIntervals", &ObserverTimeEvolution::setNIntervals)
	    .def("setIsLogarithmicScaling", &ObserverTimeEvolution::setIsLogarithmicScaling)
	    .def("getMinimum", &ObserverTimeEvolution::getMinimum)
	    .def("getMaximum", &ObserverTimeEvolution::getMaximum)
	    .def("getNIntervals", &ObserverTimeEvolution::getNIntervals)
	    .def("getIsLogarithmicScaling", &ObserverTimeEvolution::getIsLogarithmicScaling);
	py::class_<Observer, Module, ref_ptr<Observer>>(m, "Observer")
	    .def(py::init<>())
	    .def("add", &Observer::add)
	    .def("onDetection", &Observer::onDetection, py::arg("action"), py::arg("clone") = false)
	    .def("setDeactivateOnDetection", &Observer::setDeactivateOnDetection)
	    .def("setFlag", &Observer::setFlag)
	    .def("process", [](const Observer &o, Candidate *c) { o.process(c); });

	// outputs (include/crpropa/module/Output.h, TextOutput.h, ParticleCollector.h)
	{
		py::class_<Output, Module, ref_ptr<Output>> out(m, "Output");
		py::enum_<Output::OutputColumn>(out, "OutputColumn")
		    .value("TrajectoryLengthColumn", Output::TrajectoryLengthColumn)
		    .value("ColumnDensityColumn", Output::ColumnDensityColumn)
		    .value("RedshiftColumn", Output::RedshiftColumn)
		    .value("CurrentIdColumn", Output::CurrentIdColumn)
		    .value("CurrentEnergyColumn", Output::CurrentEnergyColumn)
		    .value("CurrentPositionColumn", Output::CurrentPositionColumn)
		    .value("CurrentDirectionColumn", Output::CurrentDirectionColumn)
		    .value("SourceIdColumn", Output::SourceIdColumn)
		    .value("SourceEnergyColumn", Output::SourceEnergyColumn)
		    .value("SourcePositionColumn", Output::SourcePositionColumn)
		    .value("SourceDirectionColumn", Output::SourceDirectionColumn)
		    .value("CreatedIdColumn", Output::CreatedIdColumn)
		    .value("CreatedEnergyColumn", Output::CreatedEnergyColumn)
		    .value("CreatedPositionColumn", Output::CreatedPositionColumn)
		    .value("CreatedDirectionColumn", Output::CreatedDirectionColumn)
		    .value("CandidateTagColumn", Output::CandidateTagColumn)
		    .value("SerialNumberColumn", Output::SerialNumberColumn)
		    .value("WeightColumn", Output::WeightColumn)
		    .value("TimeColumn", Output::TimeColumn)
		    .export_values();
		py::enum_<Output::OutputType>(out, "OutputType")
		    .value("Trajectory1D", Output::Trajectory1D)
		    .value("Trajectory3D", Output::Trajectory3D)
		    .value("Event1D", Output::Event1D)
		    .value("Event3D", Output::Event3D)
		    .value("Everything", Output::Everything)
		    .export_values();
		out.def(py::init<>())
		    .def(py::init<Output::OutputType>())
		    .def("setEnergyScale", &Output::setEnergyScale)
		    .def("getEnergyScale", &Output::getEnergyScale)
		    .def("setLengthScale", &Output::setLengthScale)
		    .def("getLengthScale", &Output::getLengthScale)
		    .def("setTimeScale", &Output::setTimeScale)
		    .def("getTimeScale", &Output::getTimeScale)
		    .def("setOutputType", &Output::setOutputType)
		    .def("set", &Output::set)
		    .def("enable", &Output::enable)
		    .def("disable", &Output::disable)
		    .def("enableAll", &Output::enableAll)
		    .def("disableAll", &Output::disableAll)
		    .def("set1D", &Output::set1D)
		    .def("size", &Output::size)
		    .def("enableProperty",
		         [](Output &o, const std::string &name, py::object dflt, const std::string &comment) {
			         o.enableProperty(name, py::str(dflt), comment);
		         },
		         py::arg("property"), py::arg("defaultValue"), py::arg("comment") = "")
		    .def("process", [](const Output &o, Candidate *c) { o.process(c); });
	}
	py::class_<TextOutput, Output, ref_ptr<TextOutput>>(m, "TextOutput")
	    .def(py::init<>())
	    .def(py::init<Output::OutputType>())
	    .def(py::init<const std::string &>())
	    .def(py::init<const std::string &, Output::OutputType>())
	    .def("process", [](const TextOutput &o, Candidate *c) { o.process(c); })
	    .def("close", &TextOutput::close)
	    .def("gzip", &TextOutput::gzip)
	    .def("enableRandomSeeds", &TextOutput::enableRandomSeeds)
	    .def_static("load", &TextOutput::load);
	py::class_<ParticleCollector, Module, ref_ptr<ParticleCollector>>(m, "ParticleCollector")
	    .def(py::init<>())
	    .def(py::init<std::size_t>())
	    .def(py::init<std::size_t, bool>())
	    .def(py::init<std::size_t, bool, bool>())
	    .def("process", [](const ParticleCollector &p, Candidate *c) { p.process(c); })
	    .def("reprocess", &ParticleCollector::reprocess)
	    .def("dump", &ParticleCollector::dump)
	    .def("load", &ParticleCollector::load)
	    .def("size", &ParticleCollector::size)
	    .def("__len__", &ParticleCollector::size)
	    .def("__getitem__",
	         [](const ParticleCollector &p, std::size_t i) {
		         if (i >= p.size())
			         throw py::index_error();
		         return p[i];
	         })
	    .def("__iter__",
	         [](ParticleCollector &p) { return py::make_iterator(p.getContainer().begin(), p.getContainer().end()); },
	         py::keep_alive<0, 1>())
	    .def("clearContainer", &ParticleCollector::clearContainer)
	    .def("getContainer", [](const ParticleCollector &p) { return p.getContainer(); })
	    .def("setClone", &ParticleCollector::setClone)
	    .def("getClone", &ParticleCollector::getClone)
	    .def("getTrajectory", [](const ParticleCollector &p, ModuleList *ml, std::size_t i, Module *out) {
		    p.getTrajectory(ml, i, out);
	    });

	py::class_<ModuleList, Module, ref_ptr<ModuleList>>(m, "ModuleList")
	    .def(py::init<>())
	    .def("add", &ModuleList::add, py::keep_alive<1, 2>())
	    .def("remove", &ModuleList::remove)
	    .def("size", &ModuleList::size)
	    .def("__len__", &ModuleList::size)
	    .def("__getitem__", &ModuleList::operator[])
	    .def("setShowProgress", &ModuleList::setShowProgress, py::arg("show") = true)
	    .def("setMaximumSteps", &ModuleList::setMaximumSteps)
	    .def("process", (void (ModuleList::*)(Candidate *) const) & ModuleList::process)
	    .def("run",
	         [](ModuleList &ml, ref_ptr<Candidate> c, bool recursive, bool secondariesFirst) {
		         py::gil_scoped_release nogil;
		         ml.run(c, recursive, secondariesFirst);
	         },
	         py::arg("candidate"), py::arg("recursive") = true, py::arg("secondariesFirst") = false)
	    .def("run",
	         [](ModuleList &ml, std::vector<ref_ptr<Candidate>> v, bool recursive, bool secondariesFirst) {
		         py::gil_scoped_release nogil;
		         ml.run(&v, recursive, secondariesFirst);
	         },
	         py::arg("candidates"), py::arg("recursive") = true, py::arg("secondariesFirst") = false)
	    .def("lastAcceptedSteps", &ModuleList::lastAcceptedSteps)
	    .def("lastKernelMs", &ModuleList::lastKernelMs)
	    .def("getDescription", &ModuleList::getDescription)
	    .def("showModules", &ModuleList::showModules);

	py::class_<Random>(m, "Random")
	    .def(py::init<>())
	    .def(py::init<uint32_t>())
	    .def("seed", (void (Random::*)(uint32_t)) & Random::seed)
	    .def("randInt", &Random::randInt)
	    .def("rand", &Random::rand)
	    .def("randUniform", &Random::randUniform)
	    .def("randNorm", &Random::randNorm, py::arg("mean") = 0.0, py::arg("variance") = 1.0)
	    .def("randVector", &Random::randVector)
	    .def("randPowerLaw", &Random::randPowerLaw)
	    .def("randBin", &Random::randBin)
	    .def_static("instance", &Random::instance, py::return_value_policy::reference);
}
