// Output, TextOutput and ParticleCollector (src/module/Output.cpp, TextOutput.cpp,
// ParticleCollector.cpp): the column selection of each output type, the text header and
// the printf formats of every column, so that files read the same as the reference's.
#include "crpropa/module/Output.h"
#include "crpropa/Units.h"

#include <cinttypes>
#include <cstdio>
#include <locale>
#include <mutex>
#include <sstream>
#include <stdexcept>

namespace crpropa {

const std::string g_GIT_DESC = "crpropa3_amd 0.1 (MI355X; CRPropa 3 interface)";

static bool ends_with(const std::string &s, const std::string &suffix) {
	return s.size() >= suffix.size() && s.compare(s.size() - suffix.size(), suffix.size(), suffix) == 0;
}

// ------------------------------------------------------------------ Output
Output::Output()
    : outputName(OutputTypeName(Everything)), lengthScale(Mpc), timeScale(Myr), energyScale(EeV),
      oneDimensional(false), count(0) {
	enableAll();
}

Output::Output(OutputType outputType)
    : outputName(OutputTypeName(outputType)), lengthScale(Mpc), timeScale(Myr), energyScale(EeV),
      oneDimensional(false), count(0) {
	setOutputType(outputType);
}

std::string Output::OutputTypeName(OutputType t) {
	switch (t) {
	case Trajectory1D: return "Trajectory1D";
	case Event1D: return "Event1D";
	case Trajectory3D: return "Trajectory3D";
	case Event3D: return "Event3D";
	default: return "Everything";
	}
}

// the columns of each output type (src/module/Output.cpp:37-79)
void Output::setOutputType(OutputType t) {
	modify();
	switch (t) {
	case Trajectory1D: // X, ID, E
		for (OutputColumn c : {CurrentPositionColumn, CurrentIdColumn, CurrentEnergyColumn})
			fields.set(c);
		oneDimensional = true;
		break;
	case Event1D: // D, ID, E, ID0, E0
		for (OutputColumn c : {TrajectoryLengthColumn, CurrentIdColumn, CurrentEnergyColumn, SourceIdColumn,
		                       SourceEnergyColumn})
			fields.set(c);
		oneDimensional = true;
		break;
	case Trajectory3D: // D, ID, E, X, Y, Z, Px, Py, Pz
		for (OutputColumn c : {TrajectoryLengthColumn, CurrentIdColumn, CurrentEnergyColumn, CurrentPositionColumn,
		                       CurrentDirectionColumn})
			fields.set(c);
		oneDimensional = false;
		break;
	case Event3D: // Trajectory3D and the source's ID0, E0, X0.., P0x..
		for (OutputColumn c : {TrajectoryLengthColumn, CurrentIdColumn, CurrentEnergyColumn, CurrentPositionColumn,
		                       CurrentDirectionColumn, SourceIdColumn, SourceEnergyColumn, SourcePositionColumn,
		                       SourceDirectionColumn})
			fields.set(c);
		oneDimensional = false;
		break;
	case Everything:
		fields.set();
		oneDimensional = false;
		break;
	default:
		throw std::runtime_error("Output: unknown output type");
	}
}

void Output::setEnergyScale(double s) { modify(); energyScale = s; }
void Output::setLengthScale(double s) { modify(); lengthScale = s; }
void Output::setTimeScale(double s) { modify(); timeScale = s; }
void Output::set1D(bool v) { modify(); oneDimensional = v; }
void Output::enable(OutputColumn f) { modify(); fields.set(f, true); }
void Output::disable(OutputColumn f) { modify(); fields.set(f, false); }
void Output::set(OutputColumn f, bool v) { modify(); fields.set(f, v); }
void Output::enableAll() { modify(); fields.set(); }
void Output::disableAll() { modify(); fields.reset(); }

void Output::enableProperty(const std::string &property, const std::string &defaultValue,
                            const std::string &comment) {
	modify();
	properties.push_back(Property{property, comment, defaultValue});
}

void Output::modify() {
	if (count > 0)
		throw std::runtime_error("Output: cannot change Output parameters after data has been written to file.");
}

void Output::process(Candidate *) const { count++; }

// ------------------------------------------------------------------ TextOutput
TextOutput::TextOutput() : Output(), out(&std::cout), storeRandomSeeds(false) {}
TextOutput::TextOutput(OutputType t) : Output(t), out(&std::cout), storeRandomSeeds(false) {}
TextOutput::TextOutput(std::ostream &o) : Output(), out(&o), storeRandomSeeds(false) {}
TextOutput::TextOutput(std::ostream &o, OutputType t) : Output(t), out(&o), storeRandomSeeds(false) {}

TextOutput::TextOutput(const std::string &fn)
    : Output(), outfile(fn.c_str(), std::ios::binary), out(&outfile), filename(fn), storeRandomSeeds(false) {
	if (!outfile.is_open())
		throw std::runtime_error(std::string("Cannot create file: ") + fn);
	if (ends_with(fn, ".gz"))
		gzip();
}

TextOutput::TextOutput(const std::string &fn, OutputType t)
    : Output(t), outfile(fn.c_str(), std::ios::binary), out(&outfile), filename(fn), storeRandomSeeds(false) {
	if (!outfile.is_open())
		throw std::runtime_error(std::string("Cannot create file: ") + fn);
	if (ends_with(fn, ".gz"))
		gzip();
}

TextOutput::~TextOutput() { close(); }


void TextOutput::gzip() { throw std::runtime_error("CRPropa was built without Zlib compression!"); }

void TextOutput::close() {
	if (out)
		out->flush();
}

// the header: column names, then one comment line per quantity (src/module/TextOutput.cpp:57-158)
void TextOutput::printHeader() const {
	const bool d3 = !oneDimensional;
	std::ostream &o = *out;
	o << "#";
	if (fields.test(TrajectoryLengthColumn)) o << "\tD";
	if (fields.test(TimeColumn)) o << "\ttime";
	if (fields.test(RedshiftColumn)) o << "\tz";
	if (fields.test(SerialNumberColumn)) o << "\tSN";
	if (fields.test(CurrentIdColumn)) o << "\tID";
	if (fields.test(CurrentEnergyColumn)) o << "\tE";
	if (fields.test(CurrentPositionColumn)) o << (d3 ? "\tX\tY\tZ" : "\tX");
	if (fields.test(CurrentDirectionColumn) && d3) o << "\tPx\tPy\tPz";
	if (fields.test(SerialNumberColumn)) o << "\tSN0";
	if (fields.test(SourceIdColumn)) o << "\tID0";
	if (fields.test(SourceEnergyColumn)) o << "\tE0";
	if (fields.test(SourcePositionColumn)) o << (d3 ? "\tX0\tY0\tZ0" : "\tX0");
	if (fields.test(SourceDirectionColumn) && d3) o << "\tP0x\tP0y\tP0z";
	if (fields.test(SerialNumberColumn)) o << "\tSN1";
	if (fields.test(CreatedIdColumn)) o << "\tID1";
	if (fields.test(CreatedEnergyColumn)) o << "\tE1";
	if (fields.test(CreatedPositionColumn)) o << (d3 ? "\tX1\tY1\tZ1" : "\tX1");
	if (fields.test(CreatedDirectionColumn) && d3) o << "\tP1x\tP1y\tP1z";
	if (fields.test(WeightColumn)) o << "\tW";
	if (fields.test(CandidateTagColumn)) o << "\ttag";
	for (const Property &p : properties)
		o << "\t" << p.name;
	o << "\n#\n";
	if (fields.test(TrajectoryLengthColumn))
		o << "# D             Trajectory length [" << lengthScale / Mpc << " Mpc]\n";
	if (fields.test(TimeColumn))
		o << "# time          Time [" << timeScale / Myr << " Myr]\n";
	if (fields.test(RedshiftColumn))
		o << "# z             Redshift\n";
	if (fields.test(SerialNumberColumn))
		o << "# SN/SN0/SN1    Serial number. Unique (within this run) id of the particle.\n";
	if (fields.test(CurrentIdColumn) || fields.test(CreatedIdColumn) || fields.test(SourceIdColumn))
		o << "# ID/ID0/ID1    Particle type (PDG MC numbering scheme)\n";
	if (fields.test(CurrentEnergyColumn) || fields.test(CreatedEnergyColumn) || fields.test(SourceEnergyColumn))
		o << "# E/E0/E1       Energy [" << energyScale / EeV << " EeV]\n";
	if (fields.test(CurrentPositionColumn) || fields.test(CreatedPositionColumn) || fields.test(SourcePositionColumn))
		o << "# X/X0/X1...    Position [" << lengthScale / Mpc << " Mpc]\n";
	if (fields.test(CurrentDirectionColumn) || fields.test(CreatedDirectionColumn) ||
	    fields.test(SourceDirectionColumn))
		o << "# Px/P0x/P1x... Heading (unit vector of momentum)\n";
	if (fields.test(WeightColumn))
		o << "# W             Weights \n";
	if (fields.test(CandidateTagColumn)) {
		o << "# tag           Candidate tag can be given by the source feature (user defined tag) or by the "
		     "following interaction process \n";
		o << "#\tES  \tElasticScattering \n#\tEPP \tElectronPairProduction \n#\tEMPP\tEMPairProduction\n"
		  << "#\tEMDP\tEMDoublePairProduction\n#\tEMTP\tEMTripletPairProduction \n#\tEMIC\tEMInverseComptonScattering\n"
		  << "#\tND  \tNuclearDecay\n#\tPD  \tPhotoDisintegration\n#\tPPP  \tPhotoPionProduction\n"
		  << "#\tSYN \tSynchrotronRadiation\n#\tPRIM/SEC\t primary / secondary particle\n";
	}
	for (const Property &p : properties)
		o << "# " << p.name << " " << p.comment << "\n";
	o << "# no index = current, 0 = at source, 1 = at point of creation\n#\n";
	o << "# CRPropa version: " << g_GIT_DESC << "\n#\n";
}

namespace {
struct Line {
	char buf[2048];
	size_t p = 0;
	template <typename... A> void put(const char *fmt, A... a) {
		int k = std::snprintf(buf + p, sizeof(buf) - p, fmt, a...);
		if (k > 0)
			p = std::min(sizeof(buf) - 1, p + (size_t)k);
	}
	void vec(const Vector3d &v) { put("%8.5E\t%8.5E\t%8.5E\t", v.x, v.y, v.z); }
};
std::mutex g_fileOutput; // the reference's omp critical(FileOutput)
} // namespace

// one line per candidate (src/module/TextOutput.cpp:160-292)
void TextOutput::process(Candidate *c) const {
	if (fields.none() && properties.empty())
		return;
	Line L;
	const bool d3 = !oneDimensional;
	std::locale old = std::locale::global(std::locale::classic());
	if (fields.test(TrajectoryLengthColumn)) L.put("%8.5E\t", c->getTrajectoryLength() / lengthScale);
	if (fields.test(TimeColumn)) L.put("%8.5E\t", c->getTime() / timeScale);
	if (fields.test(RedshiftColumn)) L.put("%1.5E\t", c->getRedshift());
	if (fields.test(SerialNumberColumn)) L.put("%10" PRIu64 "\t", c->getSerialNumber());
	if (fields.test(CurrentIdColumn)) L.put("%10i\t", c->current.getId());
	if (fields.test(CurrentEnergyColumn)) L.put("%8.5E\t", c->current.getEnergy() / energyScale);
	if (fields.test(CurrentPositionColumn)) {
		if (d3)
			L.vec(c->current.getPosition() / lengthScale);
		else
			L.put("%8.5E\t", c->current.getPosition().x / lengthScale);
	}
	if (fields.test(CurrentDirectionColumn) && d3) L.vec(c->current.getDirection());
	if (fields.test(SerialNumberColumn)) L.put("%10" PRIu64 "\t", c->getSourceSerialNumber());
	if (fields.test(SourceIdColumn)) L.put("%10i\t", c->source.getId());
	if (fields.test(SourceEnergyColumn)) L.put("%8.5E\t", c->source.getEnergy() / energyScale);
	if (fields.test(SourcePositionColumn)) {
		if (d3)
			L.vec(c->source.getPosition() / lengthScale);
		else
			L.put("%8.5E\t", c->source.getPosition().x / lengthScale);
	}
	if (fields.test(SourceDirectionColumn) && d3) L.vec(c->source.getDirection());
	if (fields.test(SerialNumberColumn)) L.put("%10" PRIu64 "\t", c->getCreatedSerialNumber());
	if (fields.test(CreatedIdColumn)) L.put("%10i\t", c->created.getId());
	if (fields.test(CreatedEnergyColumn)) L.put("%8.5E\t", c->created.getEnergy() / energyScale);
	if (fields.test(CreatedPositionColumn)) {
		if (d3)
			L.vec(c->created.getPosition() / lengthScale);
		else
			L.put("%8.5E\t", c->created.getPosition().x / lengthScale);
	}
	if (fields.test(CreatedDirectionColumn) && d3) L.vec(c->created.getDirection());
	if (fields.test(WeightColumn)) L.put("%8.5E\t", c->getWeight());
	if (fields.test(CandidateTagColumn)) L.put("%s\t", c->getTagOrigin().c_str());
	for (const Property &p : properties) {
		const std::string &v = c->hasProperty(p.name) ? c->getProperty(p.name) : p.defaultValue;
		L.put("%s\t", v.c_str());
	}
	L.buf[L.p - 1] = '\n';
	std::locale::global(old);
	std::lock_guard<std::mutex> lock(g_fileOutput);
	if (count == 0)
		printHeader();
	Output::process(c);
	out->write(L.buf, L.p);
}

// the columns of an Everything file, in order (src/module/TextOutput.cpp:294-365)
void TextOutput::load(const std::string &fn, ParticleCollector *collector) {
	std::ifstream in(fn.c_str());
	if (!in.good())
		throw std::runtime_error("crpropa::TextOutput: could not open file " + fn);
	if (ends_with(fn, ".gz"))
		throw std::runtime_error("CRPropa was built without Zlib compression!");
	Output defaults;
	const double ls = defaults.getLengthScale(), ts = defaults.getTimeScale(), es = defaults.getEnergyScale();
	std::string line;
	while (std::getline(in, line)) {
		std::stringstream s(line);
		if (s.peek() == '#')
			continue;
		ref_ptr<Candidate> c = new Candidate();
		double d, x, y, z;
		int i;
		s >> d;
		c->setTrajectoryLength(d * ls); // D
		s >> d;
		c->setTime(d * ts); // time
		s >> d;
		c->setRedshift(d); // z
		s >> i;
		c->setSerialNumber(i); // SN
		s >> i;
		c->current.setId(i); // ID
		s >> d;
		c->current.setEnergy(d * es); // E
		s >> x >> y >> z;
		c->current.setPosition(Vector3d(x, y, z) * ls);
		s >> x >> y >> z;
		c->current.setDirection(Vector3d(x, y, z) * ls);
		s >> i; // SN0
		s >> i;
		c->source.setId(i);
		s >> d;
		c->source.setEnergy(d * es);
		s >> x >> y >> z;
		c->source.setPosition(Vector3d(x, y, z) * ls);
		s >> x >> y >> z;
		c->source.setDirection(Vector3d(x, y, z) * ls);
		s >> i; // SN1
		s >> i;
		c->created.setId(i);
		s >> d;
		c->created.setEnergy(d * es);
		s >> x >> y >> z;
		c->created.setPosition(Vector3d(x, y, z) * ls);
		s >> x >> y >> z;
		c->created.setDirection(Vector3d(x, y, z) * ls);
		s >> d;
		c->setWeight(d); // W
		collector->process(c);
	}
}

// ------------------------------------------------------------------ ParticleCollector
ParticleCollector::ParticleCollector() : clone(false), recursive(false) { container.reserve(10000000); }
ParticleCollector::ParticleCollector(std::size_t nBuffer) : clone(false), recursive(false) {
	container.reserve(nBuffer);
}
ParticleCollector::ParticleCollector(std::size_t nBuffer, bool clone) : clone(clone), recursive(false) {
	container.reserve(nBuffer);
}
ParticleCollector::ParticleCollector(std::size_t nBuffer, bool clone, bool recursive)
    : clone(clone), recursive(recursive) {
	container.reserve(nBuffer);
}

void ParticleCollector::process(Candidate *c) const {
	static std::mutex m; // the reference's omp critical(ModifyContainer)
	std::lock_guard<std::mutex> lock(m);
	if (clone)
		container.push_back(c->clone(recursive));
	else
		container.push_back(c);
}

void ParticleCollector::reprocess(Module *action) const {
	for (const ref_ptr<Candidate> &c : container) {
		if (clone)
			action->process(c->clone(false).get());
		else
			action->process(c.get());
	}
}

void ParticleCollector::dump(const std::string &fn) const {
	TextOutput output(fn, Output::Everything);
	reprocess(&output);
	output.close();
}

void ParticleCollector::load(const std::string &fn) { TextOutput::load(fn, this); }

void ParticleCollector::getTrajectory(ModuleList *mlist, std::size_t i, Module *output) const {
	ref_ptr<Candidate> c = container[i]->clone();
	c->restart();
	mlist->add(output);
	mlist->run(c);
	mlist->remove(mlist->size() - 1);
}

} // namespace crpropa
