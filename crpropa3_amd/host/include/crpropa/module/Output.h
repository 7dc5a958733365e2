// Output modules with the reference's interface (include/crpropa/module/Output.h:12-170,
// TextOutput.h, ParticleCollector.h; src/module/Output.cpp, TextOutput.cpp,
// ParticleCollector.cpp).  They run on the host: as a detection action of a fused Observer
// (called once per detection with the candidate's state at detection) or as a module of a
// step-wise ModuleList.  Column layout and number formats are the reference's, so files are
// interchangeable.  Compressed (.gz) files and HDF5 are not built in.
#ifndef CRPROPA_AMD_OUTPUT_H
#define CRPROPA_AMD_OUTPUT_H

#include "crpropa/Module.h"
#include "crpropa/ModuleList.h"

#include <bitset>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

namespace crpropa {

/* version line written into text headers */
extern const std::string g_GIT_DESC;

class Output : public Module {
public:
	struct Property {
		std::string name;
		std::string comment;
		std::string defaultValue; // Variant::toString of the reference's default
	};
	enum OutputColumn {
		TrajectoryLengthColumn,
		ColumnDensityColumn,
		RedshiftColumn,
		CurrentIdColumn,
		CurrentEnergyColumn,
		CurrentPositionColumn,
		CurrentDirectionColumn,
		SourceIdColumn,
		SourceEnergyColumn,
		SourcePositionColumn,
		SourceDirectionColumn,
		CreatedIdColumn,
		CreatedEnergyColumn,
		CreatedPositionColumn,
		CreatedDirectionColumn,
		CandidateTagColumn,
		SerialNumberColumn,
		WeightColumn,
		TimeColumn
	};
	enum OutputType { Trajectory1D, Trajectory3D, Event1D, Event3D, Everything };

	Output();
	Output(OutputType outputType);
	static std::string OutputTypeName(OutputType outputType);
	const std::string outputName;

	void setEnergyScale(double scale);
	double getEnergyScale() const { return energyScale; }
	void setLengthScale(double scale);
	double getLengthScale() const { return lengthScale; }
	void setTimeScale(double scale);
	double getTimeScale() const { return timeScale; }
	void setOutputType(OutputType outputType);
	void set(OutputColumn field, bool value);
	void enableProperty(const std::string &property, const std::string &defaultValue,
	                    const std::string &comment = "");
	void enable(OutputColumn field);
	void disable(OutputColumn field);
	void enableAll();
	void disableAll();
	void set1D(bool value);
	size_t size() const { return count; }
	void process(Candidate *) const;

protected:
	/* configuration is fixed once data has been written (src/module/Output.cpp:28-31) */
	void modify();
	double lengthScale, timeScale, energyScale;
	std::bitset<64> fields;
	std::vector<Property> properties;
	bool oneDimensional;
	mutable size_t count;
};

class ParticleCollector;

class TextOutput : public Output {
public:
	TextOutput();                               // standard output
	TextOutput(OutputType outputType);
	TextOutput(std::ostream &out);
	TextOutput(std::ostream &out, OutputType outputType);
	TextOutput(const std::string &filename);
	TextOutput(const std::string &filename, OutputType outputType);
	~TextOutput();
	void enableRandomSeeds() { storeRandomSeeds = true; }
	void process(Candidate *c) const;
	/* reads a file written with all columns (Everything) back into candidates */
	static void load(const std::string &filename, ParticleCollector *collector);
	std::string getDescription() const { return "TextOutput"; }
	void close();
	void gzip();

protected:
	void printHeader() const;
	std::ofstream outfile;
	std::ostream *out;
	std::string filename;
	bool storeRandomSeeds;
};

class ParticleCollector : public Module {
public:
	typedef std::vector<ref_ptr<Candidate>> tContainer;
	typedef tContainer::iterator iterator;
	typedef tContainer::const_iterator const_iterator;

	ParticleCollector();
	ParticleCollector(std::size_t nBuffer);
	ParticleCollector(std::size_t nBuffer, bool clone);
	ParticleCollector(std::size_t nBuffer, bool clone, bool recursive);
	void process(Candidate *candidate) const;
	void process(ref_ptr<Candidate> c) const { process(c.get()); }
	void reprocess(Module *action) const;
	void dump(const std::string &filename) const;
	void load(const std::string &filename);
	std::size_t size() const { return container.size(); }
	ref_ptr<Candidate> operator[](std::size_t i) const { return container[i]; }
	void clearContainer() { container.clear(); }
	std::string getDescription() const { return "ParticleCollector"; }
	std::vector<ref_ptr<Candidate>> &getContainer() const { return container; }
	void setClone(bool b) { clone = b; }
	bool getClone() const { return clone; }
	iterator begin() { return container.begin(); }
	const_iterator begin() const { return container.begin(); }
	iterator end() { return container.end(); }
	const_iterator end() const { return container.end(); }
	/* re-runs the module list from the candidate's source state with `output` appended */
	void getTrajectory(ModuleList *mlist, std::size_t i, Module *output) const;

protected:
	mutable tContainer container;
	bool clone, recursive;
};

} // namespace crpropa

#endif
