// JF12Field and PlanckJF12bField with the reference's interface
// (include/crpropa/magneticField/JF12Field.h, src/magneticField/JF12Field.cpp): the Galactic field
// of Jansson & Farrar 2012, regular part with optional striated (Grid1f) and turbulent (Grid3f)
// parts.  getField runs on the GPU; PropagationBP/CK and DiffusionSDE evaluate it on the device.
#ifndef CRPROPA_AMD_JF12FIELD_H
#define CRPROPA_AMD_JF12FIELD_H

#include "crpropa/Grid.h"
#include "crpropa/magneticField/MagneticField.h"

#include "crpropa3_amd.h"

#include <mutex>

namespace crpropa {

class JF12Field : public MagneticField {
public:
	JF12Field();
	~JF12Field();

	/* striated grid of 100^3 points at 0.1 kpc with values +-1 from Random(seed) (seed 0: random) */
	void randomStriated(int seed = 0);
	/* turbulent grid of 256^3 points at 4 pc, lMin 8 pc, lMax 272 pc, Brms 1 (generated on the GPU) */
	void randomTurbulent(int seed = 0);
	void setStriatedGrid(ref_ptr<Grid1f> grid);
	void setTurbulentGrid(ref_ptr<Grid3f> grid);
	ref_ptr<Grid1f> getStriatedGrid() { return striatedGrid; }
	ref_ptr<Grid3f> getTurbulentGrid() { return turbulentGrid; }

	void setUseRegularField(bool use);
	void setUseDiskField(bool use);
	void setUseToroidalHaloField(bool use);
	void setUseXField(bool use);
	void setUseStriatedField(bool use);
	void setUseTurbulentField(bool use);
	bool isUsingRegularField() { return p.use_regular != 0; }
	bool isUsingDiskField() { return p.use_disk != 0; }
	bool isUsingToroidalHaloField() { return p.use_toroidal_halo != 0; }
	bool isUsingXField() { return p.use_x != 0; }
	bool isUsingStriatedField() { return p.use_striated != 0; }
	bool isUsingTurbulentField() { return p.use_turbulent != 0; }

	/* the parts (evaluated on the GPU with the other switches as set) */
	Vector3d getRegularField(const Vector3d &pos) const;
	Vector3d getStriatedField(const Vector3d &pos) const;
	Vector3d getTurbulentField(const Vector3d &pos) const;
	Vector3d getField(const Vector3d &pos) const;
	crp_field *deviceField() const;
	const crp_jf12_params &getParameters() const { return p; }

protected:
	explicit JF12Field(int planck);
	crp_jf12_params p;
	ref_ptr<Grid1f> striatedGrid;
	ref_ptr<Grid3f> turbulentGrid;

private:
	Vector3d evalWith(const Vector3d &pos, int regular, int striated, int turbulent) const;
	crp_field *deviceFieldLocked() const; // under mtx
	void releaseDevice() const;           // under mtx: the part views, then the field whose grids they read
	// the part getters (regular, striated, turbulent) evaluate views of the device field with their
	// switches (crp_field_create_jf12_view): no second copy of a grid
	mutable crp_field *partDev[3] = {nullptr, nullptr, nullptr};
	mutable crp_jf12_params partParams[3];
	mutable crp_field *dev = nullptr;
	mutable crp_jf12_params devParams;
	mutable const Grid1f *devStr = nullptr;
	mutable const Grid3f *devTurb = nullptr;
	mutable uint64_t devStrVersion = 0, devTurbVersion = 0;
	mutable std::mutex mtx;
};

/* JF12 with the Planck Collaboration's corrections (A&A 596 (2016) A103; JF12Field.cpp:348-363) */
class PlanckJF12bField : public JF12Field {
public:
	PlanckJF12bField() : JF12Field(1) {}
};

} // namespace crpropa

#endif
