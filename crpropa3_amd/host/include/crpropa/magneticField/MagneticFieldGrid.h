// MagneticFieldGrid (src/magneticField/MagneticFieldGrid.cpp:5-19): a field
// defined by a Grid3f, interpolated on the GPU.
#ifndef CRPROPA_AMD_MAGNETICFIELDGRID_H
#define CRPROPA_AMD_MAGNETICFIELDGRID_H

#include "crpropa/Grid.h"
#include "crpropa/magneticField/MagneticField.h"

namespace crpropa {

class MagneticFieldGrid : public MagneticField {
public:
	MagneticFieldGrid(ref_ptr<Grid3f> grid) { setGrid(grid); }
	void setGrid(ref_ptr<Grid3f> g) { grid = g; }
	ref_ptr<Grid3f> getGrid() { return grid; }
	Vector3d getField(const Vector3d &position) const;
	crp_field *deviceField() const { return grid.valid() ? grid->deviceField() : nullptr; }

private:
	ref_ptr<Grid3f> grid;
};

} // namespace crpropa

#endif
