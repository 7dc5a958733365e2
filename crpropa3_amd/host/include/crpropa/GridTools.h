// Grid tools of the reference (include/crpropa/GridTools.h:27-170) for Grid3f:
// scaling, mean and rms field strength, sampling a field onto a grid, and the
// raw float32 / text grid file formats.  Same loop orders and arithmetic types
// as src/GridTools.cpp, so results and files are bit-identical.  Sampling a
// field (fromMagneticField) evaluates every grid point in one device batch
// when the field runs on the GPU.
#ifndef CRPROPA_AMD_GRIDTOOLS_H
#define CRPROPA_AMD_GRIDTOOLS_H

#include "crpropa/Grid.h"
#include "crpropa/magneticField/MagneticField.h"

#include <array>
#include <string>

namespace crpropa {

/* every value *= a (as float), src/GridTools.cpp:17-22 */
void scaleGrid(ref_ptr<Grid3f> grid, double a);
/* src/GridTools.cpp:24-33: FP32 sum in storage order, / Nx / Ny / Nz */
Vector3f meanFieldVector(ref_ptr<Grid3f> grid);
/* src/GridTools.cpp:35-44: FP64 sum of |B| (FP32 norm) */
double meanFieldStrength(ref_ptr<Grid3f> grid);
/* src/GridTools.cpp:57-66: sqrt(FP64 sum of |B|^2 (FP32) / N) */
double rmsFieldStrength(ref_ptr<Grid3f> grid);
/* src/GridTools.cpp:79-98 */
std::array<float, 3> rmsFieldStrengthPerAxis(ref_ptr<Grid3f> grid);
/* field sampled at origin + (i + 1/2) * spacing, src/GridTools.cpp:104-117 */
void fromMagneticField(ref_ptr<Grid3f> grid, ref_ptr<MagneticField> field);

/* raw float32 x,y,z per point, ix outermost, iz innermost; values * conversion
 * (src/GridTools.cpp:134-161, 192-209) */
void loadGrid(ref_ptr<Grid3f> grid, std::string filename, double conversion = 1);
void dumpGrid(ref_ptr<Grid3f> grid, std::string filename, double conversion = 1);
/* text, one "x y z" line per point, '#' header lines (src/GridTools.cpp:238-329, 423-447) */
void loadGridFromTxt(ref_ptr<Grid3f> grid, std::string filename, double conversion = 1);
ref_ptr<Grid3f> loadGrid3fFromTxt(std::string filename, double conversion = 1);
void dumpGridToTxt(ref_ptr<Grid3f> grid, std::string filename, double conversion = 1, bool saveProp = false);

} // namespace crpropa

#endif
