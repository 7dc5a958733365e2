// Grid3f storage with the reference's interface (include/crpropa/Grid.h:75-582):
// Nx*Ny*Nz Vector3f values, z fastest (index ix*Ny*Nz + iy*Nz + iz), the volume
// starting at `origin`, sample points at origin + (i + 1/2) * spacing,
// periodic (default) or reflective continuation, optional clipVolume.
// Interpolation runs on the GPU (trilinear, tricubic or nearest neighbour, the reference's arithmetic).
#ifndef CRPROPA_AMD_GRID_H
#define CRPROPA_AMD_GRID_H

#include "crpropa/Referenced.h"
#include "crpropa/Vector3.h"

#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

struct crp_field;

namespace crpropa {

enum interpolationType { TRILINEAR = 0, TRICUBIC, NEAREST_NEIGHBOUR };

class GridProperties : public Referenced {
public:
	size_t Nx, Ny, Nz;
	Vector3d origin;
	Vector3d spacing;
	bool reflective;
	interpolationType ipol;
	bool clipVolume;

	GridProperties(Vector3d origin, size_t N, double spacing)
	    : Nx(N), Ny(N), Nz(N), origin(origin), spacing(Vector3d(spacing)), reflective(false), ipol(TRILINEAR),
	      clipVolume(false) {}
	GridProperties(Vector3d origin, size_t Nx, size_t Ny, size_t Nz, double spacing)
	    : Nx(Nx), Ny(Ny), Nz(Nz), origin(origin), spacing(Vector3d(spacing)), reflective(false), ipol(TRILINEAR),
	      clipVolume(false) {}
	GridProperties(Vector3d origin, size_t Nx, size_t Ny, size_t Nz, Vector3d spacing)
	    : Nx(Nx), Ny(Ny), Nz(Nz), origin(origin), spacing(spacing), reflective(false), ipol(TRILINEAR),
	      clipVolume(false) {}
	void setReflective(bool b) { reflective = b; }
	void setInterpolationType(interpolationType i) { ipol = i; }
	void setClipVolume(bool b) { clipVolume = b; }
};

class Grid3f : public Referenced {
public:
	Grid3f(Vector3d origin, size_t N, double spacing);
	Grid3f(Vector3d origin, size_t Nx, size_t Ny, size_t Nz, double spacing);
	Grid3f(Vector3d origin, size_t Nx, size_t Ny, size_t Nz, Vector3d spacing);
	Grid3f(const GridProperties &p);
	~Grid3f();

	void setOrigin(Vector3d o) { origin = o; changed(); }
	void setGridSize(size_t nx, size_t ny, size_t nz);
	void setSpacing(Vector3d s) { spacing = s; changed(); }
	void setReflective(bool b) { reflective = b; changed(); }
	void setClipVolume(bool b) { clipVolume = b; changed(); }
	/* TRILINEAR, TRICUBIC or NEAREST_NEIGHBOUR (Grid.h:255-264); unknown values throw. */
	void setInterpolationType(interpolationType t);
	interpolationType getInterpolationType() const { return ipolType; }

	Vector3d getOrigin() const { return origin; }
	Vector3d getSpacing() const { return spacing; }
	size_t getNx() const { return Nx; }
	size_t getNy() const { return Ny; }
	size_t getNz() const { return Nz; }
	bool isReflective() const { return reflective; }
	bool getClipVolume() const { return clipVolume; }
	size_t getSizeOf() const { return sizeof(grid) + sizeof(Vector3f) * grid.size(); }

	/* Mutators mark the device copy stale; it is re-uploaded on next use. */
	Vector3f &get(size_t ix, size_t iy, size_t iz) { changed(); return grid[ix * Ny * Nz + iy * Nz + iz]; }
	const Vector3f &get(size_t ix, size_t iy, size_t iz) const { return grid[ix * Ny * Nz + iy * Nz + iz]; }
	Vector3f getValue(size_t ix, size_t iy, size_t iz) const { return grid[ix * Ny * Nz + iy * Nz + iz]; }
	void setValue(size_t ix, size_t iy, size_t iz, Vector3f v) { changed(); grid[ix * Ny * Nz + iy * Nz + iz] = v; }
	std::vector<Vector3f> &getGrid() { changed(); return grid; }
	const std::vector<Vector3f> &getGrid() const { return grid; }
	Vector3d positionFromIndex(int index) const;

	/* Interpolated value at a position (Grid::interpolate), evaluated on the GPU. */
	Vector3f interpolate(const Vector3d &position) const;

	/* Device field for this grid (created or refreshed on demand). */
	crp_field *deviceField() const;
	/* Marks host values as modified after writing through getGrid()/data(). */
	void changed() {
		dirty = true;
		++version;
	}
	/* increases with every change (device copies made from the grid compare it) */
	uint64_t getVersion() const { return version; }

private:
	std::vector<Vector3f> grid;
	size_t Nx, Ny, Nz;
	Vector3d origin, spacing;
	bool clipVolume, reflective;
	interpolationType ipolType;
	mutable crp_field *dev;
	mutable bool dirty;
	mutable std::mutex mtx;
	uint64_t version = 0;
};

/* Grid1f = Grid<float> (include/crpropa/Grid.h:587): one float per point, the same geometry and
 * continuation rules as Grid3f.  Used as the modulation of ModulatedMagneticFieldGrid, whose
 * evaluation runs on the GPU. */
class Grid1f : public Referenced {
public:
	Grid1f(Vector3d origin, size_t N, double spacing) : Grid1f(origin, N, N, N, Vector3d(spacing)) {}
	Grid1f(Vector3d origin, size_t nx, size_t ny, size_t nz, double spacing)
	    : Grid1f(origin, nx, ny, nz, Vector3d(spacing)) {}
	Grid1f(Vector3d origin, size_t nx, size_t ny, size_t nz, Vector3d spacing)
	    : grid(nx * ny * nz, 0.f), Nx(nx), Ny(ny), Nz(nz), origin(origin), spacing(spacing), clipVolume(false),
	      reflective(false), ipolType(TRILINEAR) {}
	Grid1f(const GridProperties &p)
	    : grid(p.Nx * p.Ny * p.Nz, 0.f), Nx(p.Nx), Ny(p.Ny), Nz(p.Nz), origin(p.origin), spacing(p.spacing),
	      clipVolume(p.clipVolume), reflective(p.reflective), ipolType(TRILINEAR) {
		setInterpolationType(p.ipol);
	}
	void setOrigin(Vector3d o) { origin = o; ++version; }
	void setSpacing(Vector3d s) { spacing = s; ++version; }
	void setReflective(bool b) { reflective = b; ++version; }
	void setClipVolume(bool b) { clipVolume = b; ++version; }
	/* TRILINEAR, TRICUBIC or NEAREST_NEIGHBOUR (Grid.h:255-264); unknown values throw. */
	void setInterpolationType(interpolationType t) {
		if (t != TRILINEAR && t != TRICUBIC && t != NEAREST_NEIGHBOUR)
			throw std::runtime_error("InterpolationType: unknown interpolation type");
		ipolType = t;
		++version;
	}
	interpolationType getInterpolationType() const { return ipolType; }
	Vector3d getOrigin() const { return origin; }
	Vector3d getSpacing() const { return spacing; }
	size_t getNx() const { return Nx; }
	size_t getNy() const { return Ny; }
	size_t getNz() const { return Nz; }
	bool isReflective() const { return reflective; }
	bool getClipVolume() const { return clipVolume; }
	float &get(size_t ix, size_t iy, size_t iz) { ++version; return grid[(ix * Ny + iy) * Nz + iz]; }
	const float &get(size_t ix, size_t iy, size_t iz) const { return grid[(ix * Ny + iy) * Nz + iz]; }
	float getValue(size_t ix, size_t iy, size_t iz) const { return grid[(ix * Ny + iy) * Nz + iz]; }
	void setValue(size_t ix, size_t iy, size_t iz, float v) { ++version; grid[(ix * Ny + iy) * Nz + iz] = v; }
	std::vector<float> &getGrid() { ++version; return grid; }
	const std::vector<float> &getGrid() const { return grid; }
	void changed() { ++version; }
	uint64_t getVersion() const { return version; }

private:
	std::vector<float> grid;
	size_t Nx, Ny, Nz;
	Vector3d origin, spacing;
	bool clipVolume, reflective;
	interpolationType ipolType;
	uint64_t version = 0;
};

} // namespace crpropa

#endif
