// Grid tools for Grid3f (see crpropa/GridTools.h).  The loops, their order
// and the arithmetic types follow src/GridTools.cpp so that sums, files and
// loaded values are bit-identical to the reference's.
#include "crpropa/GridTools.h"

#include <cmath>
#include <fstream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <vector>

namespace crpropa {

void scaleGrid(ref_ptr<Grid3f> grid, double a) {
	const float f = (float)a; // Vector3<float>::operator*=(const float &)
	for (Vector3f &b : grid->getGrid())
		b *= f;
}

Vector3f meanFieldVector(ref_ptr<Grid3f> grid) {
	Vector3f mean(0.f);
	for (const Vector3f &b : static_cast<const Grid3f &>(*grid).getGrid()) {
		mean.x += b.x;
		mean.y += b.y;
		mean.z += b.z;
	}
	return mean / (float)grid->getNx() / (float)grid->getNy() / (float)grid->getNz();
}

double meanFieldStrength(ref_ptr<Grid3f> grid) {
	double mean = 0;
	for (const Vector3f &b : static_cast<const Grid3f &>(*grid).getGrid())
		mean += b.getR(); // FP32 norm, FP64 sum
	return mean / (double)grid->getNx() / (double)grid->getNy() / (double)grid->getNz();
}

double rmsFieldStrength(ref_ptr<Grid3f> grid) {
	double sumV2 = 0;
	for (const Vector3f &b : static_cast<const Grid3f &>(*grid).getGrid())
		sumV2 += b.getR2(); // FP32 square norm, FP64 sum
	return std::sqrt(sumV2 / (double)grid->getNx() / (double)grid->getNy() / (double)grid->getNz());
}

std::array<float, 3> rmsFieldStrengthPerAxis(ref_ptr<Grid3f> grid) {
	// FP32 accumulators; each square is pow(float, 2) = a double product, added in double and
	// rounded back to float (C++11 promotes pow(float, int) to double)
	float sx = 0, sy = 0, sz = 0;
	for (const Vector3f &b : static_cast<const Grid3f &>(*grid).getGrid()) {
		sx = (float)((double)sx + (double)b.x * (double)b.x);
		sy = (float)((double)sy + (double)b.y * (double)b.y);
		sz = (float)((double)sz + (double)b.z * (double)b.z);
	}
	const float nx = (float)grid->getNx(), ny = (float)grid->getNy(), nz = (float)grid->getNz();
	return {std::sqrt(sx / nx / ny / nz), std::sqrt(sy / nx / ny / nz), std::sqrt(sz / nx / ny / nz)};
}

void fromMagneticField(ref_ptr<Grid3f> grid, ref_ptr<MagneticField> field) {
	const Vector3d origin = grid->getOrigin(), spacing = grid->getSpacing();
	const size_t Nx = grid->getNx(), Ny = grid->getNy(), Nz = grid->getNz();
	const size_t n = Nx * Ny * Nz;
	std::vector<double> pos(3 * n), B(3 * n);
	size_t i = 0;
	for (size_t ix = 0; ix < Nx; ix++)
		for (size_t iy = 0; iy < Ny; iy++)
			for (size_t iz = 0; iz < Nz; iz++, i++) {
				const Vector3d p = Vector3d(double(ix) + 0.5, double(iy) + 0.5, double(iz) + 0.5) * spacing + origin;
				pos[3 * i] = p.x;
				pos[3 * i + 1] = p.y;
				pos[3 * i + 2] = p.z;
			}
	if (field->deviceField()) {
		evaluateOnDevice(*field, pos.data(), n, B.data()); // one batch for every grid point
	} else {
		// a field without a device implementation (e.g. defined in Python) is the caller's host code
		for (i = 0; i < n; i++) {
			const Vector3d b = field->getField(Vector3d(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]));
			B[3 * i] = b.x;
			B[3 * i + 1] = b.y;
			B[3 * i + 2] = b.z;
		}
	}
	std::vector<Vector3f> &g = grid->getGrid();
	for (i = 0; i < n; i++)
		g[i] = Vector3f((float)B[3 * i], (float)B[3 * i + 1], (float)B[3 * i + 2]);
}

void loadGrid(ref_ptr<Grid3f> grid, std::string filename, double c) {
	std::ifstream fin(filename.c_str(), std::ios::binary);
	if (!fin)
		throw std::runtime_error("load Grid3f: " + filename + " not found");
	fin.seekg(0, fin.end);
	const size_t length = (size_t)fin.tellg() / sizeof(float);
	fin.seekg(0, fin.beg);
	std::vector<Vector3f> &g = grid->getGrid();
	if (length != 3 * g.size())
		throw std::runtime_error("loadGrid: file and grid size do not match");
	std::vector<float> raw(length);
	fin.read(reinterpret_cast<char *>(raw.data()), (std::streamsize)(length * sizeof(float)));
	const float f = (float)c;
	for (size_t i = 0; i < g.size(); i++) {
		g[i] = Vector3f(raw[3 * i], raw[3 * i + 1], raw[3 * i + 2]);
		g[i] *= f;
	}
}

void dumpGrid(ref_ptr<Grid3f> grid, std::string filename, double c) {
	std::ofstream fout(filename.c_str(), std::ios::binary);
	if (!fout)
		throw std::runtime_error("dump Grid3f: " + filename + " not found");
	const std::vector<Vector3f> &g = static_cast<const Grid3f &>(*grid).getGrid();
	std::vector<float> raw(3 * g.size());
	const float f = (float)c; // Vector3<float>::operator*(const float &)
	for (size_t i = 0; i < g.size(); i++) {
		raw[3 * i] = g[i].x * f;
		raw[3 * i + 1] = g[i].y * f;
		raw[3 * i + 2] = g[i].z * f;
	}
	fout.write(reinterpret_cast<const char *>(raw.data()), (std::streamsize)(raw.size() * sizeof(float)));
}

void loadGridFromTxt(ref_ptr<Grid3f> grid, std::string filename, double c) {
	std::ifstream fin(filename.c_str());
	if (!fin)
		throw std::runtime_error("load Grid3f: " + filename + " not found");
	while (fin.peek() == '#') // header lines
		fin.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
	const float f = (float)c;
	for (Vector3f &b : grid->getGrid()) {
		fin >> b.x >> b.y >> b.z;
		b *= f;
		if (fin.eof())
			throw std::runtime_error("load Grid3f: file too short");
	}
}

static const char *ipolName(interpolationType t) {
	return t == TRILINEAR ? "TRILINEAR" : (t == TRICUBIC ? "TRICUBIC" : (t == NEAREST_NEIGHBOUR ? "NEAREST_NEIGHBOUR" : "NOT_UNDERSTOOD"));
}

ref_ptr<Grid3f> loadGrid3fFromTxt(std::string filename, double c) {
	std::ifstream fin(filename.c_str());
	if (!fin)
		throw std::runtime_error("load Grid3f: " + filename + " not found");
	while (fin.peek() == '#') {
		std::string line;
		std::getline(fin, line);
		if (line.rfind("GridProperties:") != 2) // "# GridProperties: Type Grid3f\torigin: ..."
			continue;
		GridProperties gp(Vector3d(0.), 1, 1, 1, 1.);
		std::stringstream ss(line);
		std::string name, type;
		ss >> name >> name >> name >> type;
		if (type != "Grid3f")
			throw std::runtime_error("Tried to load Grid3f, but Gridproperties assume grid type " + type);
		double x, y, z;
		ss >> name >> x >> y >> z;
		gp.origin = Vector3d(x, y, z);
		ss >> name >> gp.Nx >> gp.Ny >> gp.Nz;
		ss >> name >> x >> y >> z;
		gp.spacing = Vector3d(x, y, z);
		ss >> name >> gp.reflective;
		bool clip;
		ss >> name >> clip;
		gp.setClipVolume(clip);
		ss >> name >> type;
		gp.setInterpolationType(type == "TRICUBIC" ? TRICUBIC : (type == "NEAREST_NEIGHBOUR" ? NEAREST_NEIGHBOUR : TRILINEAR));
		fin.close();
		ref_ptr<Grid3f> grid = new Grid3f(gp);
		loadGridFromTxt(grid, filename, c);
		return grid;
	}
	throw std::runtime_error("could not find GridProperties in file " + filename);
}

void dumpGridToTxt(ref_ptr<Grid3f> grid, std::string filename, double c, bool saveProp) {
	std::ofstream fout(filename.c_str());
	if (!fout)
		throw std::runtime_error("dump Grid3f: " + filename + " not found");
	if (saveProp) {
		fout << "# GridProperties: Type Grid3f"
		     << "\t" << "origin: " << grid->getOrigin() << "\t" << "gridsize: " << grid->getNx() << " " << grid->getNy()
		     << " " << grid->getNz() << "\t" << "spacing: " << grid->getSpacing() << "\t"
		     << "reflective: " << grid->isReflective() << "\t" << "clipVolume: " << grid->getClipVolume() << "\t"
		     << "interpolation: " << ipolName(grid->getInterpolationType()) << "\n";
	}
	const float f = (float)c;
	for (const Vector3f &v : static_cast<const Grid3f &>(*grid).getGrid())
		fout << (v * f) << "\n";
}

} // namespace crpropa
