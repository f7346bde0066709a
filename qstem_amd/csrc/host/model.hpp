// model.hpp -- atomic model input for the driver: extended .cfg reader,
// supercell replication with Einstein frozen-phonon displacements, and the
// .img writer (libs/fileio_fftw3.cpp:721-989, 1188-1657, 586-614;
// libs/imagelib_fftw3.cpp:57-110).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "qstem.h"

namespace qsh {

struct Cell {
  double Mm[3][3] = {{0}};  // rows = cell vectors (Angstrom)
  std::vector<double> fx, fy, fz, dw, occ, q;
  std::vector<int> Z;
};

bool read_cfg(const std::string& path, Cell* cell, std::string* err);
int z_number(const std::string& symbol);

// Cartesian supercell, z-sorted; the box origin is the minimum corner.  With
// tds, every atom is displaced by N(0, sigma^2) per axis, sigma =
// sqrt(T/300) sqrt(B/(8 pi^2))/sqrt(3) (Einstein branch of phononDisplacement).
// cube > 0: keep the atoms inside [0,cube) of a supercell large enough to cover it.
void build_model(const Cell& cell, int ncx, int ncy, int ncz, const double cube[3], bool tds, double temperature,
                 uint64_t seed, double xOffset, double yOffset, std::vector<qs_atom>* atoms, double ext[3]);

struct ImgHeader {
  double t = 0, dx = 1, dy = 1;
  std::vector<double> params;
  std::string comment;
};
bool write_img(const std::string& path, const float* data, int nx, int ny, bool complex_data, const ImgHeader& h);

}  // namespace qsh
