#include "model.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>

#include "qsc.hpp"

namespace qsh {

static const char* kElements[] = {
    "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg", "Al", "Si", "P",  "S",  "Cl", "Ar", "K",
    "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr", "Rb", "Sr",
    "Y",  "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn", "Sb", "Te", "I",  "Xe", "Cs", "Ba", "La",
    "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf", "Ta", "W",  "Re", "Os",
    "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn", "Fr", "Ra", "Ac", "Th", "Pa", "U"};

int z_number(const std::string& s0) {
  std::string s;
  for (char c : s0)
    if (!isspace((unsigned char)c)) s += c;
  if (s.size() > 1 && islower((unsigned char)s[1])) s = s.substr(0, 2);
  else s = s.substr(0, 1);
  for (int i = 0; i < (int)(sizeof(kElements) / sizeof(kElements[0])); i++)
    if (s == kElements[i]) return i + 1;
  return 0;
}

bool read_cfg(const std::string& path, Cell* cell, std::string* err) {
  ParamFile pf;
  if (!pf.open(path, '#')) {
    *err = "cannot open " + path;
    return false;
  }
  std::string buf;
  int ncoord = 0;
  double scale = 1.0;
  if (pf.read("Number of particles =", &buf)) first_int(buf, &ncoord);
  if (pf.read("A =", &buf)) first_double(buf, &scale);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      char t[32];
      snprintf(t, sizeof t, "H0(%d,%d) =", i + 1, j + 1);
      double v = 0;
      if (pf.read(t, &buf) && first_double(buf, &v)) cell->Mm[i][j] = v * scale;
    }
  const bool novel = pf.read(".NO_VELOCITY.", &buf);
  int entry = 3;
  if (pf.read("entry_count =", &buf)) first_int(buf, &entry);
  if (!novel) entry += 3;
  const int off = novel ? 0 : 3;
  if (ncoord < 1) {
    *err = "Number of atoms in CFG file not specified";
    return false;
  }
  // data follows the entry_count line (readNextCFGAtom, fileio_fftw3.cpp:908-989)
  std::ifstream f(path);
  std::string line;
  bool started = false;
  double mass = 28.0;
  int element = 1;
  while (std::getline(f, line) && (int)cell->Z.size() < ncoord) {
    if (!started) {
      if (line.find("entry_count") != std::string::npos) started = true;
      continue;
    }
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.resize(hash);
    std::istringstream is(line);
    std::vector<std::string> tok;
    std::string w;
    while (is >> w) tok.push_back(w);
    if (tok.empty()) continue;
    if (tok.size() == 1 && atof(tok[0].c_str()) >= 1.0 && tok[0].find_first_not_of("0123456789.") == std::string::npos) {
      mass = atof(tok[0].c_str());
      if (!std::getline(f, line)) break;
      element = z_number(line);
      continue;
    }
    if ((int)tok.size() < entry) {
      *err = "incomplete data line in " + path;
      return false;
    }
    std::vector<double> v(entry);
    for (int i = 0; i < entry; i++) v[i] = atof(tok[i].c_str());
    cell->fx.push_back(v[0]);
    cell->fy.push_back(v[1]);
    cell->fz.push_back(v[2]);
    cell->dw.push_back(entry > 3 + off ? v[3 + off] : 0.45 * 28.0 / mass);
    cell->occ.push_back(entry > 4 + off ? v[4 + off] : 1.0);
    cell->q.push_back(entry > 5 + off ? v[5 + off] : 0.0);
    cell->Z.push_back(element);
  }
  if ((int)cell->Z.size() != ncoord) {
    *err = "number of atoms does not agree with atoms in file";
    return false;
  }
  return true;
}

void build_model(const Cell& cell, int ncx, int ncy, int ncz, const double cube[3], bool tds, double temperature,
                 uint64_t seed, double xOffset, double yOffset, std::vector<qs_atom>* atoms, double ext[3]) {
  const double(&M)[3][3] = cell.Mm;
  const bool use_cube = cube && cube[0] > 0 && cube[1] > 0 && cube[2] > 0;
  if (use_cube) {  // enough cells to cover the cube (orthogonal cells)
    ncx = (int)ceil(cube[0] / std::max(1e-9, M[0][0]));
    ncy = (int)ceil(cube[1] / std::max(1e-9, M[1][1]));
    ncz = (int)ceil(cube[2] / std::max(1e-9, M[2][2]));
  }
  // inverse of M for the fractional displacement (MmInv)
  double inv[3][3];
  const double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                     M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
  inv[0][0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) / det;
  inv[0][1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) / det;
  inv[0][2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) / det;
  inv[1][0] = (M[1][2] * M[2][0] - M[1][0] * M[2][2]) / det;
  inv[1][1] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) / det;
  inv[1][2] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) / det;
  inv[2][0] = (M[1][0] * M[2][1] - M[1][1] * M[2][0]) / det;
  inv[2][1] = (M[0][1] * M[2][0] - M[0][0] * M[2][1]) / det;
  inv[2][2] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) / det;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int a = 0; a <= ncx; a += ncx)
    for (int b = 0; b <= ncy; b += ncy)
      for (int c = 0; c <= ncz; c += ncz)
        for (int d = 0; d < 3; d++) {
          const double v = a * M[0][d] + b * M[1][d] + c * M[2][d];
          lo[d] = std::min(lo[d], v);
          hi[d] = std::max(hi[d], v);
        }
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> gauss(0.0, 1.0);
  atoms->clear();
  const size_t n = cell.Z.size();
  for (int ix = 0; ix < ncx; ix++)
    for (int iy = 0; iy < ncy; iy++)
      for (int iz = 0; iz < ncz; iz++)
        for (size_t i = 0; i < n; i++) {
          double f[3] = {cell.fx[i] + ix, cell.fy[i] + iy, cell.fz[i] + iz};
          if (tds) {
            const double sig = sqrt(temperature / 300.0) * sqrt(cell.dw[i] / (8 * M_PI * M_PI)) / sqrt(3.0);
            const double u[3] = {sig * gauss(rng), sig * gauss(rng), sig * gauss(rng)};
            for (int d = 0; d < 3; d++) f[d] += u[0] * inv[0][d] + u[1] * inv[1][d] + u[2] * inv[2][d];
          }
          qs_atom a{};
          const double x = f[0] * M[0][0] + f[1] * M[1][0] + f[2] * M[2][0] - lo[0];
          const double y = f[0] * M[0][1] + f[1] * M[1][1] + f[2] * M[2][1] - lo[1];
          const double z = f[0] * M[0][2] + f[1] * M[1][2] + f[2] * M[2][2] - lo[2];
          if (use_cube && (x < 0 || y < 0 || z < 0 || x >= cube[0] || y >= cube[1] || z >= cube[2])) continue;
          a.x = (float)(x + xOffset);
          a.y = (float)(y + yOffset);
          a.z = (float)z;
          a.dw = (float)cell.dw[i];
          a.occ = (float)cell.occ[i];
          a.q = (float)cell.q[i];
          a.Znum = cell.Z[i];
          atoms->push_back(a);
        }
  std::stable_sort(atoms->begin(), atoms->end(), [](const qs_atom& a, const qs_atom& b) { return a.z < b.z; });
  for (int d = 0; d < 3; d++) ext[d] = use_cube ? cube[d] : hi[d] - lo[d];
}

bool write_img(const std::string& path, const float* data, int nx, int ny, bool complex_data, const ImgHeader& h) {
  FILE* fp = fopen(path.c_str(), "wb");
  if (!fp) return false;
  int32_t head[8] = {56, (int32_t)h.params.size(), (int32_t)h.comment.size(), nx, ny, complex_data ? 1 : 0,
                     complex_data ? 8 : 4, 1};
  double tdd[3] = {h.t, h.dx, h.dy};
  bool ok = fwrite(head, 4, 8, fp) == 8 && fwrite(tdd, 8, 3, fp) == 3;
  if (ok && !h.params.empty()) ok = fwrite(h.params.data(), 8, h.params.size(), fp) == h.params.size();
  if (ok && !h.comment.empty()) ok = fwrite(h.comment.data(), 1, h.comment.size(), fp) == h.comment.size();
  const size_t nvals = (size_t)nx * ny * (complex_data ? 2 : 1);
  if (ok) ok = fwrite(data, 4, nvals, fp) == nvals;
  return fclose(fp) == 0 && ok;
}

}  // namespace qsh
