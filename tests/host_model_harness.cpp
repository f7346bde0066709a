// Prints the atoms of qstem3's model builder (qstem_amd/csrc/host/model.cpp) for
// tests/test_host_model.py: `host_model_harness file.cfg ncx ncy ncz tds T seed [cx cy cz]`
// writes one line "x y z dw occ q Z" per atom, then "ext ax by cz".
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "model.hpp"

int main(int argc, char** argv) {
  if (argc < 8) return 2;
  qsh::Cell cell;
  std::string err;
  if (!qsh::read_cfg(argv[1], &cell, &err)) {
    fprintf(stderr, "%s\n", err.c_str());
    return 1;
  }
  double cube[3] = {0, 0, 0};
  if (argc >= 11)
    for (int d = 0; d < 3; d++) cube[d] = atof(argv[8 + d]);
  std::vector<qs_atom> atoms;
  double ext[3];
  qsh::build_model(cell, atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), cube, atoi(argv[5]) != 0, atof(argv[6]),
                   (uint64_t)atoll(argv[7]), 0.0, 0.0, &atoms, ext);
  for (const qs_atom& a : atoms)
    printf("%.9g %.9g %.9g %.9g %.9g %.9g %d\n", a.x, a.y, a.z, a.dw, a.occ, a.q, a.Znum);
  printf("ext %.17g %.17g %.17g\n", ext[0], ext[1], ext[2]);
  return 0;
}
