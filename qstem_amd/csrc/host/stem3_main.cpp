// stem3_main.cpp -- `qstem3 file.qsc`: drop-in for QSTEM's `stem3` in STEM, TEM and
// CBED mode.
//
// Reads the same .qsc keys with the same defaults (stem3/stem3.cpp:612-1559),
// the same .cfg atomic models, and writes the same .img files, but runs the
// multislice through the C ABI of libqstem_hip.so: potential, probe and all
// scan positions on the GPU, batched.
//   STEM (doSTEM, stem3.cpp:2722-2961): detector images per output thickness
//        (stemlib.cpp:2693-2749), diffAvg_ix_iy.img and mulswav_ix_iy.img.
//   TEM  (doTEM, stem3.cpp:2354-2716): plane or tilted incident wave;
//        wave_t.img at each output thickness (interimWave, stemlib.cpp:2538-2567)
//        and the exit wave wave.img.
//   CBED (doCBED, stem3.cpp:2048-2347): probe at scan_x_start, scan_y_start;
//        diff_t.img at each output thickness (collectIntensity,
//        stemlib.cpp:2661-2675) and diffAvg_n.img, avgresults.dat.
// "sequence: r p" lines run every slab r times (runMulsSTEM's mRepeat loop) and
// p * cellDiv slabs per line, with the reference's bin and file indices.
// Differences from the reference, all deliberate:
//   * errors print a message and exit with status 1 (the reference exits 0);
//   * the frozen-phonon random stream and the CBED source-size offsets are seeded
//     ("seed:" key, default 1) instead of drawn from the clock;
//   * without a "sequence:" line the run is "sequence: 1 1" (the reference
//     propagates nothing);
//   * doTEM aborts after the slab loop reading a diff.img that nothing writes
//     (stem3.cpp:2564); qstem3 goes on to the files its code writes next:
//     diffAvg_n.img (running average of the exit diffraction pattern),
//     waveIntensity.img (run 0), image.img (later runs), avgresults.dat;
//   * doCBED's later runs re-read the diff.img of run 1 (stem3.cpp:2243-2253);
//     qstem3 averages each run's own pattern and writes no diff.img;
//   * a slab continued from the previous one keeps its wave on the device instead
//     of re-reading mulswav_ix_iy.img (which the reference only writes with
//     save level > 1 or cellDiv > 1).
#include <sys/stat.h>
#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <random>
#include <string>
#include <vector>

#include "model.hpp"
#include "qsc.hpp"
#include "qstem.h"

using namespace qsh;

namespace {

const double kMilliSecPicoAmp = 1e-3 * (1e-12 / 1.6021773e-19);  // stemtypes_fftw3.h:47-49
const double kPi = 3.1415926535897;

[[noreturn]] void die(const std::string& m) {
  fprintf(stderr, "qstem3: %s\n", m.c_str());
  exit(1);
}

double wavelength(double kev) {
  const double emass = 510.99906, hc = 12.3984244;
  return hc / sqrt(kev * (2.0 * emass + kev));
}

struct Det {
  float rin = 0, rout = 0, sx = 0, sy = 0;
  std::string name;
};

enum class Mode { STEM, TEM, CBED };

std::string dirname_of(const std::string& p) {
  const size_t k = p.find_last_of('/');
  return k == std::string::npos ? std::string(".") : p.substr(0, k);
}

std::string find_data_file(const char* argv0, const char* name) {
  if (const char* e = getenv("QSTEM_DATA")) return std::string(e) + "/" + name;
  char buf[4096];
  ssize_t n = readlink("/proc/self/exe", buf, sizeof buf - 1);
  std::string exe = n > 0 ? std::string(buf, n) : std::string(argv0);
  return dirname_of(dirname_of(exe)) + "/data/" + name;
}

std::vector<double> read_table(const char* argv0, const char* name) {
  std::vector<double> table(93 * 123);
  const std::string tp = find_data_file(argv0, name);
  FILE* f = fopen(tp.c_str(), "rb");
  if (!f || fread(table.data(), 8, table.size(), f) != table.size()) die("cannot read scattering table " + tp);
  fclose(f);
  return table;
}

// "Beam tilt X:" style value: mrad as written (used as radians by doTEM, which
// applies sin() to it), or degrees when the unit word starts with 'd' (stem3.cpp:699-717)
float tilt_value(const std::string& buf) {
  float v = 0;
  char unit[64] = {0};
  sscanf(buf.c_str(), "%g %63s", &v, unit);
  if (tolower((unsigned char)unit[0]) == 'd') v = (float)(v * kPi / 180.0);
  return v;
}

void write_or_die(const std::string& name, const float* data, int nx, int ny, bool cplx, const ImgHeader& h) {
  if (!write_img(name, data, nx, ny, cplx, h)) die("could not write " + name);
}

// fftshifted |psi|^2 / sqrt(nx ny) of a real-space wave: collectIntensity applied to
// the wave after the inverse transform (stemlib.cpp:2625-2633 as doTEM and doCBED with
// save level > 1 call it, :2476-2482)
std::vector<float> shifted_intensity(const std::vector<float>& w, int nx, int ny) {
  std::vector<float> d((size_t)nx * ny);
  const double scale = 1.0 / sqrt((double)nx * ny);
  for (int ix = 0; ix < nx; ix++)
    for (int iy = 0; iy < ny; iy++) {
      const float re = w[2 * ((size_t)ix * ny + iy)], im = w[2 * ((size_t)ix * ny + iy) + 1];
      const double in = re * re + im * im;
      d[(size_t)((ix + nx / 2) % nx) * ny + (iy + ny / 2) % ny] = (float)(in * scale);
    }
  return d;
}

}  // namespace

int main(int argc, char** argv) {
  const std::string qsc = argc > 1 ? argv[1] : "stem.qsc";
  ParamFile pf;
  if (!pf.open(qsc)) die("could not open parameter file " + qsc);
  std::string buf, w;
  // ---- readFile (stem3.cpp:612-1559) ------------------------------------------------
  Mode mode = Mode::STEM;
  if (pf.read("mode:", &buf)) {  // the reference's test order: "STEM" contains "TEM"
    if (buf.find("STEM") != std::string::npos) mode = Mode::STEM;
    else if (buf.find("TEM") != std::string::npos) mode = Mode::TEM;
    else if (buf.find("CBED") != std::string::npos) mode = Mode::CBED;
    else die("mode:" + buf + " is not built (STEM, TEM and CBED are)");
  }
  int printLevel = 2, saveLevel = 0;
  if (pf.read("print level:", &buf)) first_int(buf, &printLevel);
  if (pf.read("save level:", &buf)) first_int(buf, &saveLevel);
  if (!pf.read("filename:", &buf) || !first_word(buf, &w)) die("readFile did not find crystal .cfg file for parsing");
  std::string cfgname = w;
  if (!cfgname.empty() && cfgname[0] == '"') {
    const size_t a = buf.find('"'), b = buf.find('"', a + 1);
    cfgname = buf.substr(a + 1, b - a - 1);
  }
  int ncx = 1, ncy = 1, ncz = 1, cellDiv = 1;
  if (pf.read("NCELLX:", &buf)) first_int(buf, &ncx);
  if (pf.read("NCELLY:", &buf)) first_int(buf, &ncy);
  if (pf.read("NCELLZ:", &buf) && first_word(buf, &w)) {
    const size_t sl = w.find('/');
    if (sl != std::string::npos) cellDiv = atoi(w.c_str() + sl + 1);
    ncz = atoi(w.c_str());
  }
  if (cellDiv < 1) cellDiv = 1;
  double cube[3] = {0, 0, 0};
  if (pf.read("Cube:", &buf)) sscanf(buf.c_str(), "%lf %lf %lf", &cube[0], &cube[1], &cube[2]);
  bool tds = pf.read("tds:", &buf) && yes(buf);
  double temperature = 300.0;
  if (pf.read("temperature:", &buf)) first_double(buf, &temperature);
  double xOffset = 0, yOffset = 0;
  if (pf.read("xOffset:", &buf)) first_double(buf, &xOffset);
  if (pf.read("yOffset:", &buf)) first_double(buf, &yOffset);
  std::string cfgpath = cfgname;
  if (access(cfgpath.c_str(), R_OK) != 0) cfgpath = dirname_of(qsc) + "/" + cfgname;
  Cell cell;
  std::string err;
  if (!read_cfg(cfgpath, &cell, &err)) die(err);
  int seed = 1;
  if (pf.read("seed:", &buf)) first_int(buf, &seed);
  std::vector<qs_atom> atoms;
  double ext[3];
  build_model(cell, ncx, ncy, ncz, cube, tds, temperature, (uint64_t)seed, xOffset, yOffset, &atoms, ext);
  if (atoms.empty()) die("No atom within simulation boundaries");
  const float ax = (float)ext[0], by = (float)ext[1], cz = (float)ext[2];

  qs_params P{};
  if (!pf.read("nx:", &buf) || !first_int(buf, &P.nx)) die("nx: is required");
  P.ny = P.nx;
  if (pf.read("ny:", &buf)) first_int(buf, &P.ny);
  double d = 0;
  float resX = 0, resY = 0;
  if (pf.read("resolutionX:", &buf) && first_double(buf, &d)) resX = (float)d;
  if (pf.read("resolutionY:", &buf) && first_double(buf, &d)) resY = (float)d;
  if (!pf.read("v0:", &buf) || !first_double(buf, &d)) die("v0: is required");
  P.v0 = (float)d;
  const int centerSlices = pf.read("center slices:", &buf) && yes(buf);
  float sliceThickness = 0;
  int slices = 0;
  const float cdepth = cube[2] > 0 ? (float)cube[2] : cz;
  if (pf.read("slice-thickness:", &buf) && first_double(buf, &d)) {
    sliceThickness = (float)d;
    if (pf.read("slices:", &buf)) first_int(buf, &slices);
    else slices = (int)(cdepth / (cellDiv * sliceThickness) + 0.99);
    slices += centerSlices;
  } else if (pf.read("slices:", &buf) && first_int(buf, &slices) && slices > 0) {
    sliceThickness = cdepth / (cellDiv * slices);
  }
  if (slices <= 0) die("Number of slices = 0");
  // recompute the potential for every slab unless all cellDiv slabs are equal (stem3.cpp:930-931)
  const bool equalDivs = !tds && ncz % cellDiv == 0 && fabs(slices * sliceThickness - cdepth / cellDiv) < 1e-5;
  int outputInterval = slices;
  if (pf.read("slices between outputs:", &buf)) first_int(buf, &outputInterval);
  if (outputInterval < 1) outputInterval = slices;
  float czOffset = 0;
  if (pf.read("zOffset:", &buf) && first_double(buf, &d)) czOffset = (float)d;
  if (resX == 0) resX = ax / (float)P.nx;
  if (resY == 0) resY = by / (float)P.ny;
  const int periodicXY = pf.read("periodicXY:", &buf) && yes(buf);
  P.bandlimittrans = 1;
  if (pf.read("bandlimit f_trans:", &buf)) P.bandlimittrans = yes(buf);
  const bool savePotential = pf.read("save potential:", &buf) && yes(buf);
  if (pf.read("one time integration:", &buf) && !yes(buf)) die("only 'one time integration: yes' is built");
  if (pf.read("potential3D:", &buf) && !yes(buf)) die("only 'potential3D: yes' is built");
  int avgRuns = 10;
  if (pf.read("Runs for averaging:", &buf)) first_int(buf, &avgRuns);
  if (!tds) avgRuns = 1;
  if (avgRuns < 1) avgRuns = 1;
  // scan window (stem3.cpp:1025-1075): STEM needs all six keys; CBED reads the
  // probe position; TEM keeps the model centre
  float scanXStart = ax / 2.0f, scanYStart = by / 2.0f, scanXStop, scanYStop;
  int scanXN = 1, scanYN = 1;
  auto need = [&](const char* key, double* v) {
    if (!pf.read(key, &buf) || !first_double(buf, v)) die(std::string(key) + " is required");
  };
  if (mode == Mode::STEM) {
    need("scan_x_start:", &d); scanXStart = (float)d;
    need("scan_x_stop:", &d); scanXStop = (float)d;
    need("scan_x_pixels:", &d); scanXN = (int)d;
    need("scan_y_start:", &d); scanYStart = (float)d;
    need("scan_y_stop:", &d); scanYStop = (float)d;
    need("scan_y_pixels:", &d); scanYN = (int)d;
    if (scanXN < 1) scanXN = 1;
    if (scanYN < 1) scanYN = 1;
  } else {
    if (mode == Mode::CBED) {
      if (pf.read("scan_x_start:", &buf) && first_double(buf, &d)) scanXStart = (float)d;
      if (pf.read("scan_y_start:", &buf) && first_double(buf, &d)) scanYStart = (float)d;
    }
    scanXStop = scanXStart;
    scanYStop = scanYStart;
  }
  double dE_E = 0, dI_I = 0, dV_V = 0, Cc = 0;
  if (pf.read("dE/E:", &buf)) dE_E = atof(buf.c_str());
  if (pf.read("dI/I:", &buf)) dI_I = atof(buf.c_str());
  if (pf.read("dV/V:", &buf)) dV_V = atof(buf.c_str());
  if (pf.read("Cc:", &buf)) Cc = 1e7 * atof(buf.c_str());
  // deterministic energy-spread table (stem3.cpp:1106-1126); doSTEM steps through it,
  // doTEM and doCBED keep the dE/E read here
  const double dE0 = sqrt(dE_E * dE_E + dI_I * dI_I + dV_V * dV_V);
  std::vector<double> dEArr(avgRuns + 1, 0.0);
  {
    const double dy = 1.772453850905 / (2.0 * avgRuns), dx = kPi / ((avgRuns + 1) * 20.0);
    double x = 0, y = 0;
    for (int ix = 1; ix < avgRuns; x += dx) {
      y += exp(-x * x) * dx;
      if (y >= ix * dy) {
        dEArr[ix++] = x * 2 * dE0 / kPi;
        if (ix < avgRuns) { dEArr[ix] = -dEArr[ix - 1]; ix++; }
      }
    }
  }
  if (!pf.read("Cs:", &buf) || !first_double(buf, &d)) die("Cs: is required");
  P.Cs = (float)((float)d * 1.0e7);
  if (pf.read("C5:", &buf) && first_double(buf, &d)) P.C5 = (float)((float)d * 1.0e7);
  const float wl = (float)wavelength(P.v0);
  P.df0 = -(float)sqrt(1.5 * P.Cs * wl);
  if (pf.read("defocus:", &buf) && first_word(buf, &w)) {
    const char c0 = (char)tolower((unsigned char)w[0]);
    if (c0 == 's') P.df0 = -(float)sqrt(1.5 * P.Cs * wavelength(P.v0));
    else if (c0 == 'o') P.df0 = -(float)sqrt(P.Cs * wavelength(P.v0));
    else if (first_double(buf, &d)) P.df0 = (float)(10.0 * (float)d);
  }
  if (pf.read("astigmatism:", &buf) && first_double(buf, &d)) P.astigMag = (float)d;
  P.astigMag *= 10.0f;
  if (pf.read("astigmatism angle:", &buf) && first_double(buf, &d)) P.astigAngle = (float)d;
  P.astigAngle = (float)(P.astigAngle * kPi / 180.0);
  struct Ab { const char* key; float* v; bool phi; } abs_[] = {
      {"a_33:", &P.a33, false}, {"a_31:", &P.a31, false}, {"a_44:", &P.a44, false}, {"a_42:", &P.a42, false},
      {"a_55:", &P.a55, false}, {"a_53:", &P.a53, false}, {"a_51:", &P.a51, false}, {"a_66:", &P.a66, false},
      {"a_64:", &P.a64, false}, {"a_62:", &P.a62, false}, {"phi_33:", &P.phi33, true}, {"phi_31:", &P.phi31, true},
      {"phi_44:", &P.phi44, true}, {"phi_42:", &P.phi42, true}, {"phi_55:", &P.phi55, true},
      {"phi_53:", &P.phi53, true}, {"phi_51:", &P.phi51, true}, {"phi_66:", &P.phi66, true},
      {"phi_64:", &P.phi64, true}, {"phi_62:", &P.phi62, true}};
  for (auto& a : abs_) {
    if (pf.read(a.key, &buf) && first_double(buf, &d)) *a.v = (float)d;
    if (a.phi) *a.v /= (float)(180.0 / kPi);  // RAD2DEG
  }
  if (!pf.read("alpha:", &buf) || !first_double(buf, &d)) die("alpha: is required");
  P.alpha = (float)d;
  if (pf.read("AIS aperture:", &buf) && first_double(buf, &d)) P.aAIS = (float)d;
  double beamCurrent = 1, dwellTime = 1;
  if (pf.read("beam current:", &buf)) first_double(buf, &beamCurrent);
  if (pf.read("dwell time:", &buf)) first_double(buf, &dwellTime);
  P.electronScale = (float)((float)beamCurrent * (float)dwellTime * kMilliSecPicoAmp);
  float sourceRadius = 0;  // "Source Size (diameter):" / 2 (stem3.cpp:1258-1260)
  if (pf.read("Source Size (diameter):", &buf)) sourceRadius = (float)(atof(buf.c_str()) / 2.0);
  P.ismoth = pf.read("smooth:", &buf) && yes(buf);  // Q8: absent key -> off here
  P.gaussScale = 0.05f;
  if (pf.read("gaussian:", &buf)) {
    char ans[64] = {0};
    float gs = 0.05f;
    if (sscanf(buf.c_str(), "%63s %g", ans, &gs) >= 1) P.gaussFlag = tolower((unsigned char)ans[0]) == 'y';
    P.gaussScale = gs;
  }
  std::string folder = "data";
  if (pf.read("Folder:", &buf) && first_word(buf, &w)) {
    folder = w;
    if (!folder.empty() && folder[0] == '"') {
      const size_t a = buf.find('"'), b = buf.find('"', a + 1);
      folder = buf.substr(a + 1, b - a - 1);
    }
    if (!folder.empty() && folder.back() == '/') folder.pop_back();
  }
  // beam tilt (stem3.cpp:699-717), used by doTEM's incident wave
  float btiltx = 0, btilty = 0;
  int tiltBack = 1;
  if (pf.read("Beam tilt X:", &buf)) btiltx = tilt_value(buf);
  if (pf.read("Beam tilt Y:", &buf)) btilty = tilt_value(buf);
  if (pf.read("Tilt back:", &buf) && first_word(buf, &w)) tiltBack = tolower((unsigned char)w[0]) == 'y';
  if (mode == Mode::TEM && pf.read("Pendelloesung plot:", &buf) && yes(buf))
    die("'Pendelloesung plot: yes' is not built");
  std::vector<Det> dets;
  if (mode == Mode::STEM) {
    pf.reset();
    while (pf.read("detector:", &buf, false)) {
      Det dd;
      char name[64] = {0};
      float a = 0, b = 0, s1 = 0, s2 = 0;
      sscanf(buf.c_str(), "%g %g %63s %g %g", &a, &b, name, &s1, &s2);
      dd.rin = a; dd.rout = b; dd.name = name; dd.sx = s1; dd.sy = s2;
      dets.push_back(dd);
    }
  }
  if ((int)dets.size() > QS_MAX_DET) die("too many detectors");
  float atomRadius = 5.0f;
  if (pf.read("atom radius:", &buf) && first_double(buf, &d)) atomRadius = (float)d;
  int quirks = QS_COMPAT_Q1 | QS_COMPAT_Q2;
  if (pf.read("reference quirks:", &buf) && !yes(buf)) quirks = 0;
  // "sequence: r p" lines, read in file order without wrapping; a line that does not
  // start with a letter or 1-9 ends the list (doSTEM, stem3.cpp:2764-2793)
  struct Seq { int r, p; };
  std::vector<Seq> seqs;
  pf.reset();
  while (pf.read("sequence: ", &buf, false)) {
    const unsigned char c0 = buf.empty() ? 0 : (unsigned char)buf[0];
    if (!((c0 >= 'a' && c0 <= 'z') || (c0 >= '1' && c0 <= '9') || (c0 >= 'A' && c0 <= 'Z'))) {
      printf("Can only work with old stacking sequence\n");
      break;
    }
    int r = 1, p = 0;
    sscanf(buf.c_str(), "%d %d", &r, &p);
    if (r < 1) die("sequence: the repeat count must be at least 1");
    if (p < 1) p = 1;
    if (mode == Mode::STEM && p > 1 && cube[0] > 0 && cube[1] > 0 && cube[2] > 0) {
      printf("Warning: cube size of height %gA has been defined, ignoring sequence\n", cube[2]);
      p = 1;
    }
    seqs.push_back({r, p});
  }
  if (seqs.empty()) seqs.push_back({1, 1});

  // potential window (stem3.cpp:1431-1456) and scan check (:1459-1470)
  float potOffsetX, potOffsetY;
  int potNx, potNy;
  if (mode != Mode::TEM) {
    potOffsetX = (float)(scanXStart - 0.5 * P.nx * resX);
    potOffsetY = (float)(scanYStart - 0.5 * P.ny * resY);
    potNx = (int)((scanXStop - scanXStart) / resX);
    potNy = (int)((scanYStop - scanYStart) / resY);
    scanXStop = scanXStart + resX * potNx;
    scanYStop = scanYStart + resY * potNy;
    potNx += P.nx;
    potNy += P.ny;
  } else {
    potNx = P.nx;
    potNy = P.ny;
    potOffsetX = (float)(scanXStart - 0.5 * (double)(potNx * resX));
    potOffsetY = (float)(scanYStart - 0.5 * (double)(potNy * resY));
  }
  if (scanXStart < 0 || scanYStart < 0 || scanXStop < 0 || scanYStop < 0 || scanXStart > ax || scanYStart > by ||
      scanXStop > ax || scanYStop > by)
    die("Scanning window is outside model dimensions");
  P.resolutionX = resX;
  P.resolutionY = resY;
  P.slices = slices;
  P.sliceThickness = sliceThickness;
  P.potNx = potNx;
  P.potNy = potNy;
  P.potSizeX = potNx * resX;
  P.potSizeY = potNy * resY;
  P.by = by;
  P.Cc = (float)Cc;
  P.ndet = (int)dets.size();
  for (int i = 0; i < P.ndet; i++) {
    P.det_rin[i] = dets[i].rin; P.det_rout[i] = dets[i].rout;
    P.det_shiftx[i] = dets[i].sx; P.det_shifty[i] = dets[i].sy;
  }
  P.compat = quirks;

  // ---- doSTEM / doTEM / doCBED on the GPU --------------------------------------------
  qs_ctx* ctx = nullptr;
  if (qs_create(&P, &ctx) != QS_OK) die(std::string("qs_create: ") + qs_last_error(nullptr));
  {
    // Rez et al. neutral and ion-offset tables (stem3/scatfactsRez.h): the reference's
    // default build adds ion offsets (USE_Q_POT_OFFSETS 1, stemlib.cpp:77)
    const std::vector<double> table = read_table(argv[0], "rez_scatpar_f64.bin");
    if (qs_set_scattering_table(ctx, table.data()) != QS_OK) die(qs_last_error(ctx));
    const std::vector<double> ions = read_table(argv[0], "rez_scatparoffs_f64.bin");
    if (qs_set_ion_scattering_table(ctx, ions.data()) != QS_OK) die(qs_last_error(ctx));
  }
  mkdir(folder.c_str(), 0755);
  const int nx = P.nx, ny = P.ny;
  const size_t n2 = (size_t)nx * ny;
  const int nprobe = scanXN * scanYN;
  std::vector<int> ipos(2 * nprobe);
  for (int i = 0; i < nprobe; i++) {  // stem3.cpp:2825-2871; TEM and CBED: one window at (0, 0)
    const int ix = i / scanYN, iy = i % scanYN;
    int px = (int)(ix * (scanXStop - scanXStart) / ((float)scanXN * resX));
    int py = (int)(iy * (scanYStop - scanYStart) / ((float)scanYN * resY));
    if (px > potNx - nx) px = potNx - nx;
    if (py > potNy - ny) py = potNy - ny;
    ipos[2 * i] = px;
    ipos[2 * i + 1] = py;
  }
  const int total = slices * cellDiv;
  const int tCount = total / outputInterval;
  const int nbins = tCount + 1;
  const float slabC = sliceThickness * slices;  // make3DSlices' slab thickness c (stemlib.cpp:573)
  std::vector<float> det((size_t)nbins * P.ndet * nprobe, 0.f);
  std::vector<float> diffAvg;
  std::string fileBase = cfgname.substr(cfgname.find_last_of("/\\") == std::string::npos ? 0 : cfgname.find_last_of("/\\") + 1);
  if (fileBase.empty() || fileBase.back() != '_') {
    const size_t dot = fileBase.find('.');
    if (dot != std::string::npos) fileBase = fileBase.substr(0, dot) + "_";
    else fileBase += "_";
  }
  // projected potential of sub-cell j of the supercell (make3DSlices: atom z - c j + zOffset)
  auto build_slab = [&](int j) {
    qs_pot_params pp{};
    pp.atomRadius = atomRadius;
    pp.potOffsetX = potOffsetX;
    pp.potOffsetY = potOffsetY;
    pp.czOffset = czOffset - j * slabC;
    pp.centerSlices = centerSlices;
    pp.periodicXY = periodicXY;
    pp.tds = tds;
    pp.compat_q6 = 1;
    pp.compat_q5 = 1;  // the reference's ion-offset LUT (stemlib.cpp:1672)
    std::vector<float> V;
    if (savePotential) V.resize((size_t)slices * potNx * potNy);
    if (qs_build_potential(ctx, atoms.data(), (int)atoms.size(), &pp, savePotential ? V.data() : nullptr) != QS_OK)
      die(qs_last_error(ctx));
    if (savePotential) {  // stemlib.cpp:1276-1306: complex slices, imaginary part zero
      std::vector<float> cplx((size_t)potNx * potNy * 2, 0.f);
      for (int iz = 0; iz < slices; iz++) {
        for (size_t k = 0; k < (size_t)potNx * potNy; k++) cplx[2 * k] = V[(size_t)iz * potNx * potNy + k];
        ImgHeader h;
        h.t = sliceThickness;
        h.dx = resX;
        h.dy = resY;
        h.comment = "Projected Potential (slice " + std::to_string(iz) + ")";
        write_or_die(folder + "/" + fileBase + std::to_string(iz) + ".img", cplx.data(), potNx, potNy, true, h);
      }
    }
  };
  // header fields of the wave writers (WAVEFUNC::WriteWave/WriteDiffPat/WriteAvgArray,
  // libs/data_containers.cpp:65-92)
  auto rspace = [&](double t, const char* comment) {
    ImgHeader h;
    h.t = t;
    h.dx = resX;
    h.dy = resY;
    h.comment = comment;
    return h;
  };
  auto kspace = [&](double t, const char* comment) {
    ImgHeader h;
    h.t = t;
    h.dx = 1.0 / (nx * resX);
    h.dy = 1.0 / (ny * resY);
    h.comment = comment;
    return h;
  };
  // interimWave's parameters (stemlib.cpp:2548-2558)
  const std::vector<double> waveParams = {P.v0, P.Cs, P.df0, P.astigMag, P.astigAngle,
                                          Cc * sqrt(dE_E * dE_E + dV_V * dV_V + dI_I * dI_I), P.alpha, btiltx, btilty};
  // bin index of collectIntensity for a slice index (stemlib.cpp:2597-2607)
  auto bin_of = [&](long long slice) { return slice < total - 1 ? (int)(slice / outputInterval) : tCount; };

  std::mt19937_64 srcrng((uint64_t)seed * 2654435761ULL + 17);
  std::normal_distribution<double> gauss(0.0, 1.0);
  float waveThickness = 0;  // WAVEFUNC::thickness: kept from one run to the next
  std::vector<double> chisq(avgRuns, 0.0);
  std::vector<float> imageAvg;                  // doTEM: image intensity average
  std::map<int, std::vector<float>> cbedDiff;  // doCBED: contents of diff_t.img
  const int zero_ipos[2] = {0, 0};
  double t_all = 0;
  long long probe_slices = 0;
  auto run_slices = [&](const int* ip, int np, int start, int first, const int* bins, int n, float* dt, float* df,
                        float* ex) {
    const auto t0 = std::chrono::steady_clock::now();
    if (qs_run_slices(ctx, ip, nullptr, np, start, first, bins, n, nbins, dt, df, ex, nullptr) != QS_OK)
      die(qs_last_error(ctx));
    t_all += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    probe_slices += (long long)np * n;
  };

  for (int avgCount = 0; avgCount < avgRuns; avgCount++) {
    qs_params opt = P;
    opt.dE_E = (float)(mode == Mode::STEM ? dEArr[avgCount] : dE_E);
    if (qs_set_optics(ctx, &opt) != QS_OK) die(qs_last_error(ctx));
    int totalSliceCount = 0, cycle = 0;
    bool started = false;  // TEM, CBED: the wave continues from slab to slab and line to line
    // last collectIntensity pattern (wave->diffpat), last k-space pattern, last exit wave
    std::vector<float> diffpat(n2, 0.f), lastK(n2, 0.f), exitw;
    if (mode == Mode::CBED) {  // probe at the scan position plus the source-size offset (stem3.cpp:2093-2121)
      const double offX = sourceRadius * gauss(srcrng) * sqrt(2.0), offY = sourceRadius * gauss(srcrng) * sqrt(2.0);
      const float sx = (float)(scanXStart + offX), sy = (float)(scanYStart + offY);
      std::vector<float> pr(2 * n2);
      if (qs_make_probe_at(ctx, (double)(sx - potOffsetX), (double)(sy - potOffsetY), pr.data()) != QS_OK)
        die(qs_last_error(ctx));
      if (saveLevel > 2)
        write_or_die(folder + "/wave_probe.img", pr.data(), nx, ny, true, rspace(waveThickness, "Wavefunction"));
      if (sourceRadius > 0) {
        FILE* fp = fopen("probepos.dat", avgCount == 0 ? "w" : "a");
        if (fp) {
          fprintf(fp, "%g %g\n", sx, sy);
          fclose(fp);
        }
      }
    }
    if (mode == Mode::TEM) {  // plane wave, or tilted (stem3.cpp:2403-2420; y uses ix as there)
      std::vector<float> pw(2 * n2);
      const double lam = wavelength(P.v0);
      const double ktx = 2.0 * kPi * sin(btiltx) / lam, kty = 2.0 * kPi * sin(btilty) / lam;
      for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++) {
          const size_t k = (size_t)ix * ny + iy;
          if (btiltx == 0 && btilty == 0) {
            pw[2 * k] = 1.f;
            pw[2 * k + 1] = 0.f;
          } else {
            const double x = resX * (float)(ix - nx / 2), y = resY * (float)(ix - nx / 2);
            pw[2 * k] = (float)cos(ktx * x + kty * y);
            pw[2 * k + 1] = (float)sin(ktx * x + kty * y);
          }
        }
      if (qs_set_incident_wave(ctx, pw.data()) != QS_OK) die(qs_last_error(ctx));
    }
    for (size_t li = 0; li < seqs.size(); li++) {
      const int reps = seqs[li].r, npc = seqs[li].p * cellDiv;
      for (int pCount = 0; pCount < npc; pCount++) {
        const int j = equalDivs ? 0 : pCount % cellDiv;
        if (!equalDivs || pCount == 0) {
          // later runs re-read (re-displace) the atoms at the start of each pass over
          // the cellDiv sub-cells (stemlib.cpp:458-484)
          if (avgCount > 0 && j == 0) {
            build_model(cell, ncx, ncy, ncz, cube, tds, temperature,
                        (uint64_t)seed + 7919ULL * avgCount + 104729ULL * cycle, xOffset, yOffset, &atoms, ext);
            cycle++;
          }
          build_slab(j);
        }
        const bool lastSlab = pCount == npc - 1;
        if (mode == Mode::STEM) {
          // each sequence line starts from the probe (pCount == 0, stem3.cpp:2834-2840)
          if (pCount == 0 && qs_make_probe(ctx, nullptr) != QS_OK) die(qs_last_error(ctx));
          std::vector<int> steps((size_t)slices * reps);
          if (qs_record_bins_repeat(slices, cellDiv, outputInterval, totalSliceCount, reps, steps.data()) != nbins)
            die("bin bookkeeping failed");
          std::vector<float> diff, exw;
          if (saveLevel > 0 && lastSlab) diff.resize(n2 * nprobe);
          if ((saveLevel > 1 || cellDiv > 1) && lastSlab) exw.resize(2 * n2 * nprobe);
          for (int rep = 0; rep < reps; rep++) {
            const bool last = rep == reps - 1;
            run_slices(ipos.data(), nprobe, (pCount == 0 && rep == 0) ? 0 : 1, 0, steps.data() + (size_t)rep * slices,
                       slices, det.data(), last && !diff.empty() ? diff.data() : nullptr,
                       last && !exw.empty() ? exw.data() : nullptr);
          }
          const float thick = (totalSliceCount + slices) * sliceThickness;
          if (!diff.empty()) {  // diffAvg_ix_iy.img running average (stem3.cpp:2886-2930)
            if (diffAvg.empty()) diffAvg.assign(diff.size(), 0.f);
            for (size_t k = 0; k < diff.size(); k++)
              diffAvg[k] = avgCount == 0 ? diff[k] : ((float)avgCount * diffAvg[k] + diff[k]) / (float)(avgCount + 1);
            for (int i = 0; i < nprobe; i++)
              write_or_die(folder + "/diffAvg_" + std::to_string(i / scanYN) + "_" + std::to_string(i % scanYN) + ".img",
                           diffAvg.data() + i * n2, nx, ny, false, kspace(thick, "Average Array"));
          }
          if (!exw.empty()) {  // mulswav_ix_iy.img (stemlib.cpp:2526-2531)
            for (int i = 0; i < nprobe; i++)
              write_or_die(folder + "/mulswav_" + std::to_string(i / scanYN) + "_" + std::to_string(i % scanYN) + ".img",
                           exw.data() + i * 2 * n2, nx, ny, true, rspace(thick, "Wavefunction"));
          }
          // saveSTEMImages (stemlib.cpp:2693-2749): images keep the last run only (Navg stays 0, Q3)
          for (int t = 0; t <= tCount; t++) {
            const float thickness = t < tCount ? ((t + 1) * outputInterval) * sliceThickness : total * sliceThickness;
            for (int i = 0; i < P.ndet; i++) {
              const float* img = det.data() + ((size_t)t * P.ndet + i) * nprobe;
              float error = 0;
              double inten = 0;
              std::vector<double> params(2 + nprobe);
              for (int k = 0; k < nprobe; k++) {
                const float im2 = img[k] * img[k];  // image2 = error^2 of the pixel's raw sum
                error += (im2 - img[k] * img[k]);
                inten += img[k] * img[k];
                params[2 + k] = im2;
              }
              error /= inten;
              params[0] = avgCount + 1;
              params[1] = error;
              ImgHeader h;
              h.t = thickness;
              h.dx = (scanXStop - scanXStart) / (float)scanXN;
              h.dy = (scanYStop - scanYStart) / (float)scanYN;
              h.params = params;
              h.comment = "STEM image";
              const std::string name = t < tCount ? folder + "/" + dets[i].name + "_" + std::to_string(t) + ".img"
                                                  : folder + "/" + dets[i].name + ".img";
              write_or_die(name, img, scanXN, scanYN, false, h);
            }
          }
        } else {
          // TEM, CBED: one wave through the slab's steps (rep, islice), run in blocks
          // that end where a file needs the wave or its diffraction pattern
          for (int rep = 0; rep < reps; rep++) {
            int islice = 0;
            while (islice < slices) {
              auto interim_at = [&](int s) {  // interimWave's test (stemlib.cpp:2544)
                const long long sl = (long long)(totalSliceCount + s) * (1 + rep);
                return !(sl < total - 1 && (sl + 1) % outputInterval != 0);
              };
              auto needs = [&](int s) {
                const bool endSlab = rep == reps - 1 && s == slices - 1;
                if (mode == Mode::CBED) return saveLevel > 0 || (endSlab && (saveLevel > 2 || lastSlab));
                return interim_at(s) || (endSlab && lastSlab);
              };
              int e = islice;
              while (e < slices - 1 && !needs(e)) e++;
              std::vector<float> dk(n2);
              exitw.assign(2 * n2, 0.f);
              run_slices(zero_ipos, 1, started ? 1 : 0, islice, nullptr, e - islice + 1, nullptr, dk.data(),
                         exitw.data());
              started = true;
              const int absSlice = totalSliceCount + e;
              const float prevThickness = e > islice ? (float)(totalSliceCount + e) * sliceThickness : waveThickness;
              waveThickness = (absSlice + 1) * sliceThickness;
              diffpat = dk;  // collectIntensity after propagation (stemlib.cpp:2429)
              lastK = dk;
              if (mode == Mode::CBED) {
                // diff_t.img of collectIntensity: written in run 0, averaged into in later
                // runs at every step that reaches it (stemlib.cpp:2661-2675)
                auto cbed_update = [&](int t, const std::vector<float>& pat, float th) {
                  std::vector<float>& f = cbedDiff[t];
                  const std::string name = folder + "/diff_" + std::to_string(t) + ".img";
                  if (avgCount == 0 || f.empty()) {
                    f = pat;
                    write_or_die(name, f.data(), nx, ny, false, kspace(th, "Diffraction Pattern"));
                  } else {
                    for (size_t k = 0; k < n2; k++) f[k] = (avgCount * f[k] + pat[k]) / (avgCount + 1);
                    write_or_die(name, f.data(), nx, ny, false, kspace(th, "Average Array"));
                  }
                };
                if (saveLevel > 0) cbed_update(bin_of(totalSliceCount + (long long)e * (1 + rep)), dk, prevThickness);
                if (saveLevel > 1) {  // interimWave, then collectIntensity of the real-space wave
                  const long long sl = (long long)absSlice * (1 + rep);
                  if (interim_at(e)) {
                    const int t = (int)(sl / outputInterval);
                    const std::string name = tds ? folder + "/wave_" + std::to_string(avgCount) + "_" + std::to_string(t) + ".img"
                                                 : folder + "/wave_" + std::to_string(t) + ".img";
                    ImgHeader h = rspace(waveThickness, "Wave Function");
                    h.params = waveParams;
                    write_or_die(name, exitw.data(), nx, ny, true, h);
                  }
                  diffpat = shifted_intensity(exitw, nx, ny);
                  cbed_update(bin_of(sl), diffpat, waveThickness);
                }
                if (avgCount == 0 && saveLevel > 2 && rep == reps - 1 && e == slices - 1)
                  write_or_die(folder + "/wave_final.img", exitw.data(), nx, ny, true,
                               rspace(waveThickness, "Wavefunction"));
              } else {
                if (interim_at(e)) {  // wave_t.img (interimWave, stemlib.cpp:2538-2567)
                  const int t = (int)((long long)absSlice * (1 + rep) / outputInterval);
                  const std::string name = tds ? folder + "/wave_" + std::to_string(avgCount) + "_" + std::to_string(t) + ".img"
                                               : folder + "/wave_" + std::to_string(t) + ".img";
                  ImgHeader h = rspace(waveThickness, "Wave Function");
                  h.params = waveParams;
                  write_or_die(name, exitw.data(), nx, ny, true, h);
                }
                diffpat = shifted_intensity(exitw, nx, ny);
                if (avgCount == 0 && rep == reps - 1 && e == slices - 1 && lastSlab) {
                  // wave.img at the end of the line, after the tilt compensation that
                  // multiplies re by cos and im by sin (stem3.cpp:2494-2512)
                  if (tiltBack && (btiltx != 0 || btilty != 0)) {
                    const double lam = wavelength(P.v0);
                    const double ktx = -2.0 * kPi * sin(btiltx) / lam, kty = -2.0 * kPi * sin(btilty) / lam;
                    for (int ix = 0; ix < nx; ix++)
                      for (int iy = 0; iy < ny; iy++) {
                        const double x = resX * (float)(ix - nx / 2), y = resY * (float)(ix - nx / 2);
                        const size_t k = (size_t)ix * ny + iy;
                        exitw[2 * k] = (float)(exitw[2 * k] * cos(ktx * x + kty * y));
                        exitw[2 * k + 1] = (float)(exitw[2 * k + 1] * sin(ktx * x + kty * y));
                      }
                    // the compensated wave goes on into the next sequence line
                    if (qs_set_incident_wave(ctx, exitw.data()) != QS_OK) die(qs_last_error(ctx));
                    started = false;
                  }
                  write_or_die(folder + "/wave.img", exitw.data(), nx, ny, true,
                               rspace(waveThickness, tds ? "Test wave function for run 0"
                                                         : "Exit face wave function for no TDS"));
                }
              }
              islice = e + 1;
            }
          }
        }
        totalSliceCount += slices;
      }
    }
    if (mode == Mode::CBED) {  // diffAvg_n.img and avgresults.dat (stem3.cpp:2241-2301)
      if (avgCount == 0) {
        diffAvg = diffpat;
        write_or_die(folder + "/diffAvg_1.img", diffAvg.data(), nx, ny, false,
                     kspace(waveThickness, "DEBUG Wave intensity"));
      } else {
        double cs = 0;
        for (size_t k = 0; k < n2; k++) {
          const float t = ((float)avgCount * diffAvg[k] + diffpat[k]) / (float)(avgCount + 1);
          cs += (diffAvg[k] - t) * (diffAvg[k] - t);
          diffAvg[k] = t;
        }
        chisq[avgCount - 1] = cs / (double)n2;
        ImgHeader h = kspace(waveThickness, "Averaged Diffraction pattern, unit: 1/A");
        h.params = {0.0, 1.0 / wavelength(P.v0)};
        write_or_die(folder + "/diffAvg_" + std::to_string(avgCount + 1) + ".img", diffAvg.data(), nx, ny, false, h);
        const bool storeSeries = saveLevel != 0 && avgCount % saveLevel == 0;
        if (!storeSeries) remove((folder + "/diffAvg_" + std::to_string(avgCount) + ".img").c_str());
        FILE* fp = fopen((folder + "/avgresults.dat").c_str(), "w");
        if (fp) {
          for (int i = 0; i < avgCount; i++) fprintf(fp, "%d %g\n", i + 1, chisq[i]);
          fclose(fp);
        }
      }
    } else if (mode == Mode::TEM) {
      // what doTEM's code writes after the slab loop (stem3.cpp:2564-2684): the
      // diffraction pattern average, and the image |FFT^-1 FFT psi|^2 = (nx ny)^2 |psi|^2
      // of the exit wave (no transfer function is applied)
      std::vector<float> inten(n2);
      const double nn = (double)nx * ny;
      for (size_t k = 0; k < n2; k++)
        inten[k] = (float)(nn * nn * ((double)exitw[2 * k] * exitw[2 * k] + (double)exitw[2 * k + 1] * exitw[2 * k + 1]));
      // the diffraction pattern is the k-space one of the last step (collectIntensity's
      // second call leaves the real-space intensity in wave->diffpat)
      if (avgCount == 0) {
        diffAvg = lastK;
        imageAvg = inten;
        write_or_die(folder + "/diffAvg_1.img", diffAvg.data(), nx, ny, false, kspace(waveThickness, "Diffraction pattern"));
        write_or_die(folder + "/waveIntensity.img", inten.data(), nx, ny, false, kspace(waveThickness, "Wave intensity"));
      } else {
        double cs = 0;
        for (size_t k = 0; k < n2; k++) {
          const float t = ((float)avgCount * diffAvg[k] + lastK[k]) / (float)(avgCount + 1);
          cs += (diffAvg[k] - t) * (diffAvg[k] - t);
          diffAvg[k] = t;
          imageAvg[k] = ((float)avgCount * imageAvg[k] + inten[k]) / (float)(avgCount + 1);
        }
        chisq[avgCount - 1] = cs / (double)n2;
        write_or_die(folder + "/diffAvg_" + std::to_string(avgCount + 1) + ".img", diffAvg.data(), nx, ny, false,
                     kspace(waveThickness, "Diffraction pattern"));
        write_or_die(folder + "/image.img", imageAvg.data(), nx, ny, false, kspace(waveThickness, "Image intensity"));
        FILE* fp = fopen("avgresults.dat", "w");
        if (fp) {
          for (int i = 0; i < avgCount; i++) fprintf(fp, "%d %g\n", i + 1, chisq[i]);
          fclose(fp);
        }
      }
    }
    if (printLevel > 0)
      printf("run %d/%d: %d position(s) x %d slices, %.3f s GPU multislice\n", avgCount + 1, avgRuns, nprobe,
             totalSliceCount, t_all);
  }
  if (printLevel > 0)
    printf("qstem3: %lld probe-slices in %.3f s (%.0f probe-slices/s); output in %s/\n", probe_slices, t_all,
           probe_slices / std::max(t_all, 1e-9), folder.c_str());
  qs_destroy(ctx);
  return 0;
}
