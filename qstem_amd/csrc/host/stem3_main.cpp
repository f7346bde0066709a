// stem3_main.cpp -- `qstem3 file.qsc`: drop-in for QSTEM's `stem3` in STEM mode.
//
// Reads the same .qsc keys with the same defaults (stem3/stem3.cpp:612-1559),
// the same .cfg atomic models, and writes the same .img STEM images
// (stem3/stemlib.cpp:2693-2749), but runs the multislice through the C ABI
// of libqstem_hip.so: potential, probe and all scan positions on the GPU,
// batched.  Differences from the reference, all deliberate:
//   * errors print a message and exit with status 1 (the reference exits 0);
//   * the frozen-phonon random stream is seeded ("seed:" key, default 1)
//     instead of the clock;
//   * only `mode: STEM`; `sequence: r p` with r = 1; neutral atoms.
#include <sys/stat.h>
#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "model.hpp"
#include "qsc.hpp"
#include "qstem.h"

using namespace qsh;

namespace {

const double kMilliSecPicoAmp = 1e-3 * (1e-12 / 1.6021773e-19);  // stemtypes_fftw3.h:47-49

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

}  // namespace

int main(int argc, char** argv) {
  const std::string qsc = argc > 1 ? argv[1] : "stem.qsc";
  ParamFile pf;
  if (!pf.open(qsc)) die("could not open parameter file " + qsc);
  std::string buf, w;
  // ---- readFile (stem3.cpp:612-1559), STEM subset --------------------------------
  if (pf.read("mode:", &buf) && buf.find("STEM") == std::string::npos) die("only mode: STEM is built");
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
  float scanXStart, scanXStop, scanYStart, scanYStop;
  int scanXN, scanYN;
  auto need = [&](const char* key, double* v) {
    if (!pf.read(key, &buf) || !first_double(buf, v)) die(std::string(key) + " is required");
  };
  need("scan_x_start:", &d); scanXStart = (float)d;
  need("scan_x_stop:", &d); scanXStop = (float)d;
  need("scan_x_pixels:", &d); scanXN = (int)d;
  need("scan_y_start:", &d); scanYStart = (float)d;
  need("scan_y_stop:", &d); scanYStop = (float)d;
  need("scan_y_pixels:", &d); scanYN = (int)d;
  if (scanXN < 1) scanXN = 1;
  if (scanYN < 1) scanYN = 1;
  double dE_E = 0, dI_I = 0, dV_V = 0, Cc = 0;
  if (pf.read("dE/E:", &buf)) dE_E = atof(buf.c_str());
  if (pf.read("dI/I:", &buf)) dI_I = atof(buf.c_str());
  if (pf.read("dV/V:", &buf)) dV_V = atof(buf.c_str());
  if (pf.read("Cc:", &buf)) Cc = 1e7 * atof(buf.c_str());
  // deterministic energy-spread table (stem3.cpp:1106-1126)
  const double dE0 = sqrt(dE_E * dE_E + dI_I * dI_I + dV_V * dV_V);
  std::vector<double> dEArr(avgRuns + 1, 0.0);
  {
    const double pi = 3.1415926535897;
    const double dy = 1.772453850905 / (2.0 * avgRuns), dx = pi / ((avgRuns + 1) * 20.0);
    double x = 0, y = 0;
    for (int ix = 1; ix < avgRuns; x += dx) {
      y += exp(-x * x) * dx;
      if (y >= ix * dy) {
        dEArr[ix++] = x * 2 * dE0 / pi;
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
  P.astigAngle = (float)(P.astigAngle * 3.1415926535897 / 180.0);
  struct Ab { const char* key; float* v; bool phi; } abs_[] = {
      {"a_33:", &P.a33, false}, {"a_31:", &P.a31, false}, {"a_44:", &P.a44, false}, {"a_42:", &P.a42, false},
      {"a_55:", &P.a55, false}, {"a_53:", &P.a53, false}, {"a_51:", &P.a51, false}, {"a_66:", &P.a66, false},
      {"a_64:", &P.a64, false}, {"a_62:", &P.a62, false}, {"phi_33:", &P.phi33, true}, {"phi_31:", &P.phi31, true},
      {"phi_44:", &P.phi44, true}, {"phi_42:", &P.phi42, true}, {"phi_55:", &P.phi55, true},
      {"phi_53:", &P.phi53, true}, {"phi_51:", &P.phi51, true}, {"phi_66:", &P.phi66, true},
      {"phi_64:", &P.phi64, true}, {"phi_62:", &P.phi62, true}};
  for (auto& a : abs_) {
    if (pf.read(a.key, &buf) && first_double(buf, &d)) *a.v = (float)d;
    if (a.phi) *a.v /= (float)(180.0 / 3.1415926535897);  // RAD2DEG
  }
  if (!pf.read("alpha:", &buf) || !first_double(buf, &d)) die("alpha: is required");
  P.alpha = (float)d;
  if (pf.read("AIS aperture:", &buf) && first_double(buf, &d)) P.aAIS = (float)d;
  double beamCurrent = 1, dwellTime = 1;
  if (pf.read("beam current:", &buf)) first_double(buf, &beamCurrent);
  if (pf.read("dwell time:", &buf)) first_double(buf, &dwellTime);
  P.electronScale = (float)((float)beamCurrent * (float)dwellTime * kMilliSecPicoAmp);
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
  std::vector<Det> dets;
  pf.reset();
  while (pf.read("detector:", &buf, false)) {
    Det dd;
    char name[64] = {0};
    float a = 0, b = 0, s1 = 0, s2 = 0;
    sscanf(buf.c_str(), "%g %g %63s %g %g", &a, &b, name, &s1, &s2);
    dd.rin = a; dd.rout = b; dd.name = name; dd.sx = s1; dd.sy = s2;
    dets.push_back(dd);
  }
  if ((int)dets.size() > QS_MAX_DET) die("too many detectors");
  float atomRadius = 5.0f;
  if (pf.read("atom radius:", &buf) && first_double(buf, &d)) atomRadius = (float)d;
  int quirks = QS_COMPAT_Q1 | QS_COMPAT_Q2;
  if (pf.read("reference quirks:", &buf) && !yes(buf)) quirks = 0;
  pf.reset();
  if (pf.read("sequence:", &buf, false)) {
    int r = 1, p = 1;
    sscanf(buf.c_str(), "%d %d", &r, &p);
    if (r != 1) die("sequence: repeats other than 1 are not built");
  }

  // potential window (stem3.cpp:1431-1447) and scan check (:1459-1470)
  const float potOffsetX = (float)(scanXStart - 0.5 * P.nx * resX);
  const float potOffsetY = (float)(scanYStart - 0.5 * P.ny * resY);
  int potNx = (int)((scanXStop - scanXStart) / resX);
  int potNy = (int)((scanYStop - scanYStart) / resY);
  scanXStop = scanXStart + resX * potNx;
  scanYStop = scanYStart + resY * potNy;
  potNx += P.nx;
  potNy += P.ny;
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

  // ---- doSTEM (stem3.cpp:2722-2961) on the GPU --------------------------------------
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
  const int nprobe = scanXN * scanYN;
  std::vector<int> ipos(2 * nprobe);
  for (int i = 0; i < nprobe; i++) {  // stem3.cpp:2825-2871
    const int ix = i / scanYN, iy = i % scanYN;
    int px = (int)(ix * (scanXStop - scanXStart) / ((float)scanXN * resX));
    int py = (int)(iy * (scanYStop - scanYStart) / ((float)scanYN * resY));
    if (px > potNx - P.nx) px = potNx - P.nx;
    if (py > potNy - P.ny) py = potNy - P.ny;
    ipos[2 * i] = px;
    ipos[2 * i + 1] = py;
  }
  const int total = slices * cellDiv;
  const int tCount = total / outputInterval;
  const int nbins = tCount + 1;
  std::vector<float> det((size_t)nbins * P.ndet * nprobe, 0.f);
  std::vector<float> diffAvg;
  std::vector<int> bins(slices);
  std::string fileBase = cfgname.substr(cfgname.find_last_of("/\\") == std::string::npos ? 0 : cfgname.find_last_of("/\\") + 1);
  if (fileBase.empty() || fileBase.back() != '_') {
    const size_t dot = fileBase.find('.');
    if (dot != std::string::npos) fileBase = fileBase.substr(0, dot) + "_";
    else fileBase += "_";
  }
  double t_all = 0;
  long long probe_slices = 0;
  for (int avgCount = 0; avgCount < avgRuns; avgCount++) {
    if (avgCount > 0) {
      build_model(cell, ncx, ncy, ncz, cube, tds, temperature, (uint64_t)seed + 7919ULL * avgCount, xOffset, yOffset,
                  &atoms, ext);
    }
    qs_params opt = P;
    opt.dE_E = (float)dEArr[avgCount];
    if (qs_set_optics(ctx, &opt) != QS_OK) die(qs_last_error(ctx));
    for (int pCount = 0; pCount < cellDiv; pCount++) {
      qs_pot_params pp{};
      pp.atomRadius = atomRadius;
      pp.potOffsetX = potOffsetX;
      pp.potOffsetY = potOffsetY;
      pp.czOffset = czOffset - pCount * sliceThickness * slices;
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
          write_img(folder + "/" + fileBase + std::to_string(iz) + ".img", cplx.data(), potNx, potNy, true, h);
        }
      }
      if (pCount == 0 && qs_make_probe(ctx, nullptr) != QS_OK) die(qs_last_error(ctx));
      const int nb = qs_record_bins(slices, cellDiv, outputInterval, pCount * slices, bins.data());
      const bool last_slab = pCount == cellDiv - 1;
      std::vector<float> diff, exitw;
      if (saveLevel > 0 && last_slab) diff.resize((size_t)nprobe * P.nx * P.ny);
      if (saveLevel > 1 && last_slab) exitw.resize((size_t)nprobe * P.nx * P.ny * 2);
      const auto t0 = std::chrono::steady_clock::now();
      if (qs_run_stem(ctx, ipos.data(), nprobe, pCount == 0 ? 0 : 1, bins.data(), slices, nb, det.data(),
                      diff.empty() ? nullptr : diff.data(), exitw.empty() ? nullptr : exitw.data(), nullptr) != QS_OK)
        die(qs_last_error(ctx));
      const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      t_all += sec;
      probe_slices += (long long)nprobe * slices;
      const float thick = total * sliceThickness;
      if (!diff.empty()) {  // diffAvg_ix_iy.img running average (stem3.cpp:2886-2930)
        const size_t n2 = (size_t)P.nx * P.ny;
        if (diffAvg.empty()) diffAvg.assign(diff.size(), 0.f);
        for (size_t k = 0; k < diff.size(); k++)
          diffAvg[k] = avgCount == 0 ? diff[k] : ((float)avgCount * diffAvg[k] + diff[k]) / (float)(avgCount + 1);
        for (int i = 0; i < nprobe; i++) {
          ImgHeader h;
          h.t = thick;
          h.dx = 1.0 / (P.nx * resX);
          h.dy = 1.0 / (P.ny * resY);
          h.comment = "Average Array";
          write_img(folder + "/diffAvg_" + std::to_string(i / scanYN) + "_" + std::to_string(i % scanYN) + ".img",
                    diffAvg.data() + i * n2, P.nx, P.ny, false, h);
        }
      }
      if (!exitw.empty()) {  // mulswav_ix_iy.img (stemlib.cpp:2526-2531)
        const size_t n2 = (size_t)P.nx * P.ny * 2;
        for (int i = 0; i < nprobe; i++) {
          ImgHeader h;
          h.t = thick;
          h.dx = resX;
          h.dy = resY;
          h.comment = "Wavefunction";
          write_img(folder + "/mulswav_" + std::to_string(i / scanYN) + "_" + std::to_string(i % scanYN) + ".img",
                    exitw.data() + i * n2, P.nx, P.ny, true, h);
        }
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
          if (!write_img(name, img, scanXN, scanYN, false, h)) die("could not write " + name);
        }
      }
    }
    if (printLevel > 0)
      printf("run %d/%d: %d positions x %d slices, %.3f s GPU multislice\n", avgCount + 1, avgRuns, nprobe,
             slices * cellDiv, t_all);
  }
  if (printLevel > 0)
    printf("qstem3: %lld probe-slices in %.3f s (%.0f probe-slices/s); images in %s/\n", probe_slices, t_all,
           probe_slices / std::max(t_all, 1e-9), folder.c_str());
  qs_destroy(ctx);
  return 0;
}
