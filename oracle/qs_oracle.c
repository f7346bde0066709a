/*
 * qs_oracle.c -- FP64 CPU restatement of QSTEM's STEM multislice path.
 *
 * TEST INFRASTRUCTURE ONLY (see qs_oracle.h).  Loaded by tests/, the smoke
 * check and bench.py's cpu_baseline leg; never by the product library.
 *
 * Pinning (what checks each function independently of its own formula):
 *   qo_fft1d/2d     brute-force DFT (n<=64), Parseval, impulse, numpy.fft   tests/test_oracle_math.py
 *   qo_wavelength   closed-form values at 200/300 kV                         tests/test_oracle_math.py
 *   qo_splinh/seval exact on cubics (Akima reproduces polynomials <=3 only
 *                   locally) -> checked on linear data + interpolation of nodes
 *   qo_probe        sum|psi|^2 = N^2; Cs=df=0 -> real positive aperture sum;
 *                   peak at window centre; chi closed form                   tests/test_oracle_probe.py
 *   qo_make_trans   |t|=1 without band limit; mask pixel count; FFT identity  tests/test_oracle_trans.py
 *   qo_multislice   empty specimen = free propagation (analytic in k);
 *                   detector covering all k = electronScale*sum|psi|^2/N^2;
 *                   reference fixture mulswav_16_2.img norm (statistical)    tests/test_oracle_multislice.py
 *   qo_atom_lut     slab sums of the continuous radial transform of f_e    tests/test_oracle_potential.py
 *   qo_atom_lut2    absolute level: slice sum = 47.8658 f_e(0)/(4 L_z) with
 *                   L_z the LUT z period (clamp off, to 3.5 %); norm switch
 *                   -> f_e(0); clamp only adds                               tests/test_oracle_potential_level.py
 *                   ion offsets: same path with the ion table (equality with
 *                   the neutral LUT given one table), charge term linear,
 *                   Q5 level near the neutral one; offset values and Q5
 *                   pinned statistically by the reference's charged-specimen
 *                   output                                                   tests/test_gpu_reference_fixtures.py
 *   qo_make_slices  translation invariance, x<->y symmetry, peak slice       tests/test_oracle_potential.py
 * Parity unpinned: the exact z-sampling offsets of make3DSlices (rounding of
 * negative offsets, the +Nr row) have no reference-produced fixture.
 */
#include "qs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* FFT: iterative radix-2 for powers of two, Bluestein otherwise.  Sign -1 is
 * forward, +1 backward, both unnormalised (FFTW convention used by QSTEM,
 * libs/data_containers.cpp:50-51). */
static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

static void fft_pow2(double complex *a, int n, int sign) {
  for (int i = 1, j = 0; i < n; i++) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) { double complex t = a[i]; a[i] = a[j]; a[j] = t; }
  }
  for (int len = 2; len <= n; len <<= 1) {
    int half = len >> 1;
    for (int k = 0; k < half; k++) {
      double ang = sign * 2.0 * M_PI * k / len;
      double complex w = cos(ang) + I * sin(ang);
      for (int i = 0; i < n; i += len) {
        double complex u = a[i + k], v = a[i + k + half] * w;
        a[i + k] = u + v;
        a[i + k + half] = u - v;
      }
    }
  }
}

static void fft_bluestein(double complex *a, int n, int sign) {
  int m = 1;
  while (m < 2 * n - 1) m <<= 1;
  double complex *w = (double complex *)malloc(sizeof(double complex) * n);
  double complex *u = (double complex *)calloc(m, sizeof(double complex));
  double complex *v = (double complex *)calloc(m, sizeof(double complex));
  for (int k = 0; k < n; k++) {
    /* exp(sign*i*pi*k^2/n) with k^2 reduced mod 2n for accuracy */
    long long k2 = ((long long)k * k) % (2LL * n);
    double ang = sign * M_PI * (double)k2 / n;
    w[k] = cos(ang) + I * sin(ang);
  }
  for (int k = 0; k < n; k++) u[k] = a[k] * w[k];
  v[0] = conj(w[0]);
  for (int k = 1; k < n; k++) { v[k] = conj(w[k]); v[m - k] = conj(w[k]); }
  fft_pow2(u, m, -1);
  fft_pow2(v, m, -1);
  for (int k = 0; k < m; k++) u[k] *= v[k];
  fft_pow2(u, m, +1);
  for (int k = 0; k < n; k++) a[k] = u[k] * w[k] / m;
  free(w); free(u); free(v);
}

void qo_fft1d(double complex *a, int n, int sign) {
  if (n <= 1) return;
  if (is_pow2(n)) fft_pow2(a, n, sign);
  else fft_bluestein(a, n, sign);
}

/* 2D transform of a row-major [nx][ny] array (ny fastest), as fftw_plan_dft_2d(nx,ny). */
void qo_fft2d(double complex *a, int nx, int ny, int sign) {
  for (int ix = 0; ix < nx; ix++) qo_fft1d(a + (size_t)ix * ny, ny, sign);
  double complex *col = (double complex *)malloc(sizeof(double complex) * nx);
  for (int iy = 0; iy < ny; iy++) {
    for (int ix = 0; ix < nx; ix++) col[ix] = a[(size_t)ix * ny + iy];
    qo_fft1d(col, nx, sign);
    for (int ix = 0; ix < nx; ix++) a[(size_t)ix * ny + iy] = col[ix];
  }
  free(col);
}

/* ------------------------------------------------------------------------ */
/* stem3/stemutil.cpp:714-725 */
double qo_wavelength(double kev) {
  const double emass = 510.99906, hc = 12.3984244;
  return hc / sqrt(kev * (2.0 * emass + kev));
}

/* stem3/stemutil.cpp:927-941 */
double qo_sigma(double kev) {
  const double emass = 510.99906;
  double x = (emass + kev) / (2.0 * emass + kev);
  return 2.0 * M_PI * x / (qo_wavelength(kev) * kev);
}

/* Akima quasi-Hermite spline, stem3/stemutil.cpp:975-1041 (H. Akima, J.ACM 17 (1970) 589). */
void qo_splinh(const double *x, const double *y, double *b, double *c, double *d, int n) {
  const double small = 1.0e-25;
  if (n < 4) return;
  double m5 = (y[3] - y[2]) / (x[3] - x[2]);
  double m4 = (y[2] - y[1]) / (x[2] - x[1]);
  double m3 = (y[1] - y[0]) / (x[1] - x[0]);
  double m2 = 2.0 * m3 - m4;
  double m1 = 2.0 * m2 - m3;
  double m54 = fabs(m5 - m4), m43 = fabs(m4 - m3), m32 = fabs(m3 - m2), m21 = fabs(m2 - m1);
  double t1 = (m43 + m21 > small) ? (m43 * m2 + m21 * m3) / (m43 + m21) : 0.5 * (m2 + m3);
  for (int i = 0; i < n - 1; i++) {
    double t2 = (m54 + m32 > small) ? (m54 * m3 + m32 * m4) / (m54 + m32) : 0.5 * (m3 + m4);
    double h = x[i + 1] - x[i];
    b[i] = t1;
    c[i] = (3.0 * m3 - 2.0 * t1 - t2) / h;
    d[i] = (t1 + t2 - 2.0 * m3) / (h * h);
    m1 = m2; m2 = m3; m3 = m4; m4 = m5;
    m5 = (i < n - 4) ? (y[i + 4] - y[i + 3]) / (x[i + 4] - x[i + 3]) : 2.0 * m4 - m3;
    m21 = m32; m32 = m43; m43 = m54; m54 = fabs(m5 - m4);
    t1 = t2;
  }
}

/* stem3/stemutil.cpp:1069-1092: binary search, end intervals clamp. */
double qo_seval(const double *x, const double *y, const double *b, const double *c,
                const double *d, int n, double x0) {
  int i;
  if (x0 <= x[0]) i = 0;
  else if (x0 >= x[n - 2]) i = n - 2;
  else {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
      int k = (lo + hi) / 2;
      if (x0 < x[k]) hi = k; else lo = k;
    }
    i = lo;
  }
  double z = x0 - x[i];
  return y[i] + (b[i] + (c[i] + d[i] * z) * z) * z;
}

/* ------------------------------------------------------------------------ */
/* probe(), stem3/stemlib.cpp:1952-2191.  The wave is built at the window
 * centre as the driver does (stem3.cpp:2838): dx = nx/2*resX.  Output is the
 * normalised real-space probe [nx][ny] with sum |psi|^2 = nx*ny. */
void qo_probe(const qo_params *p, double complex *psi) { qo_probe_d(p, p->df0, p->Cs, psi); }

static void probe_impl(const qo_params *p, double df0, double Cs, double px, double py, double complex *psi);

/* probe() with defocus and Cs given in double precision (finite-difference checks). */
void qo_probe_d(const qo_params *p, double df0, double Cs, double complex *psi) {
  /* the driver's argument nx/2*resX, float-valued (stem3.cpp:2838) */
  const float cx = (float)(p->nx / 2) * p->resolutionX, cy = (float)(p->ny / 2) * p->resolutionY;
  probe_impl(p, df0, Cs, (double)cx, (double)cy, psi);
}

/* probe() centred at (px, py) Angstrom from the window corner (doCBED, stem3.cpp:2097). */
void qo_probe_at(const qo_params *p, double px, double py, double complex *psi) {
  probe_impl(p, p->df0, p->Cs, px, py, psi);
}

static void probe_impl(const qo_params *p, double df0, double Cs, double px, double py, double complex *psi) {
  const int nx = p->nx, ny = p->ny;
  const double ax = nx * (double)p->resolutionX, by = ny * (double)p->resolutionY;
  /* stemlib.cpp:1974-1975: dx = ax - dx */
  const double dx = ax - px, dy = by - py;
  const double v0 = p->v0;
  const double wavlen = 12.26 / sqrt(v0 * 1.e3 + v0 * v0 * 0.9788); /* :2014 */
  const double rx = 1.0 / ax, ry = 1.0 / by, rx2 = rx * rx, ry2 = ry * ry;
  const double alpha = 0.001 * p->alpha;
  double k2max = sin(alpha) / wavlen;
  k2max *= k2max;
  const double pixel = rx2 + ry2;
  double scale = 1.0 / sqrt((double)nx * (double)ny);
  const double delta = (double)p->Cc * (double)p->dE_E;
  const double a62 = p->compat_q2 ? 0.0 : (double)p->a62;
  for (int iy = 0; iy < ny; iy++) {
    double ky = (iy > ny / 2) ? (double)(iy - ny) : (double)iy;
    double ky2 = ky * ky * ry2;
    for (int ix = 0; ix < nx; ix++) {
      double kx = (ix > nx / 2) ? (double)(ix - nx) : (double)ix;
      double k2 = kx * kx * rx2 + ky2;
      double kt2 = k2 * wavlen * wavlen;
      double kt = sqrt(kt2);
      double phi = atan2(ry * ky, rx * kx);
      double chi = kt2 * (df0 + delta + (double)p->astigMag * cos(2.0 * (phi - (double)p->astigAngle))) / 2.0;
      double ktp = kt2 * kt; /* theta^3 */
      if (p->a33 > 0 || p->a31 > 0)
        chi += ktp * (p->a33 * cos(3.0 * (phi - p->phi33)) + p->a31 * cos(phi - p->phi31)) / 3.0;
      ktp *= kt; /* theta^4 */
      if (p->a44 > 0 || p->a42 > 0 || Cs != 0)
        chi += ktp * (p->a44 * cos(4.0 * (phi - p->phi44)) + p->a42 * cos(2.0 * (phi - p->phi42)) + Cs) / 4.0;
      ktp *= kt; /* theta^5 */
      if (p->a55 > 0 || p->a53 > 0 || p->a51 > 0)
        chi += ktp * (p->a55 * cos(5.0 * (phi - p->phi55)) + p->a53 * cos(3.0 * (phi - p->phi53)) +
                      p->a51 * cos(phi - p->phi51)) / 5.0;
      ktp *= kt; /* theta^6 */
      /* Q2: the reference's condition contains an assignment (a62 = 0) that
       * zeroes a62 and is false; the C6 term is evaluated when any of the
       * other conditions hold.  compat_q2 reproduces it. */
      int c6 = p->a66 > 0 || p->a64 > 0 || p->C5 != 0 || (!p->compat_q2 && p->a62 > 0);
      if (c6)
        chi += ktp * (p->a66 * cos(6.0 * (phi - p->phi66)) + p->a64 * cos(4.0 * (phi - p->phi64)) +
                      a62 * cos(2.0 * (phi - p->phi62)) + p->C5) / 6.0;
      chi *= 2.0 * M_PI / wavlen;
      chi -= 2.0 * M_PI * ((dx * kx / ax) + (dy * ky / by));
      double complex v;
      if (p->ismoth != 0 && fabs(k2 - k2max) <= pixel) v = 0.5 * scale * (cos(chi) - I * sin(chi));
      else if (k2 <= k2max) v = scale * (cos(chi) - I * sin(chi));
      else v = 0.0;
      psi[(size_t)ix * ny + iy] = v;
    }
  }
  qo_fft2d(psi, nx, ny, +1);
  if (p->gaussFlag) { /* :2118-2127 */
    for (int ix = 0; ix < nx; ix++)
      for (int iy = 0; iy < ny; iy++) {
        double r = exp(-((double)((ix - nx / 2) * (ix - nx / 2) + (iy - ny / 2) * (iy - ny / 2))) /
                       ((double)nx * nx * p->gaussScale));
        psi[(size_t)ix * ny + iy] *= r;
      }
  }
  if (p->aAIS > 0) { /* :2130-2148 */
    double avgRes = sqrt(0.5 * ((double)p->resolutionX * p->resolutionX + (double)p->resolutionY * p->resolutionY));
    double edge = 5.0 * avgRes; /* SMOOTH_EDGE = 5 pixels (stemlib.cpp:1939) */
    for (int ix = 0; ix < nx; ix++)
      for (int iy = 0; iy < ny; iy++) {
        double x = ix * (double)p->resolutionX - dx, y = iy * (double)p->resolutionY - dy;
        double r = sqrt(x * x + y * y);
        double dd = r - 0.5 * p->aAIS + edge;
        if (dd > 0) psi[(size_t)ix * ny + iy] = 0;
        else if (dd >= -edge) psi[(size_t)ix * ny + iy] *= 0.5 * (1 - cos(M_PI * dd / edge));
      }
  }
  double sum = 0;
  for (size_t i = 0; i < (size_t)nx * ny; i++) sum += creal(psi[i]) * creal(psi[i]) + cimag(psi[i]) * cimag(psi[i]);
  scale = sqrt((double)nx * (double)ny / sum);
  for (size_t i = 0; i < (size_t)nx * ny; i++) psi[i] *= scale;
}

/* k^2 tables and band limit exactly as propagate_slow builds them in single
 * precision (stem3/stemlib.cpp:2782-2822); collectIntensity reuses them. */
void qo_k_tables(const qo_params *p, float *kx2, float *ky2, float *k2max) {
  const int nx = p->nx, ny = p->ny;
  const float ax = p->resolutionX * nx, by = p->resolutionY * ny;
  for (int i = 0; i < nx; i++) {
    float k = (i > nx / 2) ? (float)(i - nx) / ax : (float)i / ax;
    kx2[i] = k * k;
  }
  for (int i = 0; i < ny; i++) {
    float k = (i > ny / 2) ? (float)(i - ny) / by : (float)i / by;
    ky2[i] = k * k;
  }
  float km = nx / (2.0F * ax);
  if (ny / (2.0F * by) < km) km = ny / (2.0F * by);
  km = (float)(2.0 / 3.0 * km);
  *k2max = km * km;
}

/* stem3.cpp:1338-1347: detector limits in (1/A)^2, single precision. */
void qo_det_k2(const qo_params *p, float *k2in, float *k2out) {
  double wl = qo_wavelength(p->v0);
  for (int i = 0; i < p->ndet; i++) {
    float a = (float)(sin(p->det_rin[i] * 0.001) / wl);
    float b = (float)(sin(p->det_rout[i] * 0.001) / wl);
    k2in[i] = a * a;
    k2out[i] = b * b;
  }
}

/* initSTEMSlices, stem3/stemlib.cpp:2199-2357.  V is the real projected
 * potential [nslice][potNx][potNy]; trans receives exp(i*gamma*lambda*V),
 * band limited to 2/3 of Nyquist when bandlimittrans is set. */
void qo_make_trans(const qo_params *p, const float *V, int nslice, double complex *trans) {
  const size_t n = (size_t)nslice * p->potNx * p->potNy;
  double *Vd = (double *)malloc(sizeof(double) * n);
  for (size_t i = 0; i < n; i++) Vd[i] = V[i];
  qo_make_trans_d(p, Vd, nslice, trans);
  free(Vd);
}

/* Same with a double-precision potential (finite-difference checks). */
/* Interaction constant of initSTEMSlices: gamma*lambda with gamma = 1 + v0/511 in float
 * (stemlib.cpp:2271-2280). */
static double trans_sigma(const qo_params *p) {
  float mm0 = 1.0F + p->v0 / 511.0F;
  return mm0 * qo_wavelength(p->v0);
}

static void bandlimit_d(const qo_params *p, double complex *trans, int nslice);

void qo_make_trans_d(const qo_params *p, const double *V, int nslice, double complex *trans) {
  const size_t n2 = (size_t)p->potNx * p->potNy;
  const double scale = trans_sigma(p);
  for (size_t i = 0; i < (size_t)nslice * n2; i++) {
    double vz = V[i] * scale;
    trans[i] = cos(vz) + I * sin(vz);
  }
  if (p->bandlimittrans) bandlimit_d(p, trans, nslice);
}

/* The band limit of initSTEMSlices (stemlib.cpp:2290-2336): FFT, 2/3 disc with
 * 1/(potNx potNy), IFFT, in place on [nslice][potNx][potNy]; a real diagonal mask
 * between F and its inverse, so the operator is self-adjoint. */
static void bandlimit_d(const qo_params *p, double complex *trans, int nslice) {
  const int nx = p->potNx, ny = p->potNy;
  const size_t n2 = (size_t)nx * ny;
  /* k-tables (:2221-2238) in float; Q1 uses muls->by for negative ky */
  float *kx2 = (float *)malloc(sizeof(float) * nx), *ky2 = (float *)malloc(sizeof(float) * ny);
  for (int i = 0; i < nx; i++) {
    float k = (i > nx / 2) ? (float)(i - nx) / p->potSizeX : (float)i / p->potSizeX;
    kx2[i] = k * k;
  }
  for (int i = 0; i < ny; i++) {
    float k = (i > ny / 2) ? (float)(i - ny) / (p->compat_q1 ? p->by : p->potSizeY) : (float)i / p->potSizeY;
    ky2[i] = k * k;
  }
  float k2max = nx / (2.0F * p->potSizeX);
  float t = ny / (2.0F * p->potSizeY);
  if (t < k2max) k2max = t;
  k2max = (float)((2.0F / 3.0F) * k2max);
  k2max = k2max * k2max;
  double fftScale = 1.0 / ((double)nx * ny);
  for (int s = 0; s < nslice; s++) {
    double complex *a = trans + (size_t)s * n2;
    qo_fft2d(a, nx, ny, -1);
    for (int ix = 0; ix < nx; ix++)
      for (int iy = 0; iy < ny; iy++) {
        float k2 = ky2[iy] + kx2[ix];
        if (k2 < k2max) a[(size_t)ix * ny + iy] *= fftScale;
        else a[(size_t)ix * ny + iy] = 0;
      }
    qo_fft2d(a, nx, ny, +1);
  }
  free(kx2); free(ky2);
}

/* stem3.cpp:2860-2871: integer window offset of a scan position. */
void qo_scan_offsets(int ix, int iy, float scanXStart, float scanXStop, int scanXN,
                     float scanYStart, float scanYStop, int scanYN,
                     float resX, float resY, int potNx, int potNy, int nx, int ny,
                     int *iPosX, int *iPosY) {
  int px = (int)(ix * (scanXStop - scanXStart) / ((float)scanXN * resX));
  int py = (int)(iy * (scanYStop - scanYStart) / ((float)scanYN * resY));
  if (px > potNx - nx) px = potNx - nx;
  if (py > potNy - ny) py = potNy - ny;
  *iPosX = px;
  *iPosY = py;
}

/* collectIntensity bin assignment, stem3/stemlib.cpp:2588-2607, combined
 * with "last write wins" (Detector::Navg stays 0, Q3).  A slab runs
 * mulsRepeat1 passes over its slices (stemlib.cpp:2403-2405); step
 * k = mRepeat*slices + islice passes slice index
 * totalSliceCount + islice*(1+mRepeat) (:2429).  bin_of_step[k] receives the
 * bin that step k leaves as final value within this slab, else -1.
 * Returns the number of bins (tCount+1). */
int qo_record_bins(int slices, int cellDiv, int outputInterval, int totalSliceCount,
                   int mulsRepeat1, int *bin_of_step) {
  int total = slices * cellDiv;
  int tCount = outputInterval > 0 ? (int)ceil((double)(total / outputInterval)) : 0;
  int nb = tCount + 1;
  int nsteps = slices * (mulsRepeat1 > 0 ? mulsRepeat1 : 1);
  int *last = (int *)malloc(sizeof(int) * nb);
  for (int t = 0; t < nb; t++) last[t] = -1;
  for (int k = 0; k < nsteps; k++) {
    int rep = k / slices, s = k % slices;
    int slice = totalSliceCount + s * (1 + rep);
    int t;
    if (outputInterval == 0) t = 0;
    else if (slice < total - 1) t = slice / outputInterval;
    else t = tCount;
    bin_of_step[k] = -1;
    if (t >= 0 && t < nb) last[t] = k;
  }
  for (int t = 0; t < nb; t++)
    if (last[t] >= 0) bin_of_step[last[t]] = t;
  free(last);
  return nb;
}

/* runMulsSTEM, stem3/stemlib.cpp:2375-2534, with transmit (:2867-2888),
 * propagate_slow (:2760-2850), collectIntensity (:2576-2687) and
 * fft_normalize (:2890-2906), for a batch of probes. */
void qo_multislice(const qo_params *p, const double complex *trans, int nslice,
                   const int *ipos, int nprobe, double complex *psi_io,
                   const int *bin_of_slice, int nbins, double *det, double *diff) {
  const int nx = p->nx, ny = p->ny;
  const size_t n2 = (size_t)nx * ny;
  float *kx2 = (float *)malloc(sizeof(float) * nx), *ky2 = (float *)malloc(sizeof(float) * ny);
  float k2max;
  qo_k_tables(p, kx2, ky2, &k2max);
  float k2in[QO_MAX_DET], k2out[QO_MAX_DET];
  qo_det_k2(p, k2in, k2out);
  /* propagator e^{-i pi dz lambda k^2} in double, separable like propxr/propyr */
  const double wl = qo_wavelength(p->v0), dz = p->sliceThickness;
  const double ax = (double)p->resolutionX * nx, by = (double)p->resolutionY * ny;
  double complex *px = (double complex *)malloc(sizeof(double complex) * nx);
  double complex *py = (double complex *)malloc(sizeof(double complex) * ny);
  for (int i = 0; i < nx; i++) {
    double k = ((i > nx / 2) ? (double)(i - nx) : (double)i) / ax;
    double t = M_PI * dz * wl * k * k;
    px[i] = cos(t) - I * sin(t);
  }
  for (int i = 0; i < ny; i++) {
    double k = ((i > ny / 2) ? (double)(i - ny) : (double)i) / by;
    double t = M_PI * dz * wl * k * k;
    py[i] = cos(t) - I * sin(t);
  }
  const double escale = (double)p->electronScale / ((double)n2 * (double)n2);
  const double dscale = 1.0 / sqrt((double)n2);
  double complex *w = (double complex *)malloc(sizeof(double complex) * n2);
  for (int ip = 0; ip < nprobe; ip++) {
    memcpy(w, psi_io + ip * n2, sizeof(double complex) * n2);
    const int posx = ipos[2 * ip], posy = ipos[2 * ip + 1];
    for (int s = 0; s < nslice; s++) {
      const double complex *t = trans + (size_t)s * p->potNx * p->potNy;
      for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++)
          w[(size_t)ix * ny + iy] *= t[(size_t)(ix + posx) * p->potNy + iy + posy];
      qo_fft2d(w, nx, ny, -1);
      for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++) {
          size_t k = (size_t)ix * ny + iy;
          if (kx2[ix] < k2max && (kx2[ix] + ky2[iy]) < k2max) w[k] *= px[ix] * py[iy];
          else w[k] = 0;
        }
      int bin = bin_of_slice ? bin_of_slice[s] : -1;
      if (det && bin >= 0) {
        for (int d = 0; d < p->ndet; d++) {
          int sx = (int)p->det_shiftx[d], sy = (int)p->det_shifty[d];
          double acc = 0;
          for (int ix = 0; ix < nx; ix++)
            for (int iy = 0; iy < ny; iy++) {
              float k2 = kx2[ix] + ky2[iy];
              if (k2 >= k2in[d] && k2 <= k2out[d]) {
                size_t k;
                if (sx == 0 && sy == 0) k = (size_t)ix * ny + iy;
                else k = (size_t)((ix + sx + nx) % nx) * ny + (size_t)((iy + sy + ny) % ny);
                acc += creal(w[k]) * creal(w[k]) + cimag(w[k]) * cimag(w[k]);
              }
            }
          det[((size_t)bin * p->ndet + d) * nprobe + ip] = acc * escale;
        }
      }
      if (diff && s == nslice - 1) {
        double *dp = diff + ip * n2;
        for (int ix = 0; ix < nx; ix++)
          for (int iy = 0; iy < ny; iy++) {
            double complex v = w[(size_t)ix * ny + iy];
            dp[(size_t)((ix + nx / 2) % nx) * ny + (iy + ny / 2) % ny] =
                (creal(v) * creal(v) + cimag(v) * cimag(v)) * dscale;
          }
      }
      qo_fft2d(w, nx, ny, +1);
      for (size_t k = 0; k < n2; k++) w[k] /= (double)n2;
    }
    memcpy(psi_io + ip * n2, w, sizeof(double complex) * n2);
  }
  free(w); free(px); free(py); free(kx2); free(ky2);
  (void)nbins;
}

/* FP64 adjoint of the multislice for the loss of the library's qs_adjoint,
 *   L = sum_p sum_k (I_p(k) - T_p(k))^2,  I_p = |Phi_p|^2,  Phi_p = FFT(psi_exit,p) / N^2,
 * the diffraction intensity after the last slice (N^2 = nx*ny), with every slice of V
 * and the incident waves psi0 [nx][ny] shared by the nprobe window offsets ipos.  Not a
 * reference routine (QSTEM has no adjoint): reverse mode of qo_multislice and
 * qo_make_trans_d.  Conventions: for complex z, g_z = dL/d(conj z), so that
 * dL = 2 Re sum conj(g_z) dz; through y = F x (unnormalised DFT) g_x = F^H g_y = IDFT(g_y);
 * through y = x c, g_x = conj(c) g_y and g_c += conj(x) g_y; t = BL(u) with the
 * self-adjoint band limit, u = exp(i sigma V): dL/dV = 2 sigma Im(g_u conj(u)).
 * Outputs: loss; gV [nslice][potNx][potNy] (nullable); gpsi0 [nx][ny] = g of the incident
 * wave summed over probes (nullable).  Waves of all slices are kept per probe (FP64). */
void qo_adjoint(const qo_params *p, const double *V, int nslice, const int *ipos, int nprobe,
                const double complex *psi0, const double *target, double *loss, double *gV,
                double complex *gpsi0) {
  const int nx = p->nx, ny = p->ny, pnx = p->potNx, pny = p->potNy;
  const size_t n2 = (size_t)nx * ny, slice = (size_t)pnx * pny;
  double complex *t = (double complex *)malloc(sizeof(double complex) * slice * nslice);
  qo_make_trans_d(p, V, nslice, t);
  double complex *gt = (double complex *)calloc(slice * nslice, sizeof(double complex));
  float *kx2 = (float *)malloc(sizeof(float) * nx), *ky2 = (float *)malloc(sizeof(float) * ny);
  float k2max;
  qo_k_tables(p, kx2, ky2, &k2max);
  const double wl = qo_wavelength(p->v0), dz = p->sliceThickness;
  const double ax = (double)p->resolutionX * nx, by = (double)p->resolutionY * ny;
  double complex *P = (double complex *)malloc(sizeof(double complex) * n2); /* propagator x mask */
  for (int ix = 0; ix < nx; ix++) {
    double kx = ((ix > nx / 2) ? (double)(ix - nx) : (double)ix) / ax;
    for (int iy = 0; iy < ny; iy++) {
      double ky = ((iy > ny / 2) ? (double)(iy - ny) : (double)iy) / by;
      double tx = M_PI * dz * wl * kx * kx, ty = M_PI * dz * wl * ky * ky;
      int in = kx2[ix] < k2max && (kx2[ix] + ky2[iy]) < k2max;
      P[(size_t)ix * ny + iy] = in ? (cos(tx) - I * sin(tx)) * (cos(ty) - I * sin(ty)) : 0;
    }
  }
  double complex *psi = (double complex *)malloc(sizeof(double complex) * n2 * (nslice + 1));
  double complex *g = (double complex *)malloc(sizeof(double complex) * n2);
  if (gpsi0) memset(gpsi0, 0, sizeof(double complex) * n2);
  const double N2 = (double)n2;
  double L = 0;
  for (int ip = 0; ip < nprobe; ip++) {
    const int px = ipos[2 * ip], py = ipos[2 * ip + 1];
    memcpy(psi, psi0, sizeof(double complex) * n2);
    /* forward: psi[s+1] = IDFT(P DFT(psi[s] t_s)) / N^2 */
    for (int s = 0; s < nslice; s++) {
      const double complex *ts = t + (size_t)s * slice;
      double complex *a = psi + (size_t)(s + 1) * n2;
      for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++)
          a[(size_t)ix * ny + iy] = psi[(size_t)s * n2 + (size_t)ix * ny + iy] * ts[(size_t)(ix + px) * pny + iy + py];
      qo_fft2d(a, nx, ny, -1);
      for (size_t k = 0; k < n2; k++) a[k] *= P[k];
      qo_fft2d(a, nx, ny, +1);
      for (size_t k = 0; k < n2; k++) a[k] /= N2;
    }
    /* loss and its seed: Phi = DFT(psi_S)/N^2, g_Phi = 2 (I - T) Phi, g_psiS = IDFT(g_Phi)/N^2 */
    memcpy(g, psi + (size_t)nslice * n2, sizeof(double complex) * n2);
    qo_fft2d(g, nx, ny, -1);
    const double *T = target + (size_t)ip * n2;
    for (size_t k = 0; k < n2; k++) {
      double complex phi = g[k] / N2;
      double I_ = creal(phi) * creal(phi) + cimag(phi) * cimag(phi);
      L += (I_ - T[k]) * (I_ - T[k]);
      g[k] = 2.0 * (I_ - T[k]) * phi / N2;
    }
    qo_fft2d(g, nx, ny, +1);
    /* reverse: through psi[s+1] = IDFT(P DFT(a_s)) / N^2 and a_s = psi[s] t_s */
    for (int s = nslice - 1; s >= 0; s--) {
      qo_fft2d(g, nx, ny, -1); /* (IDFT)^H = DFT */
      for (size_t k = 0; k < n2; k++) g[k] = conj(P[k]) * g[k] / N2;
      qo_fft2d(g, nx, ny, +1); /* DFT^H = IDFT: g is now g_a */
      const double complex *ts = t + (size_t)s * slice;
      double complex *gts = gt + (size_t)s * slice;
      const double complex *ps = psi + (size_t)s * n2;
      for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++) {
          const size_t k = (size_t)ix * ny + iy, w = (size_t)(ix + px) * pny + iy + py;
          gts[w] += conj(ps[k]) * g[k];
          g[k] = conj(ts[w]) * g[k];
        }
    }
    if (gpsi0)
      for (size_t k = 0; k < n2; k++) gpsi0[k] += g[k];
  }
  *loss = L;
  if (gV) {
    /* g_u = BL(g_t) (self-adjoint), dL/dV = 2 sigma Im(g_u conj(u)), u = exp(i sigma V) */
    if (p->bandlimittrans) bandlimit_d(p, gt, nslice);
    const double sg = trans_sigma(p);
    for (size_t i = 0; i < slice * nslice; i++) {
      const double complex u = cos(V[i] * sg) + I * sin(V[i] * sg);
      gV[i] = 2.0 * sg * cimag(gt[i] * conj(u));
    }
  }
  free(t); free(gt); free(kx2); free(ky2); free(P); free(psi); free(g);
}

/* ------------------------------------------------------------------------ */
/* getAtomPotential3D geometry, stem3/stemlib.cpp:1366-1388 (OVERSAMP_X = 2). */
void qo_lut_geometry(float atomRadius, float resX, float sliceThickness, qo_lut_geom *g) {
  g->nx = 2 * 2 * (int)ceil(atomRadius / resX);
  int nzp = (int)floor(2 * sliceThickness / resX);
  if (2.0 * (nzp >> 1) == nzp) nzp += 1;
  g->nzPerSlice = nzp;
  g->nz = (2 * (int)ceil(atomRadius / sliceThickness)) * nzp;
  g->dkx = 0.5 * 2 / (double)(float)(g->nx * resX);          /* int*float product in float */
  g->dkz = nzp / (double)(float)(g->nz * sliceThickness);
  double kmax = 0.5 * g->nx * g->dkx / 2.0;
  g->smax2 = kmax; /* squared by the caller after tapering; stored unsquared here */
  g->Nr = g->nx / 2;
  g->Nz_lut = g->nz / 2;
}

/* :1390-1413: overwrite the last three s-grid nodes and taper the tail. */
void qo_taper_scatpar(double *sgrid, double *rows, int nrows, double smax) {
  const int nsf = 123;
  sgrid[nsf - 1] = 1.2 * smax;
  sgrid[nsf - 2] = 1.1 * smax;
  sgrid[nsf - 3] = smax;
  if (sgrid[nsf - 4] > sgrid[nsf - 3]) {
    for (int ix = 0; ix < nsf - 10; ix++) {
      if (sgrid[nsf - 4 - ix] < sgrid[nsf - 3] - 0.001 * (ix + 1)) break;
      sgrid[nsf - 4 - ix] = sgrid[nsf - 3] - 0.001 * (ix + 1);
      for (int r = 0; r < nrows; r++) rows[(size_t)r * nsf + nsf - 4 - ix] = 0;
    }
  }
}

/* LUT V(r,z) of one element, stem3/stemlib.cpp:1419-1532.  The reference
 * builds the k-space map in float and inverse transforms with fftwf; this
 * restatement keeps everything double. */
void qo_atom_lut(const qo_lut_geom *g, const double *sgrid, const double *frow, double B,
                 float sliceThickness, float *lut) {
  qo_atom_lut2(g, sgrid, frow, sgrid, frow, g->dkx, B, sliceThickness, 1.0, 1, lut);
}

/* Scale that turns the reference's LUT normalisation 47.8658 dkx dkz / nz
 * (stemlib.cpp:1516, :1720) into 4 dkx dkz dz/nzPerSlice: the slice-summed
 * projected potential of an atom then integrates to f_e(0) (the phase integral
 * gamma*lambda*f_e(0)) at every slice thickness, instead of to
 * 47.8658 f_e(0) / (4 L_z) with L_z = nz dz / nzPerSlice = 2 ceil(R/dz) dz the
 * LUT's z period.  The 4 undoes the radial factor 1/4 of reading the LUT (radial
 * spacing dx) at dr = dx/2 (stemlib.cpp:576).  Not a reference routine. */
double qo_lut_norm_fix(const qo_lut_geom *g, float sliceThickness) {
  return 4.0 * g->nz * ((double)sliceThickness / g->nzPerSlice) / 47.8658;
}

/* General LUT builder: getAtomPotential3D (stemlib.cpp:1419-1532) when
 * (sgrid_y, frow_y) = (sgrid, frow) and dky = dkx; getAtomPotentialOffset3D
 * (stemlib.cpp:1629-1725) with (sgrid, frow) the ion-offset table.  The spline
 * is fitted to (sgrid, frow) and evaluates the (kx, kz) term; the q_y sum
 * evaluates seval(sgrid_y, frow_y, <that spline's b, c, d>) with step dky, which
 * for the ion offsets with frow_y the neutral row is the reference's quirk Q5
 * (stemlib.cpp:1672).  `scale` multiplies the result (1 = reference); clamp = 1
 * sets negative slab sums to zero as the reference does (stemlib.cpp:1508); the
 * tests turn it off to compare the LUT's integral with f_e(0) exactly. */
void qo_atom_lut2(const qo_lut_geom *g, const double *sgrid, const double *frow, const double *sgrid_y,
                  const double *frow_y, double dky, double B, float sliceThickness, double scale, int clamp,
                  float *lut) {
  const int nsf = 123, nx = g->nx, nz = g->nz, nzp = g->nzPerSlice;
  const double smax2 = g->smax2 * g->smax2, dkx = g->dkx, dkz = g->dkz;
  double b[123], c[123], d[123];
  qo_splinh(sgrid, frow, b, c, d, nsf);
  double complex *tmp = (double complex *)calloc((size_t)nx * nz, sizeof(double complex));
  const int izOffset = (nzp - 1) / 2;
  const double zPos = -2.0 * M_PI * ((double)sliceThickness / nzp * izOffset);
  for (int iz = 0; iz < nz; iz++) {
    double kz = dkz * (iz < nz / 2 ? iz : iz - nz);
    for (int ix = 0; ix < nx; ix++) {
      double kx = dkx * (ix < nx / 2 ? ix : ix - nx);
      double s2 = kx * kx + kz * kz;
      if (s2 < smax2) {
        double f = qo_seval(sgrid, frow, b, c, d, nsf, sqrt(s2)) * exp(-s2 * B * 0.25);
        for (int iy = 1; iy < nx; iy++) {
          double s3 = dky * iy;
          s3 = s3 * s3 + s2;
          if (s3 < smax2) f += 2 * qo_seval(sgrid_y, frow_y, b, c, d, nsf, sqrt(s3)) * exp(-s3 * B * 0.25);
          else break;
        }
        f *= dkx;
        double phase = kz * zPos;
        tmp[ix + (size_t)iz * nx] = f * (cos(phase) + I * sin(phase));
      }
    }
  }
  qo_fft2d(tmp, nz, nx, +1);
  for (int ix = 0; ix < nx / 2; ix++)
    for (int iz = 0; iz < nz / 2; iz++) {
      double zs = 0;
      for (int iiz = -izOffset; iiz <= izOffset; iiz++)
        if (iz + izOffset + iiz < nz / 2) zs += creal(tmp[ix + (size_t)(iz + izOffset + iiz) * nx]);
      if (clamp && zs < 0) zs = 0;
      lut[ix + (size_t)iz * (nx / 2)] = (float)(47.8658 * dkx * dkz / nz * zs * scale);
    }
  free(tmp);
}

/* make3DSlices, fftpotential + potential3D branches (stem3/stemlib.cpp:883-1213),
 * single slab (cellDiv = 1), non-periodic z, Z_INTERPOLATION = 0, neutral atoms. */
void qo_make_slices(const qo_params *p, const qo_lut_geom *g, const float *const *luts,
                    const int *kind_of_atom, const qo_atom *atoms, int natom,
                    float potOffsetX, float potOffsetY, float czOffset, int centerSlices,
                    float atomRadius, int periodic, float *V) {
  qo_make_slices_q(p, g, luts, NULL, kind_of_atom, atoms, natom, potOffsetX, potOffsetY, czOffset, centerSlices,
                   atomRadius, periodic, V);
}

/* qo_make_slices with ion charges (USE_Q_POT_OFFSETS, stemlib.cpp:77): an atom of
 * charge q != 0 adds q times the same r-linear / z-nearest sample of its
 * element's offset LUT offs_luts[kind] (stemlib.cpp:966-979, :993-1006,
 * :1150-1163, :1176-1189).  offs_luts may be NULL (no charged atoms). */
void qo_make_slices_q(const qo_params *p, const qo_lut_geom *g, const float *const *luts,
                      const float *const *offs_luts, const int *kind_of_atom, const qo_atom *atoms, int natom,
                      float potOffsetX, float potOffsetY, float czOffset, int centerSlices,
                      float atomRadius, int periodic, float *V) {
  const int nx = p->potNx, ny = p->potNy, nslc = p->slices;
  const float dx = p->resolutionX, dy = p->resolutionY, dzs = p->sliceThickness;
  const float dr = p->resolutionX / 2;
  const int iRadX = (int)ceil(atomRadius / dx), iRadY = (int)ceil(atomRadius / dy);
  const int iRadZ = (int)ceil(atomRadius / dzs);
  const float c = dzs * nslc;
  const int Nr = g->Nr, Nz = g->Nz_lut, nzSub = g->nzPerSlice;
  const int limHi = Nr * (Nz - 1), limLo = -Nr * (Nz - 1), step = nzSub * Nr;
  memset(V, 0, sizeof(float) * (size_t)nslc * nx * ny);
  for (int ia = 0; ia < natom; ia++) {
    if (atoms[ia].Znum == 0) continue;
    /* :886-889 then :945 (the -dz and +dz cancel up to float rounding) */
    float atomZ = (float)(atoms[ia].z + czOffset - (0.5 * dzs * (1 - centerSlices)));
    atomZ -= dzs;
    if (atomZ - atomRadius > c) break;
    if (atomZ + atomRadius + dzs < 0) continue;
    atomZ += dzs;
    const float atomX = atoms[ia].x - potOffsetX, atomY = atoms[ia].y - potOffsetY;
    const int iAtomX = (int)floor(atomX / dx), iAtomY = (int)floor(atomY / dy);
    const int iAtomZ = (int)floor(atomZ / dzs + 0.5);
    int iax0, iax1, iay0, iay1;
    if (!periodic) {
      iax0 = iAtomX - iRadX < 0 ? 0 : iAtomX - iRadX;
      iax1 = iAtomX + iRadX >= nx ? nx - 1 : iAtomX + iRadX;
      iay0 = iAtomY - iRadY < 0 ? 0 : iAtomY - iRadY;
      iay1 = iAtomY + iRadY >= ny ? ny - 1 : iAtomY + iRadY;
      if (!(iax0 < nx && iax1 >= 0 && iay0 < ny && iay1 >= 0)) continue;
    } else { /* :1106-1109, upper bounds exclusive */
      iax0 = iAtomX - iRadX; iax1 = iAtomX + iRadX - 1;
      iay0 = iAtomY - iRadY; iay1 = iAtomY + iRadY - 1;
    }
    const int iaz0 = iAtomZ - iRadZ < 0 ? -iAtomZ : -iRadZ;
    const int iaz1 = iAtomZ + iRadZ >= nslc ? nslc - iAtomZ - 1 : iRadZ;
    if (!(iAtomZ + iaz0 < nslc && iAtomZ + iaz1 >= 0)) continue;
    const float *lut = luts[kind_of_atom[ia]];
    const float q = atoms[ia].q;
    const float *olut = (q != 0 && offs_luts) ? offs_luts[kind_of_atom[ia]] : NULL;
    const float zrel = atomZ / dzs;
    const double dOffsZ = (float)((float)(iAtomZ + iaz0) - zrel) * nzSub;
    const int iOffsZ0 = (int)(dOffsZ + 0.5) * Nr;
    for (int iax = iax0; iax <= iax1; iax++) {
      double x2 = (float)(iax * dx - atomX);
      x2 *= x2;
      const int ixw = periodic ? (((iax % nx) + nx) % nx) : iax;
      for (int iay = iay0; iay <= iay1; iay++) {
        double y2 = (float)(iay * dy - atomY);
        y2 *= y2;
        double r = sqrt(x2 + y2);
        double ddr = r / dr;
        int ir = (int)floor(ddr);
        if (ir >= Nr - 1) continue;
        ddr -= ir;
        const int iyw = periodic ? (((iay % ny) + ny) % ny) : iay;
        int iOffsZ = iOffsZ0;
        for (int iaz = iaz0; iaz <= iaz1; iaz++) {
          double pv = 0;
          if (iOffsZ < 0) {
            if (iOffsZ > limLo) {
              pv = (1 - ddr) * lut[ir - iOffsZ + Nr] + ddr * lut[ir + 1 - iOffsZ + Nr];
              if (olut) pv += q * ((1 - ddr) * olut[ir - iOffsZ + Nr] + ddr * olut[ir + 1 - iOffsZ + Nr]);
            }
          } else if (iOffsZ < limHi) {
            pv = (1 - ddr) * lut[ir + iOffsZ] + ddr * lut[ir + 1 + iOffsZ];
            if (olut) pv += q * ((1 - ddr) * olut[ir + iOffsZ] + ddr * olut[ir + 1 + iOffsZ]);
          }
          float *cell = V + ((size_t)(iAtomZ + iaz) * nx + ixw) * ny + iyw;
          *cell = (float)(*cell + pv);
          iOffsZ += step;
        }
      }
    }
  }
}

/* make3DSlices with double-precision atom positions and position arithmetic
 * (same discrete sampling as qo_make_slices; used for finite differences of
 * the potential with respect to atom positions). */
void qo_make_slices_d(const qo_params *p, const qo_lut_geom *g, const float *const *luts,
                      const int *kind_of_atom, const double *xyz, const int *Znum, int natom,
                      double potOffsetX, double potOffsetY, double czOffset, int centerSlices,
                      double atomRadius, double dr, double *V) {
  const int nx = p->potNx, ny = p->potNy, nslc = p->slices;
  const double dx = p->resolutionX, dy = p->resolutionY, dzs = p->sliceThickness;
  const int iRadX = (int)ceil(atomRadius / dx), iRadY = (int)ceil(atomRadius / dy);
  const int iRadZ = (int)ceil(atomRadius / dzs);
  const double c = dzs * nslc;
  const int Nr = g->Nr, Nz = g->Nz_lut, nzSub = g->nzPerSlice;
  const int limHi = Nr * (Nz - 1), limLo = -Nr * (Nz - 1), step = nzSub * Nr;
  memset(V, 0, sizeof(double) * (size_t)nslc * nx * ny);
  for (int ia = 0; ia < natom; ia++) {
    if (Znum[ia] == 0) continue;
    double atomZ = xyz[3 * ia + 2] + czOffset - 0.5 * dzs * (1 - centerSlices);
    if (atomZ - dzs - atomRadius > c) continue;
    if (atomZ + atomRadius < 0) continue;
    const double atomX = xyz[3 * ia] - potOffsetX, atomY = xyz[3 * ia + 1] - potOffsetY;
    const int iAtomX = (int)floor(atomX / dx), iAtomY = (int)floor(atomY / dy);
    const int iAtomZ = (int)floor(atomZ / dzs + 0.5);
    int iax0 = iAtomX - iRadX < 0 ? 0 : iAtomX - iRadX;
    int iax1 = iAtomX + iRadX >= nx ? nx - 1 : iAtomX + iRadX;
    int iay0 = iAtomY - iRadY < 0 ? 0 : iAtomY - iRadY;
    int iay1 = iAtomY + iRadY >= ny ? ny - 1 : iAtomY + iRadY;
    if (!(iax0 < nx && iax1 >= 0 && iay0 < ny && iay1 >= 0)) continue;
    const int iaz0 = iAtomZ - iRadZ < 0 ? -iAtomZ : -iRadZ;
    const int iaz1 = iAtomZ + iRadZ >= nslc ? nslc - iAtomZ - 1 : iRadZ;
    if (!(iAtomZ + iaz0 < nslc && iAtomZ + iaz1 >= 0)) continue;
    const float *lut = luts[kind_of_atom[ia]];
    const double dOffsZ = ((double)(iAtomZ + iaz0) - atomZ / dzs) * nzSub;
    const int iOffsZ0 = (int)(dOffsZ + 0.5) * Nr;
    for (int iax = iax0; iax <= iax1; iax++) {
      double x2 = iax * dx - atomX;
      x2 *= x2;
      for (int iay = iay0; iay <= iay1; iay++) {
        double y2 = iay * dy - atomY;
        y2 *= y2;
        double r = sqrt(x2 + y2);
        double ddr = r / dr;
        int ir = (int)floor(ddr);
        if (ir >= Nr - 1) continue;
        ddr -= ir;
        int iOffsZ = iOffsZ0;
        for (int iaz = iaz0; iaz <= iaz1; iaz++) {
          double pv = 0;
          if (iOffsZ < 0) {
            if (iOffsZ > limLo) pv = (1 - ddr) * lut[ir - iOffsZ + Nr] + ddr * lut[ir + 1 - iOffsZ + Nr];
          } else if (iOffsZ < limHi) {
            pv = (1 - ddr) * lut[ir + iOffsZ] + ddr * lut[ir + 1 + iOffsZ];
          }
          V[((size_t)(iAtomZ + iaz) * nx + iax) * ny + iay] += pv;
          iOffsZ += step;
        }
      }
    }
  }
}
