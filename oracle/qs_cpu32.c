/*
 * qs_cpu32.c -- plain CPU FP32 multislice: the CPU baseline of the benchmark.
 *
 * BENCHMARK AND TEST INFRASTRUCTURE ONLY, like the rest of oracle/: bench.py's
 * cpu_baseline leg times it and tests/ check it against the FP64 oracle; the
 * product library (libqstem_hip.so) never links or calls it.
 *
 * Restates the reference's scan loop on the host: doSTEM's OpenMP parallel
 * for over probe positions, one wave per thread (stem3/stem3.cpp:2820-2946),
 * and per slice runMulsSTEM's transmit, forward FFT, propagate_slow,
 * collectIntensity, inverse FFT and fft_normalize (stem3/stemlib.cpp:2405-2443,
 * 2576-2687, 2760-2906), in single precision as the reference computes the
 * wave (FLOAT_PRECISION 1, fftwf).  FFTW is not available here, so the 2D
 * transforms are this file's own Stockham FFT (radix-4 passes, one radix-2 pass
 * for an odd power of two, radix-3 and radix-5 passes for those factors) with
 * precomputed twiddles, on split real/imaginary arrays, batched so that the
 * innermost loop runs over contiguous transforms and vectorises.  The wave is
 * cache-blocked into two passes per slice:
 *   rows    (blocks of 16 rows): row IFFT of the previous slice with the 1/(nx ny)
 *           of fft_normalize, x transmission row, row FFT;
 *   columns (strips of 32 columns): column FFT, x propagator with the 2/3 band
 *           limit, detector annulus sums, column IFFT.
 * Sizes from 16 to 4096 whose prime factors are 2, 3 and 5 (400 = 2^4 5^2 of the
 * reference's own example).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "qs_oracle.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

enum { RB = 16, CB = 32 }; /* rows per row block, columns per column strip (the last may be narrower) */

/* One Stockham pass of radix 3 or 5 over x[n][B] -> y[n][B] (split arrays, B contiguous
 * transforms): butterflies j < n/R, k = j mod Ns, inputs j + r n/R times
 * w^r = exp(sign 2 pi i r k/(R Ns)) = tw[r k n/(R Ns)] (conjugated for sign +1), a
 * length-R DFT, outputs (j/Ns) R Ns + k + r Ns. */
static void pass35(const float *restrict ar, const float *restrict ai, float *restrict br, float *restrict bi, int n,
                   int B, int R, int Ns, const float *restrict twr, const float *restrict twi, int sign) {
  const int q = n / R, step = n / (R * Ns);
  const float sg = (float)sign;
  for (int j = 0; j < q; j++) {
    const int k = j % Ns;
    float wr[5] = {1, 1, 1, 1, 1}, wi[5] = {0, 0, 0, 0, 0};
    for (int r = 1; r < R; r++) {
      wr[r] = twr[r * k * step];
      wi[r] = sign < 0 ? twi[r * k * step] : -twi[r * k * step];
    }
    const int o = (j / Ns) * R * Ns + k;
    if (R == 3) {
      const float h = 0.86602540378443865f; /* sin(2 pi/3) */
      for (int b = 0; b < B; b++) {
        const float x0r = ar[(size_t)j * B + b], x0i = ai[(size_t)j * B + b];
        const float u1r = ar[(size_t)(j + q) * B + b], u1i = ai[(size_t)(j + q) * B + b];
        const float u2r = ar[(size_t)(j + 2 * q) * B + b], u2i = ai[(size_t)(j + 2 * q) * B + b];
        const float x1r = u1r * wr[1] - u1i * wi[1], x1i = u1r * wi[1] + u1i * wr[1];
        const float x2r = u2r * wr[2] - u2i * wi[2], x2i = u2r * wi[2] + u2i * wr[2];
        const float sr = x1r + x2r, si = x1i + x2i, dr = x1r - x2r, di = x1i - x2i;
        const float tr = x0r - 0.5f * sr, ti = x0i - 0.5f * si;
        const float ur = -sg * h * di, ui = sg * h * dr; /* sign i (sqrt3/2) (x1 - x2) */
        br[(size_t)o * B + b] = x0r + sr; bi[(size_t)o * B + b] = x0i + si;
        br[(size_t)(o + Ns) * B + b] = tr + ur; bi[(size_t)(o + Ns) * B + b] = ti + ui;
        br[(size_t)(o + 2 * Ns) * B + b] = tr - ur; bi[(size_t)(o + 2 * Ns) * B + b] = ti - ui;
      }
    } else {
      const float c1 = 0.30901699437494742f, c2 = -0.80901699437494742f;
      const float s1 = 0.95105651629515357f, s2 = 0.58778525229247313f;
      for (int b = 0; b < B; b++) {
        float xr[5], xi[5];
        for (int r = 0; r < 5; r++) {
          const float vr = ar[(size_t)(j + r * q) * B + b], vi = ai[(size_t)(j + r * q) * B + b];
          xr[r] = vr * wr[r] - vi * wi[r];
          xi[r] = vr * wi[r] + vi * wr[r];
        }
        const float b1r = xr[1] + xr[4], b1i = xi[1] + xi[4], b2r = xr[2] + xr[3], b2i = xi[2] + xi[3];
        const float d1r = xr[1] - xr[4], d1i = xi[1] - xi[4], d2r = xr[2] - xr[3], d2i = xi[2] - xi[3];
        const float r1r = xr[0] + c1 * b1r + c2 * b2r, r1i = xi[0] + c1 * b1i + c2 * b2i;
        const float r2r = xr[0] + c2 * b1r + c1 * b2r, r2i = xi[0] + c2 * b1i + c1 * b2i;
        /* sign i (s1 d1 + s2 d2) and sign i (s2 d1 - s1 d2) */
        const float e1r = s1 * d1r + s2 * d2r, e1i = s1 * d1i + s2 * d2i;
        const float e2r = s2 * d1r - s1 * d2r, e2i = s2 * d1i - s1 * d2i;
        const float i1r = -sg * e1i, i1i = sg * e1r, i2r = -sg * e2i, i2i = sg * e2r;
        br[(size_t)o * B + b] = xr[0] + b1r + b2r; bi[(size_t)o * B + b] = xi[0] + b1i + b2i;
        br[(size_t)(o + Ns) * B + b] = r1r + i1r; bi[(size_t)(o + Ns) * B + b] = r1i + i1i;
        br[(size_t)(o + 4 * Ns) * B + b] = r1r - i1r; bi[(size_t)(o + 4 * Ns) * B + b] = r1i - i1i;
        br[(size_t)(o + 2 * Ns) * B + b] = r2r + i2r; bi[(size_t)(o + 2 * Ns) * B + b] = r2i + i2i;
        br[(size_t)(o + 3 * Ns) * B + b] = r2r - i2r; bi[(size_t)(o + 3 * Ns) * B + b] = r2i - i2i;
      }
    }
  }
}

/* Radix list of a length: 4s, one 2 for an odd power of two, then 3s and 5s; 0 when n
 * has another prime factor. */
static int fft_plan(int n, int *rad) {
  int m = n, e2 = 0, np = 0;
  while (m % 2 == 0) { m /= 2; e2++; }
  for (int i = 0; i < e2 / 2; i++) rad[np++] = 4;
  if (e2 % 2) rad[np++] = 2;
  while (m % 3 == 0) { m /= 3; rad[np++] = 3; }
  while (m % 5 == 0) { m /= 5; rad[np++] = 5; }
  return m == 1 ? np : 0;
}

/* Batched Stockham transform of x[n][B] (split arrays, B contiguous transforms);
 * y is scratch of the same size.  tw[m] = exp(-2 pi i m / n).  sign -1 forward,
 * +1 backward, unnormalised.  Returns 1 when the result is in y, 0 when in x. */
static int fft_batched(float *restrict xr, float *restrict xi, float *restrict yr, float *restrict yi, int n, int B,
                       const float *restrict twr, const float *restrict twi, int sign) {
  float *ar = xr, *ai = xi, *br = yr, *bi = yi;
  int Ns = 1, where = 0, rad[32];
  const int np = fft_plan(n, rad);
  const float sg = (float)sign;
  for (int ip = 0; ip < np; ip++) {
    const int R = rad[ip];
    if (R == 4) {
      const int q = n / 4, step = n / (4 * Ns);
      for (int j = 0; j < q; j++) {
        const int k = j % Ns;
        /* w^r = exp(sign 2 pi i r k / (4 Ns)) = tw[r k step]: tw holds exp(-2 pi i m/n), so the
         * imaginary part is tw_i for sign -1 and -tw_i for sign +1 */
        const float w1r = twr[k * step], w2r = twr[2 * k * step], w3r = twr[3 * k * step];
        const float v1i = sign < 0 ? twi[k * step] : -twi[k * step];
        const float v2i = sign < 0 ? twi[2 * k * step] : -twi[2 * k * step];
        const float v3i = sign < 0 ? twi[3 * k * step] : -twi[3 * k * step];
        const float *x0r = ar + (size_t)j * B, *x0i = ai + (size_t)j * B;
        const float *x1r = ar + (size_t)(j + q) * B, *x1i = ai + (size_t)(j + q) * B;
        const float *x2r = ar + (size_t)(j + 2 * q) * B, *x2i = ai + (size_t)(j + 2 * q) * B;
        const float *x3r = ar + (size_t)(j + 3 * q) * B, *x3i = ai + (size_t)(j + 3 * q) * B;
        const int o = (j / Ns) * 4 * Ns + k;
        float *y0r = br + (size_t)o * B, *y0i = bi + (size_t)o * B;
        float *y1r = br + (size_t)(o + Ns) * B, *y1i = bi + (size_t)(o + Ns) * B;
        float *y2r = br + (size_t)(o + 2 * Ns) * B, *y2i = bi + (size_t)(o + 2 * Ns) * B;
        float *y3r = br + (size_t)(o + 3 * Ns) * B, *y3i = bi + (size_t)(o + 3 * Ns) * B;
        for (int b = 0; b < B; b++) {
          const float a0r = x0r[b], a0i = x0i[b];
          const float a1r = x1r[b] * w1r - x1i[b] * v1i, a1i = x1r[b] * v1i + x1i[b] * w1r;
          const float a2r = x2r[b] * w2r - x2i[b] * v2i, a2i = x2r[b] * v2i + x2i[b] * w2r;
          const float a3r = x3r[b] * w3r - x3i[b] * v3i, a3i = x3r[b] * v3i + x3i[b] * w3r;
          const float t0r = a0r + a2r, t0i = a0i + a2i, t1r = a0r - a2r, t1i = a0i - a2i;
          const float t2r = a1r + a3r, t2i = a1i + a3i;
          /* (a1 - a3) * (sign i): sign -1 -> (r, i) -> (i, -r) */
          const float dr = a1r - a3r, di = a1i - a3i;
          const float t3r = -sg * di, t3i = sg * dr;
          y0r[b] = t0r + t2r; y0i[b] = t0i + t2i;
          y2r[b] = t0r - t2r; y2i[b] = t0i - t2i;
          y1r[b] = t1r + t3r; y1i[b] = t1i + t3i;
          y3r[b] = t1r - t3r; y3i[b] = t1i - t3i;
        }
      }
    } else if (R == 2) {
      const int h = n / 2, step = n / (2 * Ns);
      for (int j = 0; j < h; j++) {
        const int k = j % Ns;
        const float wr = twr[k * step], wi = sign < 0 ? twi[k * step] : -twi[k * step];
        const int o = (j / Ns) * 2 * Ns + k;
        const float *x0r = ar + (size_t)j * B, *x0i = ai + (size_t)j * B;
        const float *x1r = ar + (size_t)(j + h) * B, *x1i = ai + (size_t)(j + h) * B;
        float *y0r = br + (size_t)o * B, *y0i = bi + (size_t)o * B;
        float *y1r = br + (size_t)(o + Ns) * B, *y1i = bi + (size_t)(o + Ns) * B;
        for (int b = 0; b < B; b++) {
          const float ur = x1r[b] * wr - x1i[b] * wi, ui = x1r[b] * wi + x1i[b] * wr;
          y0r[b] = x0r[b] + ur; y0i[b] = x0i[b] + ui;
          y1r[b] = x0r[b] - ur; y1i[b] = x0i[b] - ui;
        }
      }
    } else {
      pass35(ar, ai, br, bi, n, B, R, Ns, twr, twi, sign);
    }
    float *tr = ar, *ti = ai;
    ar = br; ai = bi; br = tr; bi = ti;
    where ^= 1;
    Ns *= R;
  }
  return where;
}

typedef struct {
  int nx, ny, potNy, ndet;
  float *twxr, *twxi, *twyr, *twyi; /* exp(-2 pi i m / n) */
  float *pxr, *pxi, *pyr, *pyi;     /* propagator, 1/(nx ny) folded into px */
  float *kx2, *ky2;
  float k2max;
  float k2in[QO_MAX_DET], k2out[QO_MAX_DET];
  int sx[QO_MAX_DET], sy[QO_MAX_DET];
  double escale;
} setup32;

/* Threads the OpenMP runtime would use (the bench reports it). */
int qc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

static int supported(int n) {
  int rad[32];
  return n >= 16 && n <= 4096 && fft_plan(n, rad) > 0;
}

/* Per-thread buffers: the wave (split), and a block buffer with its scratch. */
typedef struct {
  float *wr, *wi;
  float *b0r, *b0i, *b1r, *b1i;
} work32;

/* Row pass over rows [x0, x0 + RB): optional inverse row FFT (previous slice; the
 * 1/(nx ny) is already in the wave), optional transmission multiply, optional forward
 * row FFT.  first: the wave is the probe (no inverse transform). */
static void row_pass(const setup32 *S, work32 *W, int x0, int RB, const float *trans, int posx, int posy, int inv,
                     int fwd) {
  const int ny = S->ny;
  float *br = W->b0r, *bi = W->b0i, *sr = W->b1r, *si = W->b1i;
  for (int y = 0; y < ny; y++) /* transposing gather: buf[y][b] = wave[x0 + b][y] */
    for (int b = 0; b < RB; b++) {
      br[y * RB + b] = W->wr[(size_t)(x0 + b) * ny + y];
      bi[y * RB + b] = W->wi[(size_t)(x0 + b) * ny + y];
    }
  if (inv && fft_batched(br, bi, sr, si, ny, RB, S->twyr, S->twyi, +1)) {
    float *t = br; br = sr; sr = t;
    t = bi; bi = si; si = t;
  }
  if (trans) { /* transmit: psi[x][y] *= t[x + posx][y + posy] (stemlib.cpp:2880-2887) */
    for (int y = 0; y < ny; y++)
      for (int b = 0; b < RB; b++) {
        const float *t = trans + 2 * ((size_t)(x0 + b + posx) * S->potNy + y + posy);
        const float vr = br[y * RB + b], vi = bi[y * RB + b];
        br[y * RB + b] = vr * t[0] - vi * t[1];
        bi[y * RB + b] = vr * t[1] + vi * t[0];
      }
  }
  if (fwd && fft_batched(br, bi, sr, si, ny, RB, S->twyr, S->twyi, -1)) {
    float *t = br; br = sr; sr = t;
    t = bi; bi = si; si = t;
  }
  for (int y = 0; y < ny; y++)
    for (int b = 0; b < RB; b++) {
      W->wr[(size_t)(x0 + b) * ny + y] = br[y * RB + b];
      W->wi[(size_t)(x0 + b) * ny + y] = bi[y * RB + b];
    }
}

/* Column pass over the CB columns [y0, y0 + CB): column FFT, propagator and band limit
 * (propagate_slow, stemlib.cpp:2831-2849), detector sums into acc (collectIntensity,
 * stemlib.cpp:2623-2659; nullable), column IFFT. */
static void col_pass(const setup32 *S, work32 *W, int y0, int CB, double *acc) {
  const int nx = S->nx, ny = S->ny;
  float *br = W->b0r, *bi = W->b0i, *sr = W->b1r, *si = W->b1i;
  for (int x = 0; x < nx; x++)
    for (int b = 0; b < CB; b++) {
      br[x * CB + b] = W->wr[(size_t)x * ny + y0 + b];
      bi[x * CB + b] = W->wi[(size_t)x * ny + y0 + b];
    }
  if (fft_batched(br, bi, sr, si, nx, CB, S->twxr, S->twxi, -1)) {
    float *t = br; br = sr; sr = t;
    t = bi; bi = si; si = t;
  }
  for (int x = 0; x < nx; x++) {
    const float kx2 = S->kx2[x];
    for (int b = 0; b < CB; b++) {
      const int y = y0 + b;
      float vr = br[x * CB + b], vi = bi[x * CB + b];
      if (kx2 < S->k2max && kx2 + S->ky2[y] < S->k2max) {
        const float pr = S->pxr[x] * S->pyr[y] - S->pxi[x] * S->pyi[y];
        const float pi = S->pxr[x] * S->pyi[y] + S->pxi[x] * S->pyr[y];
        const float tr = vr * pr - vi * pi;
        vi = vr * pi + vi * pr;
        vr = tr;
      } else {
        vr = vi = 0.f;
      }
      br[x * CB + b] = vr;
      bi[x * CB + b] = vi;
    }
  }
  if (acc) {
    for (int d = 0; d < S->ndet; d++) {
      double s = 0;
      for (int x = 0; x < nx; x++) {
        /* pixel (x, y) counts for detector d when k^2 of (x - sx, y - sy) is in its annulus */
        const float kx2 = S->kx2[((x - S->sx[d]) % nx + nx) % nx];
        float row = 0.f;
        for (int b = 0; b < CB; b++) {
          const float k2 = kx2 + S->ky2[((y0 + b - S->sy[d]) % ny + ny) % ny];
          const float v = br[x * CB + b] * br[x * CB + b] + bi[x * CB + b] * bi[x * CB + b];
          row += (k2 >= S->k2in[d] && k2 <= S->k2out[d]) ? v : 0.f;
        }
        s += row;
      }
      acc[d] += s;
    }
  }
  if (fft_batched(br, bi, sr, si, nx, CB, S->twxr, S->twxi, +1)) {
    float *t = br; br = sr; sr = t;
    t = bi; bi = si; si = t;
  }
  for (int x = 0; x < nx; x++)
    for (int b = 0; b < CB; b++) {
      W->wr[(size_t)x * ny + y0 + b] = br[x * CB + b];
      W->wi[(size_t)x * ny + y0 + b] = bi[x * CB + b];
    }
}

/* Multislice of nprobe probes over nslice slices, FP32.  trans: [nslice][potNx][potNy]
 * complex (interleaved); probe: [nx][ny] complex; det (nullable): [nbins][ndet][nprobe],
 * written for slices with bin_of_slice[s] >= 0; exit_wave (nullable): [nprobe][nx][ny]
 * complex.  nthreads <= 0: the OpenMP default.  Returns 0, or -1 for an unsupported size. */
int qc_multislice(const qo_params *p, const float *trans, int nslice, const float *probe, const int *ipos,
                  int nprobe, const int *bin_of_slice, int nbins, float *det, float *exit_wave, int nthreads) {
  const int nx = p->nx, ny = p->ny;
  if (!supported(nx) || !supported(ny) || p->ndet > QO_MAX_DET) return -1;
  (void)nbins;
  setup32 S;
  memset(&S, 0, sizeof S);
  S.nx = nx; S.ny = ny; S.potNy = p->potNy; S.ndet = p->ndet;
  float *tab = (float *)malloc(sizeof(float) * (size_t)(6 * nx + 6 * ny));
  S.twxr = tab; S.twxi = tab + nx; S.pxr = tab + 2 * nx; S.pxi = tab + 3 * nx; S.kx2 = tab + 4 * nx;
  S.twyr = tab + 6 * nx; S.twyi = S.twyr + ny; S.pyr = S.twyr + 2 * ny; S.pyi = S.twyr + 3 * ny;
  S.ky2 = S.twyr + 4 * ny;
  qo_k_tables(p, S.kx2, S.ky2, &S.k2max);
  qo_det_k2(p, S.k2in, S.k2out);
  for (int d = 0; d < p->ndet; d++) { S.sx[d] = (int)p->det_shiftx[d]; S.sy[d] = (int)p->det_shifty[d]; }
  const double wl = qo_wavelength(p->v0), dz = p->sliceThickness;
  const double ax = (double)p->resolutionX * nx, by = (double)p->resolutionY * ny, norm = 1.0 / ((double)nx * ny);
  for (int i = 0; i < nx; i++) {
    S.twxr[i] = (float)cos(-2 * M_PI * i / nx);
    S.twxi[i] = (float)sin(-2 * M_PI * i / nx);
    const double k = ((i > nx / 2) ? (double)(i - nx) : (double)i) / ax, t = M_PI * dz * wl * k * k;
    S.pxr[i] = (float)(cos(t) * norm);
    S.pxi[i] = (float)(-sin(t) * norm);
  }
  for (int i = 0; i < ny; i++) {
    S.twyr[i] = (float)cos(-2 * M_PI * i / ny);
    S.twyi[i] = (float)sin(-2 * M_PI * i / ny);
    const double k = ((i > ny / 2) ? (double)(i - ny) : (double)i) / by, t = M_PI * dz * wl * k * k;
    S.pyr[i] = (float)cos(t);
    S.pyi[i] = (float)(-sin(t));
  }
  S.escale = p->electronScale; /* the 1/(nx ny)^2 of collectIntensity is in px */
  const size_t n2 = (size_t)nx * ny;
  const int mx = nx > ny ? nx : ny, blk = (RB > CB ? RB : CB) * mx;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : qc_max_threads())
  {
    work32 W;
    float *mem = (float *)aligned_alloc(64, sizeof(float) * (2 * n2 + 4 * (size_t)blk));
    W.wr = mem; W.wi = mem + n2;
    W.b0r = mem + 2 * n2; W.b0i = W.b0r + blk; W.b1r = W.b0i + blk; W.b1i = W.b1r + blk;
    double acc[QO_MAX_DET];
#pragma omp for schedule(dynamic, 1)
    for (int ip = 0; ip < nprobe; ip++) {
      for (size_t k = 0; k < n2; k++) { W.wr[k] = probe[2 * k]; W.wi[k] = probe[2 * k + 1]; }
      const int posx = ipos[2 * ip], posy = ipos[2 * ip + 1];
      for (int s = 0; s < nslice; s++) {
        const float *t = trans + 2 * (size_t)s * p->potNx * p->potNy;
        for (int x0 = 0; x0 < nx; x0 += RB) row_pass(&S, &W, x0, nx - x0 < RB ? nx - x0 : RB, t, posx, posy, s > 0, 1);
        const int bin = (det && bin_of_slice) ? bin_of_slice[s] : -1;
        for (int d = 0; d < S.ndet; d++) acc[d] = 0;
        for (int y0 = 0; y0 < ny; y0 += CB) col_pass(&S, &W, y0, ny - y0 < CB ? ny - y0 : CB, bin >= 0 ? acc : NULL);
        if (bin >= 0)
          for (int d = 0; d < S.ndet; d++) det[((size_t)bin * S.ndet + d) * nprobe + ip] = (float)(acc[d] * S.escale);
      }
      /* inverse row transforms of the last slice: the exit wave */
      if (exit_wave) {
        for (int x0 = 0; x0 < nx; x0 += RB) row_pass(&S, &W, x0, nx - x0 < RB ? nx - x0 : RB, NULL, 0, 0, 1, 0);
        float *e = exit_wave + 2 * n2 * (size_t)ip;
        for (size_t k = 0; k < n2; k++) { e[2 * k] = W.wr[k]; e[2 * k + 1] = W.wi[k]; }
      }
    }
    free(mem);
  }
  free(tab);
  return 0;
}

