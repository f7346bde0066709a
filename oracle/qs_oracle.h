/*
 * qs_oracle.h -- FP64 CPU restatement of QSTEM's STEM multislice path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the FP64 reference the HIP product
 * path is compared with: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product library (libqstem_hip.so) never links or calls it.
 *
 * Every function cites the reference routine it restates (paths relative to
 * the QSTEM source tree).  Arithmetic is double precision throughout, except
 * the discrete decisions (band-limit aperture, detector annuli, scan-window
 * offsets) which reproduce the reference's single-precision comparisons so
 * that the same pixels are selected.
 *
 * Pinning status of each function is recorded in the header comment of
 * qs_oracle.c.
 */
#ifndef QS_ORACLE_H
#define QS_ORACLE_H

#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QO_MAX_DET 16

/* Mirrors the MULS fields that the STEM path reads (libs/data_containers.h:103-308).
 * Fields that MULS stores as float_tt are float here too, so the same rounded
 * values enter the arithmetic. */
typedef struct qo_params {
  int nx, ny;               /* wave grid */
  float resolutionX, resolutionY;
  float v0;                 /* kV */
  int slices;               /* slices per slab */
  float sliceThickness;     /* dz = cz[0] */
  int potNx, potNy;         /* transmission array size */
  float potSizeX, potSizeY; /* potNx*resolutionX, potNy*resolutionY */
  float by;                 /* model size in y (used by the Q1 quirk) */
  int compat_q1;            /* 1: initSTEMSlices uses muls->by for negative ky */
  int bandlimittrans;
  /* probe (stemlib.cpp:1952-2191) */
  float Cs, C5, df0, astigMag, astigAngle;
  float a33, a31, a44, a42, a55, a53, a51, a66, a64, a62;
  float phi33, phi31, phi44, phi42, phi55, phi53, phi51, phi66, phi64, phi62;
  float alpha;              /* mrad */
  float aAIS;               /* Angstrom */
  float Cc, dE_E;
  int ismoth, gaussFlag;
  float gaussScale;
  int compat_q2;            /* 1: a62 forced to 0 and C6 term always on (stemlib.cpp:2082) */
  /* detectors (stem3.cpp:1323-1354) */
  int ndet;
  float det_rin[QO_MAX_DET], det_rout[QO_MAX_DET];
  float det_shiftx[QO_MAX_DET], det_shifty[QO_MAX_DET];
  float electronScale;
} qo_params;

/* ---- numerics ---- */
void qo_fft1d(double complex *a, int n, int sign);
void qo_fft2d(double complex *a, int nx, int ny, int sign);
double qo_wavelength(double kev);
double qo_sigma(double kev);
void qo_splinh(const double *x, const double *y, double *b, double *c, double *d, int n);
double qo_seval(const double *x, const double *y, const double *b, const double *c,
                const double *d, int n, double x0);

/* ---- hot path pieces ---- */
void qo_probe(const qo_params *p, double complex *psi);
void qo_probe_d(const qo_params *p, double df0, double Cs, double complex *psi);
void qo_probe_at(const qo_params *p, double px, double py, double complex *psi);
void qo_k_tables(const qo_params *p, float *kx2, float *ky2, float *k2max);
void qo_det_k2(const qo_params *p, float *k2in, float *k2out);
void qo_make_trans(const qo_params *p, const float *V, int nslice, double complex *trans);
void qo_make_trans_d(const qo_params *p, const double *V, int nslice, double complex *trans);
void qo_scan_offsets(int ix, int iy, float scanXStart, float scanXStop, int scanXN,
                     float scanYStart, float scanYStop, int scanYN,
                     float resX, float resY, int potNx, int potNy, int nx, int ny,
                     int *iPosX, int *iPosY);
int qo_record_bins(int slices, int cellDiv, int outputInterval, int totalSliceCount,
                   int mulsRepeat1, int *bin_of_step);
/* Multislice over nslice slices for nprobe probes.  trans is [nslice][potNx][potNy].
 * psi_io is [nprobe][nx][ny]: incident waves in, exit waves out.
 * det (nullable) is [nbins][ndet][nprobe], written for slices whose bin_of_slice >= 0.
 * diff (nullable) is [nprobe][nx][ny] fftshifted |Psi|^2/sqrt(nx*ny) of the last slice. */
void qo_multislice(const qo_params *p, const double complex *trans, int nslice,
                   const int *ipos, int nprobe, double complex *psi_io,
                   const int *bin_of_slice, int nbins, double *det, double *diff);

/* FP64 adjoint of qo_multislice for L = sum (|DFT(psi_exit)/N^2|^2 - target)^2 (no reference
 * counterpart): loss, dL/dV [nslice][potNx][potNy], and g_psi0 = dL/d(conj psi0) summed over
 * probes [nx][ny] (nullable outputs). */
void qo_adjoint(const qo_params *p, const double *V, int nslice, const int *ipos, int nprobe,
                const double complex *psi0, const double *target, double *loss, double *gV,
                double complex *gpsi0);

/* ---- potential (stemlib.cpp:383-1541) ---- */
typedef struct qo_lut_geom { int nx, nz, nzPerSlice, Nr, Nz_lut; double dkx, dkz, smax2; } qo_lut_geom;
void qo_lut_geometry(float atomRadius, float resX, float sliceThickness, qo_lut_geom *g);
/* taper the shared Rez s-grid like the reference (stemlib.cpp:1390-1413); sgrid has 123 entries
 * and rows is [nrows][123]; returns nothing, modifies in place */
void qo_taper_scatpar(double *sgrid, double *rows, int nrows, double smax);
void qo_atom_lut(const qo_lut_geom *g, const double *sgrid, const double *frow, double B,
                 float sliceThickness, float *lut /* [Nz_lut][Nr] */);
void qo_atom_lut2(const qo_lut_geom *g, const double *sgrid, const double *frow, const double *sgrid_y,
                  const double *frow_y, double dky, double B, float sliceThickness, double scale, int clamp,
                  float *lut);
double qo_lut_norm_fix(const qo_lut_geom *g, float sliceThickness);
typedef struct qo_atom { float x, y, z, dw, occ, q; int Znum; } qo_atom;
/* 3D potential scatter for non-periodic (periodic=0) or periodic (1) xy; V is [slices][potNx][potNy].
 * luts[k] is the LUT for element kind k; kind_of_atom[i] indexes it. Atoms must be z-sorted. */
void qo_make_slices(const qo_params *p, const qo_lut_geom *g, const float *const *luts,
                    const int *kind_of_atom, const qo_atom *atoms, int natom,
                    float potOffsetX, float potOffsetY, float czOffset, int centerSlices,
                    float atomRadius, int periodic, float *V);

void qo_make_slices_q(const qo_params *p, const qo_lut_geom *g, const float *const *luts,
                      const float *const *offs_luts, const int *kind_of_atom, const qo_atom *atoms, int natom,
                      float potOffsetX, float potOffsetY, float czOffset, int centerSlices,
                      float atomRadius, int periodic, float *V);

void qo_make_slices_d(const qo_params *p, const qo_lut_geom *g, const float *const *luts,
                      const int *kind_of_atom, const double *xyz, const int *Znum, int natom,
                      double potOffsetX, double potOffsetY, double czOffset, int centerSlices,
                      double atomRadius, double dr, double *V);

#ifdef __cplusplus
}
#endif
#endif
