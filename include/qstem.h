/*
 * qstem.h -- C ABI of the MI355X multislice library (libqstem_hip.so).
 *
 * Drop-in boundary for QSTEM's STEM hot path.  The reference exposes it as
 * C++ free functions over a process-global MULS and per-thread WAVEFUNCs
 * (stem3/stemlib.h:36-62); this ABI replaces them with an opaque context
 * that owns the device buffers and two HIP streams (qs_set_streams).  Plain pointers and sizes
 * only; every call returns a status code instead of the reference's
 * printf+exit(0).
 *
 * Reference interface each entry point replaces:
 *   qs_create          readFile() MULS sizing + WAVEFUNC ctor      stem3/stem3.cpp:612-1559, libs/data_containers.cpp:24-58
 *   qs_set_potential   initSTEMSlices(MULS*, int nlayer)           stem3/stemlib.cpp:2199-2357
 *   qs_set_trans       direct write of muls.trans                  libs/data_containers.h:112-116
 *   qs_get_trans       read back of muls.trans                     stem3/stemlib.cpp:1280-1306 (save potential)
 *   qs_make_probe      probe(MULS*, WavePtr, double dx, double dy) stem3/stemlib.cpp:1952-2191
 *   qs_make_probe_at   probe() at any window position (doCBED)     stem3/stem3.cpp:2097
 *   qs_run_stem        doSTEM scan loop body: runMulsSTEM +        stem3/stem3.cpp:2820-2946,
 *                      transmit/propagate_slow/collectIntensity/   stem3/stemlib.cpp:2375-2534, 2576-2687,
 *                      fft_normalize for a batch of positions     2760-2906
 *   qs_build_potential make3DSlices + getAtomPotential3D            stem3/stemlib.cpp:383-1541
 *   qs_run_slices      runMulsSTEM over a sub-range of the slab    stem3/stemlib.cpp:2403-2482
 *   qs_record_bins     tCount/interval bookkeeping of collectIntensity stem3/stemlib.cpp:2588-2607
 *   qs_record_bins_repeat  the same with "sequence: r p" repeats    stem3/stemlib.cpp:2403, 2429
 *   qs_last_error      (none: the reference prints and exits)
 *
 * Threading: one context = one stream; calls on one context must not overlap.
 * Distinct contexts may be used from distinct threads.
 */
#ifndef QSTEM_H
#define QSTEM_H

#ifdef __cplusplus
extern "C" {
#endif

#define QS_OK 0
#define QS_ERR_ARG 1          /* invalid argument or inconsistent sizes */
#define QS_ERR_UNSUPPORTED 2  /* size or option the library does not handle */
#define QS_ERR_HIP 3          /* HIP runtime failure (message in qs_last_error) */
#define QS_ERR_STATE 4        /* call order violated (e.g. run before potential) */

#define QS_MAX_DET 16

/* Bit flags for qs_params.compat: reproduce reference quirks. */
#define QS_COMPAT_Q1 1 /* initSTEMSlices uses muls->by for negative ky (stemlib.cpp:2237) */
#define QS_COMPAT_Q2 2 /* probe(): "(a62 = 0)" assignment in the C6 condition (stemlib.cpp:2082) */

/* The MULS fields the STEM path reads (libs/data_containers.h:103-308).
 * Float where MULS stores float_tt, so the same rounded values are used. */
typedef struct qs_params {
  int nx, ny;                 /* wave grid: powers of two 64..4096 take the fused sweeps; any
                               * other size in 16..2048 (e.g. 400) an unfused path (qs_run_stem
                               * only: no adjoint, uneven batch, autotuner or no-allocation run) */
  float resolutionX, resolutionY;
  float v0;                   /* kV */
  int slices;                 /* slices per slab */
  float sliceThickness;       /* Angstrom (cz[0]) */
  int potNx, potNy;           /* transmission array, >= nx, ny */
  float potSizeX, potSizeY;   /* potNx*resolutionX, potNy*resolutionY */
  float by;                   /* model extent in y (only used with QS_COMPAT_Q1) */
  int bandlimittrans;         /* "bandlimit f_trans:" */
  float Cs, C5, df0, astigMag, astigAngle; /* Angstrom / rad, as stored in MULS */
  float a33, a31, a44, a42, a55, a53, a51, a66, a64, a62;
  float phi33, phi31, phi44, phi42, phi55, phi53, phi51, phi66, phi64, phi62;
  float alpha;                /* mrad */
  float aAIS;                 /* Angstrom, 0 = off */
  float Cc, dE_E;
  int ismoth, gaussFlag;
  float gaussScale;
  int ndet;                   /* <= QS_MAX_DET */
  float det_rin[QS_MAX_DET], det_rout[QS_MAX_DET]; /* mrad */
  float det_shiftx[QS_MAX_DET], det_shifty[QS_MAX_DET];
  float electronScale;        /* beamCurrent*dwellTime*MILLISEC_PICOAMP */
  int compat;                 /* QS_COMPAT_* bits */
} qs_params;

typedef struct qs_ctx qs_ctx;

/* Creates a context on the current HIP device.  On failure *out is NULL and
 * qs_last_error(NULL) describes the problem. */
int qs_create(const qs_params *p, qs_ctx **out);
void qs_destroy(qs_ctx *ctx);
/* Last error message of ctx, or of the calling thread's last failed qs_create when ctx is NULL. */
const char *qs_last_error(const qs_ctx *ctx);

/* Projected potential V[nslice][potNx][potNy] (float, host memory or device memory of
 * the context's device) -> transmission functions exp(i*gamma*lambda*V), band limited
 * when bandlimittrans is set.  A device V must be complete when the call is made (the
 * call does not wait for other streams). */
int qs_set_potential(qs_ctx *ctx, const float *V, int nslice);
/* Transmission functions directly, complex64 [nslice][potNx][potNy] (interleaved re,im). */
int qs_set_trans(qs_ctx *ctx, const float *trans, int nslice);
int qs_get_trans(qs_ctx *ctx, float *trans_out);

/* One atom as make3DSlices consumes it (libs/stemtypes_fftw3.h:69-77): Cartesian
 * position in Angstrom, Debye-Waller factor B (A^2), occupancy, charge, Z (0 = vacancy). */
typedef struct qs_atom {
  float x, y, z, dw, occ, q;
  int Znum;
} qs_atom;

/* The make3DSlices inputs beyond qs_params (stem3.cpp:827-1394). */
typedef struct qs_pot_params {
  float atomRadius;             /* "atom radius:" (default 5 A) */
  float potOffsetX, potOffsetY; /* scan_x_start - nx/2*resolutionX, likewise y (stem3.cpp:1436-1437) */
  float czOffset;               /* "zOffset:" */
  int centerSlices;             /* "center slices:" */
  int periodicXY;               /* "periodicXY:" */
  int tds;                      /* frozen phonons: B := 0 in the LUT (stemlib.cpp:914) */
  int compat_q6;                /* cache the LUT by Z only, first B in z order wins (stemlib.cpp:1422) */
  int lut_dr_fix;               /* 0: sample the LUT with dr = resolutionX/2 as make3DSlices does
                                 * (stemlib.cpp:576); the LUT is built with f_e(s) read at s = |k|
                                 * (a potential stretched 2x in r and z) on a radial spacing of
                                 * resolutionX, so this read restores the radial shape (Q10);
                                 * 1: sample at the LUT's own spacing (atoms 2x wider, 4x the
                                 * projected integral), a comparison variant */
  int compat_q5;                /* ion-offset LUT of charged atoms: 1 = the reference's q_y sum
                                 * over the neutral table's nodes with the ion table's spline
                                 * (stemlib.cpp:1672, quirk Q5); 0 = over the ion table */
  int lut_norm_fix;             /* 1: scale the LUTs by 4 L_z / 47.8658 (L_z = 2 ceil(R/dz) dz, the
                                 * LUT's z period) so that an atom's slice-summed projected
                                 * potential integrates to f_e(0) at every slice thickness; the
                                 * reference's 47.8658 dkx dkz / nz (stemlib.cpp:1516) gives
                                 * 47.8658 f_e(0) / (4 L_z) (with dr = resolutionX/2); 0 = reference */
} qs_pot_params;

/* Rez et al. electron scattering factors: 93 rows x 123 doubles, row 0 = s grid,
 * row Z = f_e(s) (stem3/scatfactsRez.h:22-25).  Required before qs_build_potential. */
int qs_set_scattering_table(qs_ctx *ctx, const double *table);
/* Ion-offset scattering factors df_e/dq, same layout (stem3/scatfactsRez.h:122,
 * scatParOffs).  Required before qs_build_potential with charged atoms. */
int qs_set_ion_scattering_table(qs_ctx *ctx, const double *table);

/* make3DSlices + getAtomPotential3D (+ getAtomPotentialOffset3D for charged atoms, the
 * reference's USE_Q_POT_OFFSETS build, stemlib.cpp:77, :1548-1741) + initSTEMSlices:
 * builds the projected potential of the given atoms into qs_params.slices slices on
 * the device, then the transmission functions.  3D potential only.  V_out (nullable)
 * receives V [slices][potNx][potNy].  Each voxel is the FP32 sum of the atoms' LUT
 * samples (plus q times the offset LUT's sample for charged atoms), gathered per
 * voxel from atoms binned by pixel tile, in an unspecified order. */
int qs_build_potential(qs_ctx *ctx, const qs_atom *atoms, int natom, const qs_pot_params *pp, float *V_out);

/* Replaces the probe optics of the context (Cs, C5, df0, astigmatism, a_nm/phi_nm,
 * alpha, aAIS, Cc, dE_E, smooth, gaussian) by those of p; other fields are ignored.
 * doSTEM changes dE_E between frozen-phonon runs (stem3.cpp:2750). */
int qs_set_optics(qs_ctx *ctx, const qs_params *p);

/* Forms the probe at the window centre, (nx/2)*resolutionX and (ny/2)*resolutionY in
 * float as doSTEM passes them (stem3.cpp:2838); keeps it on the device; copies it to
 * out (nx*ny complex64) when non-NULL. */
int qs_make_probe(qs_ctx *ctx, float *out);
/* qs_make_probe with the probe centred at (x, y) Angstrom from the window's corner:
 * doCBED places it at scan_x_start - potOffsetX plus the source-size offset
 * (stem3.cpp:2093-2097). */
int qs_make_probe_at(qs_ctx *ctx, double x, double y, float *out);

/* Bin bookkeeping of collectIntensity: for a slab of `slices` slices starting at
 * absolute slice totalSliceCount, bin_of_slice[s] receives the output-thickness
 * bin that slice s leaves as its final value (or -1).  Returns the number of bins
 * (tCount+1), or -1 on bad arguments. */
int qs_record_bins(int slices, int cellDiv, int outputInterval, int totalSliceCount, int *bin_of_slice);
/* qs_record_bins for a slab run mulsRepeat1 times ("sequence: r p", runMulsSTEM's
 * mRepeat loop): bin_of_step has slices*mulsRepeat1 entries, step k = repeat k/slices,
 * slice k%slices, recorded under the reference's slice index totalSliceCount +
 * (k%slices)*(1 + k/slices) (stemlib.cpp:2429).  Returns tCount+1, or -1. */
int qs_record_bins_repeat(int slices, int cellDiv, int outputInterval, int totalSliceCount, int mulsRepeat1,
                          int *bin_of_step);

/* Incident wave [nx][ny] complex64 (host) in place of the STEM probe, e.g. the
 * plane wave psi = 1 of TEM/CBED-style runs (the reference's doTEM initialises
 * the wave to 1 before its slice loop, stem3/stem3.cpp:2545-2672).  qs_run_stem
 * then starts every item from this wave; qs_adjoint gives no probe gradients. */
int qs_set_incident_wave(qs_ctx *ctx, const float *psi);

/* Device memory held by the library over all contexts, current and peak since
 * the last reset (reset_peak != 0 resets it to the current value).  No reference
 * counterpart. */
int qs_device_memory(long long *current_bytes, long long *peak_bytes, int reset_peak);

/* Runs nslice slices for nprobe scan positions.
 *   ipos          [nprobe][2] integer window offsets into the transmission array
 *   start         0: start from the probe; 1: continue the waves left by the previous call
 *   bin_of_slice  [nslice], -1 or the detector bin to record after that slice
 *   det           in/out [nbins][ndet][nprobe]; recorded bins are overwritten (nullable)
 *   diff          out [nprobe][nx][ny] fftshifted |Psi|^2/sqrt(nx*ny) after the last slice (nullable)
 *   exit_wave     out [nprobe][nx][ny] complex64 real-space waves after the last slice (nullable)
 *   int_intensity out [nprobe] sum|psi|^2/(nx*ny) of the exit waves (nullable)
 * All pointers are host pointers. */
int qs_run_stem(qs_ctx *ctx, const int *ipos, int nprobe, int start, const int *bin_of_slice, int nslice, int nbins,
                float *det, float *diff, float *exit_wave, float *int_intensity);

/* qs_run_stem without allocation: no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under HIP stream capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same results as qs_run_stem with
 * start = 0, configuration 0 and no diff or exit-wave output (det and int_intensity
 * as there, host pointers, nullable); the context's transmission functions and
 * probe must be set.  Every device buffer of the call (window offsets, wave state,
 * detector partial sums and outputs) is carved from a static device array of
 * QS_NOALLOC_SCRATCH_BYTES that each GPU holds until the process ends; a call that
 * needs more (about nprobe * nx * 2/3 ny * 8 bytes) fails with QS_ERR_ARG.  The call
 * calls no hipMalloc or hipFree and runs on the context's stream, one probe group
 * after another.  All contexts on one GPU share that GPU's array: calls that
 * overlap on one GPU race. */
#define QS_NOALLOC_SCRATCH_BYTES (64LL << 20)
int qs_run_stem_noalloc(qs_ctx *ctx, const int *ipos, int nprobe, const int *bin_of_slice, int nslice, int nbins,
                        float *det, float *int_intensity);

/* Frozen-phonon configurations as a batch axis.  The context holds nconfig
 * transmission functions (one per phonon configuration, same slices); the
 * potential setters (qs_set_potential, qs_set_trans, qs_build_potential) fill
 * the configuration chosen by qs_select_config, configuration 0 first (it fixes
 * the slice count).  qs_set_configs drops all loaded configurations.  Replaces
 * the reference's serial TDS loop of doSTEM (stem3/stem3.cpp:2769-2950), which
 * rebuilds one potential per run. */
int qs_set_configs(qs_ctx *ctx, int nconfig);
int qs_select_config(qs_ctx *ctx, int config);

/* qs_run_stem with item i = (window offset ipos[i], configuration config[i]);
 * config NULL means configuration 0 for every item. */
int qs_run_stem_cfg(qs_ctx *ctx, const int *ipos, const int *config, int nprobe, int start, const int *bin_of_slice,
                    int nslice, int nbins, float *det, float *diff, float *exit_wave, float *int_intensity);
/* qs_run_stem_cfg over transmission slices first_slice .. first_slice+nslice-1
 * (bin_of_slice[s] belongs to slice first_slice+s): a slab run in pieces, e.g. up to
 * each output thickness of TEM and CBED runs, which write the wave or diffraction
 * pattern there (interimWave, collectIntensity, stemlib.cpp:2476-2482, 2663-2676). */
int qs_run_slices(qs_ctx *ctx, const int *ipos, const int *config, int nprobe, int start, int first_slice,
                  const int *bin_of_slice, int nslice, int nbins, float *det, float *diff, float *exit_wave,
                  float *int_intensity);

/* Use src's transmission functions (all configurations) instead of dst's own:
 * contexts of different window sizes over one potential grid (same potNx/potNy,
 * resolution, voltage and band limit).  src must outlive dst's runs; src NULL
 * returns dst to its own.  No reference counterpart (one window size per run). */
int qs_share_trans(qs_ctx *dst, qs_ctx *src);

/* One window size of an uneven batch: a context (its nx, ny, probe, detectors)
 * and its items.  det [nbins][ndet][nitem] and int_intensity [nitem] as in
 * qs_run_stem (nullable). */
typedef struct qs_group {
  qs_ctx *ctx;
  const int *ipos;   /* [nitem][2] */
  const int *config; /* [nitem] or NULL */
  int nitem;
  float *det;
  float *int_intensity;
} qs_group;

/* Uneven batch: every group starts from its probe and advances through nslice
 * slices of the shared transmission functions; within each slice the groups
 * follow each other on the first group's stream, so a slice is read by every
 * window size while it is cache resident.  All contexts must share one set of
 * transmission functions (qs_share_trans); each group uses its own batch size
 * and launch tuning.  qs_last_run_ms of the first context reports the device
 * time.  No reference counterpart (the reference runs one window size per
 * program invocation). */
int qs_run_groups(const qs_group *groups, int ngroup, const int *bin_of_slice, int nslice, int nbins);

/* Adjoint of qs_run_stem for the loss L = sum_p sum_k (I_p(k) - target_p(k))^2, where
 * I_p(k) = |Psi_p(k) P(k)|^2/(nx ny)^2 is the diffraction intensity after the last slice
 * (fraction of the incident beam per pixel, FFTW order [nprobe][nx][ny]).  All
 * transmission slices are used, starting from the probe.  The forward waves are kept
 * every ckpt_every slices and recomputed inside each segment; ckpt_every = 0 lets the
 * library choose the smallest interval K whose ceil(nslice/K) + K wave states fit in
 * half the free device memory (K = 1: every slice kept, nothing recomputed), at most
 * ceil(sqrt(nslice)); qs_last_ckpt_interval reports the interval used.
 * Outputs (nullable): loss; grad_V [nslice][potNx][potNy] = dL/dV (needs the potential
 * set by qs_set_potential or qs_build_potential); grad_probe[2] = dL/d(df0), dL/d(Cs) in
 * the units of qs_params (Angstrom).  Device memory: ceil(nslice/K) + K wave states
 * (K = checkpoint interval), dL/dV of all slices, and dL/dt of one slice (all slices
 * when the transmission grid is band-limited but not a power of two; where one window
 * covers a band-limited power-of-two slice, instead two band-compact buffers of its
 * row transforms, the band limit of one slice running on the second stream of
 * qs_set_streams beside the reverse sweeps of the next).  No
 * reference counterpart (QSTEM has no adjoint). */
int qs_adjoint(qs_ctx *ctx, const int *ipos, int nprobe, const float *target, int ckpt_every, float *loss,
               float *grad_V, float *grad_probe);

/* qs_adjoint in two halves, for callers that evaluate the loss first and ask for
 * gradients later (a torch.autograd.Function's forward and backward).
 * qs_forward_save: the forward pass of qs_adjoint with its checkpoints, and the
 * loss; the context keeps them.  qs_adjoint_saved: the reverse pass from what the
 * last qs_forward_save kept (grad_V, grad_probe as in qs_adjoint); it consumes
 * them, and fails with QS_ERR_STATE when none are kept or when the potential,
 * probe, optics or configurations changed since (qs_run_stem in between is
 * allowed).  qs_last_run_ms reports each half's device time.  No reference
 * counterpart. */
int qs_forward_save(qs_ctx *ctx, const int *ipos, int nprobe, const float *target, int ckpt_every, float *loss);
/* (target of qs_forward_save and qs_adjoint, and grad_V of qs_adjoint_saved and
 * qs_adjoint, may be host memory or device memory of the context's device; both calls
 * have finished with them when they return.) */
/* Checkpoint interval of the last qs_adjoint / qs_forward_save (0 before the first).
 * No reference counterpart. */
int qs_last_ckpt_interval(const qs_ctx *ctx);
int qs_adjoint_saved(qs_ctx *ctx, float *grad_V, float *grad_probe);
/* dL/d(x, y, z) of the atoms of the last qs_build_potential (same atoms, same pp), from the
 * grad_V of the last qs_adjoint; z is 0 (the potential samples z by nearest slice).  Fails
 * with QS_ERR_ARG when the atoms' count, Z or charge or pp differ from that build, and with
 * QS_ERR_STATE when the potential was set otherwise since. */
int qs_atom_gradients(qs_ctx *ctx, const qs_atom *atoms, int natom, const qs_pot_params *pp, float *grad_xyz);

/* Probes advanced together through all slices before the next group starts
 * (0 = library default).  batch = 1 is the per-probe launch sequence. */
int qs_set_batch(qs_ctx *ctx, int batch);
int qs_get_batch(const qs_ctx *ctx);

/* Streams of qs_run_stem and qs_run_groups: 2 (default) lets consecutive probe
 * groups (qs_run_groups: contexts, by the setting of the first group's context)
 * run on two HIP streams, so that one group's launches fill the other's ramp-up
 * and drain; 1 keeps every launch on the context stream.  Each call waits for
 * both streams before it returns.  Results do not depend on it.  No reference
 * counterpart. */
int qs_set_streams(qs_ctx *ctx, int nstreams);
int qs_get_streams(const qs_ctx *ctx);

/* Replay the slice loop of qs_run_stem as a HIP graph: captured on the first call
 * of a given shape (probe count, slices, bins, outputs, buffers) and relaunched
 * with one hipGraphLaunch while the shape repeats; any setter above drops it.
 * Per-launch profiling (qs_set_profiling) bypasses the graph.  Replaces the
 * per-pixel runMulsSTEM calls of the scan loop (stem3/stem3.cpp:2820-2946);
 * the reference has no launch graph. */
int qs_set_graph(qs_ctx *ctx, int on);

/* Launch shape of the two sweeps (0 = library default): row-sweep steps per
 * wavefront (a power of two) and column tiles per block; qs_get_tuning also
 * reports the probe group size of qs_set_batch.  For grids on the mixed-radix path
 * (not powers of two, prime factors 2, 3, 5, 7) the two values are the rows per block
 * of the row kernel and the columns per tile of the column kernel (powers of two up
 * to 16 that fit the block's LDS; QS_ERR_ARG otherwise); where a length has a compiled
 * plan (288, 300, 320, 360, 384, 400, 441, 448, 500, 540, 600, 640, 768, 900, 1280, 1350,
 * 1536, 1800:
 * wavefront-resident kernels, the default unless QSTEM_COMPILED_PLANS=0 is set when
 * the context is created) the first is the rows per transform slot of the row kernel
 * and the second is unused.  No reference counterpart. */
int qs_set_tuning(qs_ctx *ctx, int rows_per_wave, int tiles_per_block);
int qs_get_tuning(const qs_ctx *ctx, int *rows_per_wave, int *tiles_per_block, int *batch);

/* Autotuner: times the slice loop on this context's transmission function and
 * probe (first 8 slices, up to 512 of the given window offsets) over probe group
 * sizes, row-sweep steps per wavefront and column tiles per block, applies the
 * fastest (as qs_set_batch + qs_set_tuning) and reports it (outputs nullable;
 * rate in probe-slices/s).  On the mixed-radix path it tunes the group size, rows
 * per block, columns per tile and the radix plan of each axis (few large radices,
 * radix 8, radices up to 8, or the compiled plan where the length has one: results
 * then differ by FP32 rounding only).  Results are cached per device model and wave geometry
 * for the life of the process, and across processes in the JSON file named by the
 * environment variable QSTEM_TUNE_CACHE when it is set (read at the first call,
 * rewritten after each new measurement; no file is touched when it is unset).  No
 * reference counterpart. */
int qs_autotune(qs_ctx *ctx, const int *ipos, int nprobe, int *batch, int *rows_per_wave, int *tiles_per_block,
                float *rate);

/* Device time of the last qs_run_stem in milliseconds (HIP events on the context stream). */
float qs_last_run_ms(const qs_ctx *ctx);

/* Per-launch timing of the row sweep (transmit + row transforms) and the
 * column sweep (column transforms + propagate + detect): when on, each launch
 * is bracketed by HIP events on the context stream and the durations are
 * accumulated until the next qs_set_profiling call.  No reference counterpart
 * (the reference only prints coarse CPU timers, stem3.cpp:2937-2945). */
int qs_set_profiling(qs_ctx *ctx, int on);
int qs_kernel_times(const qs_ctx *ctx, double *ms_rows, double *ms_cols, long *n_rows, long *n_cols);

#ifdef __cplusplus
}
#endif
#endif
