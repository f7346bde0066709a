// qs_internal.hpp -- shared declarations between the C-ABI layer and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace qs {

constexpr int kMaxDet = 16;

// A float in device global memory: pointers read from device arrays (the LUT
// table) are cast to it so their loads are global, not flat, loads.
using gfloat = __attribute__((address_space(1))) float;

// Arguments of both multislice sweeps (passed by value, lives in the kernarg segment).
struct SweepArgs {
  int nx, ny;          // wave grid (powers of two)
  int nprobe;          // probes in this launch
  int probe0;          // global index of the first probe of this launch
  int kc, kcp;         // band columns stored per row, and padded row length
  int band_a, band_b;  // band = iy in [0, band_a] U [ny-band_b, ny-1]
  int ntiles;          // column tiles per probe in sweep B
  int potNy;           // row pitch of the transmission slice
  int rw;              // sweep A: transform steps per wavefront (set by the launcher)
  int tpb;             // sweep B: column tiles per block (0: launcher default)
  const float2* probe;     // shared real-space probe [nx][ny] (sweep A mode 0)
  const float2* trans;     // transmission slice [potNx][potNy]
  const int2* ipos;        // window offset per global probe
  const int* cfg;          // nullable: frozen-phonon configuration per global probe
  long long cfg_stride;    // elements between consecutive configurations' transmission functions
  float2* state;           // [nprobe_total][nx][kcp]
  const float2* state_in = nullptr;  // sweep A reads the previous state here (if set) and writes a.state
  float2* out_wave;        // exit waves [nprobe_total][nx][ny] (sweep A mode 2)
  const float2* twx;       // forward twiddles W_nx^m
  const float2* twy;       // forward twiddles W_ny^m
  const float2* px;        // propagator x factor (includes 1/(nx ny))
  const float2* py;        // propagator y factor
  const float* kx2;        // single-precision k^2 tables as propagate_slow builds them
  const float* ky2;
  float k2max;
  int rec_bin;             // >= 0: record detector sums on this slice
  int ndet;
  int det_sx[kMaxDet], det_sy[kMaxDet];
  float det_k2in[kMaxDet], det_k2out[kMaxDet];
  float* partial;          // [nprobe][ntiles][kMaxDet+1]
  float* diff;             // nullable: fftshifted diffraction intensities [nprobe_total][nx][ny]
  float diff_scale;
  const float* vpot = nullptr;  // sweep A mode 4: projected potential [nprobe_total][nx][ny]
  float* gv = nullptr;          // sweep A mode 4: dL/dV out [nprobe_total][nx][ny]
  float gv_sigma = 0.f;         // sweep A mode 4: interaction constant sigma
};

// Backward sweeps of the adjoint (reverse of SweepArgs' forward sweeps).
struct AdjArgs {
  SweepArgs f;              // geometry, tables, trans slice, ipos, probe, twiddles
  float2* gx;               // adjoint state, same layout as the forward state
  const float2* xprev;      // forward state X_{s-1} (rows sweep, s > 0) or X_{S-1} (last column sweep)
  const float* target;      // [nprobe_total][nx][ny] target intensities, FFTW order (last column sweep)
  float* loss_partial;      // [nprobe][ntiles] (last column sweep)
  float* gT;                // [potNx][potNy][2] gradient w.r.t. the transmission slice (atomics)
  float* gprobe;            // [nx][ny][2] sum over probes of the gradient w.r.t. psi_0 (atomics)
  int single_writer;        // 1: one probe per launch, gT updated without atomics;
                            // 2: and its window is the whole slice, gT stored (no prior clear)
  // With single_writer 2 and the transmission band limit on band-compact sweeps: the
  // row sweep stores the row FFT of each dL/dt row, band-compact in the potential
  // grid's band (bl_band_a, bl_gap, rows of bl_kcp), into blstate instead of the row
  // into gT (the first pass of the band limit); nullptr: gT as above.
  float2* blstate;
  int bl_band_a, bl_gap, bl_kcp;
};

struct ProbeArgs {
  int nx, ny;
  double ax, by, rx, ry, rx2, ry2;
  double wavlen, k2max, pixel, scale;
  double df_eff, astigMag, astigAngle;
  double a33, a31, a44, a42, a55, a53, a51, a66, a64, a62;
  double phi33, phi31, phi44, phi42, phi55, phi53, phi51, phi66, phi64, phi62;
  double Cs, C5;
  int on3, on4, on5, on6;
  double sx, sy;  // probe centre terms (ax - nx/2*resX, by - ny/2*resY)
  int ismoth, gaussFlag;
  double gaussScale, aAIS, edge, resX, resY;
};

// LUT of one element (getAtomPotential3D, stem3/stemlib.cpp:1342-1541).
struct LutArgs {
  int nx, nz, nzPerSlice;
  double dkx, dkz, smax2;  // smax2 already squared
  double zPos;             // -2 pi dz/nzPerSlice * izOffset
  double dky;              // q_y step of the q_y sum (dkx for the neutral LUT)
  double scale;            // factor on the reference's 47.8658 dkx dkz / nz (1 = reference)
};

struct QsAtomDev {
  float x, y, z, dw, occ, q;
  int Znum;
};

// Scatter of atoms into slices (make3DSlices, stem3/stemlib.cpp:883-1213).
struct ScatterArgs {
  int potNx, potNy, slices;
  float dx, dy, sliceThickness;
  float potOffsetX, potOffsetY, czOffset, atomRadius;
  int centerSlices, periodic;
  int iRadX, iRadY, iRadZ;
  int Nr, Nz, nzSub;
  float dr;  // radial LUT step used for the lookup
};

void launch_lut_fk(const double* spl, const LutArgs& g, double B, float2* out, hipStream_t st);
void launch_lut_extract(const float2* tmp, const LutArgs& g, float* lut, hipStream_t st);
// Sum over atoms in gather form: per-atom reach, atoms binned by
// the kPotTile x kPotTile pixel tiles their box reaches (tile tx * tiles_y + ty) and
// the blocks of kPotZc slices they reach (bin tile * zblocks + zb), and one thread per
// pixel summing the kPotZc slices of its block.
constexpr int kPotTile = 16, kPotZc = 16;
struct AtomReach {
  float atomX, atomY;          // position relative to the potential array (Angstrom)
  int iax0, iax1, iay0, iay1;  // pixel box (inclusive)
  int iAtomZ, iaz0, iaz1;      // slices iAtomZ + [iaz0, iaz1]; iAtomZ = INT_MIN: reaches nothing
  int iOffsZ0;                 // LUT z offset at iaz0
  int kind;                    // LUT index
  float q;                     // ion charge: q times the kind's offset LUT is added
};
void launch_atom_reach(const QsAtomDev* atoms, int natom, const int* kind, const ScatterArgs& s, AtomReach* reach,
                       hipStream_t st);
void launch_bin_atoms(const AtomReach* reach, int natom, const ScatterArgs& s, int zblocks, int* count_or_cursor,
                      int* list, bool fill, hipStream_t st);
// oluts[kind]: ion-offset LUT of the kind, or null (no charged atom of that kind)
void launch_gather(const AtomReach* reach, const int* tstart, const int* list, const float* const* luts,
                   const float* const* oluts, const ScatterArgs& s, float* V, hipStream_t st);

// Multislice of non-power-of-two wave grids (qs_generic.hip): full-size state
// [probe][nx][ny], separate transmit / row-transform / transpose / propagate passes.
struct GenArgs {
  int nx, ny, nprobe, probe0, potNy;
  const float2* probe;  // [nx][ny] (first slice)
  const float2* trans;  // transmission slice [potNx][potNy] of configuration 0
  const int2* ipos;     // window offset per global probe
  const int* cfg;       // nullable: configuration per global probe
  long long cfg_stride;
  float2* state;        // [nprobe][nx][ny]
};
struct GenPropArgs {
  int nx, ny, probe0;
  float2* data;                // [nprobe][ny][nx] spectrum (transposed)
  const float2* px;            // propagator x factor (includes 1/(nx ny))
  const float2* py;
  const float* kx2;
  const float* ky2;
  float k2max;
  int rec, ndet;
  int det_sx[kMaxDet], det_sy[kMaxDet];
  float det_k2in[kMaxDet], det_k2out[kMaxDet];
  float* partial;              // [nprobe][gen_chunks(ny)][kMaxDet+1] when rec
  float* diff;                 // nullable: [nprobe_total][nx][ny]
  float diff_scale;
};
void launch_gen_transmit(const GenArgs& a, bool first, hipStream_t st);
int gen_chunks(int ny);
void launch_gen_propagate(const GenPropArgs& a, int nprobe, hipStream_t st);

// Mixed-radix fused path (sizes whose prime factors are 2, 3, 5, 7): the state is the
// y-spectrum [probe][x][ky], stored for the ky columns inside the 2/3 band only.
constexpr int kGenMaxPasses = 16;
struct GenPlan {
  int n, np;
  int r[kGenMaxPasses];  // radices of the Stockham passes (2, 3, 4, 5, 7, 8, 9, 16, 25), product n
};
struct GenRowArgs {
  int nx, ny, nprobe, probe0, potNy, rpb, mode;
  const float2* probe;  // [nx][ny] (mode 0)
  const float2* trans;  // transmission slice [potNx][potNy] of configuration 0
  const int2* ipos;     // window offset per global probe
  const int* cfg;       // nullable: configuration per global probe
  long long cfg_stride;
  float2* state;        // [nprobe_total][nx][ny] y-spectrum, in-band columns
  float2* exitw;        // mode 2: [nprobe_total][nx][ny] real-space waves
  const float* ky2;
  float k2max;
  GenPlan plan;         // length ny
  const float2* tw;     // W_ny^m, m < ny
};
struct GenColArgs {
  int nx, ny, nprobe, probe0, tc;
  float2* state;
  const int* tiles;     // first column of each tile
  const float2* px;     // propagator x factor (includes 1/(nx ny))
  const float2* py;
  const float* kx2;
  const float* ky2;
  float k2max;
  int rec, ndet;
  int det_sx[kMaxDet], det_sy[kMaxDet];
  float det_k2in[kMaxDet], det_k2out[kMaxDet];
  float* partial;       // [nprobe][ntiles][kMaxDet+1] when rec
  float* diff;          // nullable: [nprobe_total][nx][ny]
  float diff_scale;
  GenPlan plan;         // length nx
  const float2* tw;     // W_nx^m, m < nx
};
bool gen_plan(int n, GenPlan* pl);  // false: n has a prime factor above 7
// radix plan variants (0 = gen_plan's): 0 few passes (radix 16, 25, 9 first),
// 1 radix 8 for powers of two, 2 radices 2-8 only, 3 (kGenPlanCompiled) the
// wavefront-resident kernels with a compiled plan (k_mix_rows / k_mix_cols; the
// lengths of mix_compiled only; the GenPlan is then variant 0's, unused)
constexpr int kGenPlanVariants = 4, kGenPlanCompiled = 3;
bool gen_plan_variant(int n, int variant, GenPlan* pl);
// Lengths with a compiled radix plan (qs_mix.hpp; one translation unit qs_mix.hip per
// length, which the build reads from this list): those measured faster than the
// runtime-planned kernels (720, 800, 1200 and 1620 have plans too and were not).
#define QS_MIX_SIZES(X) \
  X(288) X(300) X(320) X(360) X(384) X(400) X(441) X(448) X(500) X(540) X(600) X(640) X(768) X(900) X(1280) \
  X(1350) X(1536) X(1800)
bool mix_compiled(int n);         // n has a compiled radix plan (QS_MIX_SIZES)
int mix_cols_per_block(int n);    // band columns per block of k_mix_cols<n>
// a.rpb: rows per transform slot (k_mix_rows, ny = a.ny)
void launch_mix_rows(const GenRowArgs& a, hipStream_t st);
// a.tiles: the in-band ky columns (a.tc of them), nblocks = ceil(a.tc / mix_cols_per_block(nx))
void launch_mix_cols(const GenColArgs& a, int nblocks, hipStream_t st);
int gen_rows_per_block(int ny);   // default rows per block of k_gen_rows (power of two)
int gen_cols_per_tile(int nx);    // default columns per tile of k_gen_cols (power of two)
int gen_max_vectors(int n);       // most vectors of length n one block holds in LDS
// Row transforms through the mixed-radix LDS passes (qs_generic.hip): length n
// directly (M == 0, plan of n) or Bluestein through a smooth length M (plan of M).
struct GenRowXformArgs {
  float2* data;            // [nrows][n]
  long nrows;
  int n, M, rpb;           // rpb: rows per block, a power of two (gen_rows_per_block of n or M)
  float scale;
  GenPlan plan;
  const float2* tw;        // W_L^m, L = M or n
  const float2* chirp;     // Bluestein: [n], exp(DIR i pi k^2/n)
  const float2* vhat;      // Bluestein: [M]
};
void launch_gen_rowxform(const GenRowXformArgs& a, int dir, hipStream_t st);
void launch_gen_rows(const GenRowArgs& a, hipStream_t st);
void launch_gen_cols(const GenColArgs& a, int ntiles, hipStream_t st);

bool supported_size(int n);
int col_tile(int nx);

void launch_sweep_rows(const SweepArgs& a, int mode, hipStream_t st);
void launch_sweep_cols(const SweepArgs& a, bool rec, bool diff, hipStream_t st);
void launch_reduce_partials(const float* partial, int nprobe, int ntiles, int ndet, float escale, float* det_bin,
                            int det_stride, float* total, int probe0, hipStream_t st);
void launch_rows_fft(float2* d, long nrows, int L, int dir, const float2* tw, float scale, hipStream_t st);
void launch_rows_bluestein(float2* d, long nrows, int n, int M, const float2* chirp, const float2* vhat,
                           const float2* tw, float scale, hipStream_t st);
void launch_transpose(const float2* src, float2* dst, int batch, int na, int nb, hipStream_t st);
void launch_phase_grating(const float* V, float2* trans, size_t n, double scale, hipStream_t st);
void launch_bandlimit_T(float2* d, int nslice, int nky, int nkx, const float* kx2, const float* ky2, float k2max,
                        float scale, hipStream_t st);
void launch_probe_kspace(const ProbeArgs& a, float2* out, hipStream_t st);
int probe_window_blocks(int n);
void launch_probe_window(const ProbeArgs& a, float2* psi, double* partial, hipStream_t st);
void launch_fill_zero(float2* p, size_t n, hipStream_t st);
hipError_t set_col_lds_limits();
void launch_adj_rows(const AdjArgs& a, bool first_slice, hipStream_t st);
void launch_adj_cols(const AdjArgs& a, bool last_slice, hipStream_t st);
void launch_grad_v(const float2* gu, const float* V, float* gV, size_t n, double sigma, hipStream_t st);
void launch_target_outside_band(const float* target, int nprobe, int nx, int ny, int band_a, int band_b,
                                double* out, hipStream_t st);
void launch_probe_grad(const ProbeArgs& a, const float2* G, double cscale, double wavlen_p, double* out2,
                       hipStream_t st);
void launch_atom_grad(const QsAtomDev* atoms, int natom, const float* const* luts, const float* const* oluts,
                      const int* kind, const ScatterArgs& s, const float* gV, float* grad, hipStream_t st);
void launch_cplx_fill_zero(float* p, size_t nfloats, hipStream_t st);

}  // namespace qs
