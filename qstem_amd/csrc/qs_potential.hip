// qs_potential.hip -- projected-potential slices on the GPU.
//
// Restates make3DSlices (stem3/stemlib.cpp:383-1334, fftpotential +
// potential3D branches, Z_INTERPOLATION 0) and getAtomPotential3D
// (stem3/stemlib.cpp:1342-1541):
//   1. per element, f_e(k) from an Akima spline of the Rez table times the
//      Debye-Waller factor, summed over q_y (k_lut_fk, FP64),
//   2. a 2D inverse DFT of the (kz, kx) map (Stockham/Bluestein row kernels),
//   3. z-integration over nzPerSlice sub-samples, clamp >= 0, scale (k_lut_extract),
//   4. r-linear / z-nearest LUT samples of every atom summed into the slices it
//      reaches, gathered per voxel from atoms binned by pixel tile and block of
//      slices (k_atom_reach, k_bin_atoms, k_gather): one store per voxel.
#include <hip/hip_runtime.h>

#include <climits>

#include "qs_internal.hpp"

namespace qs {

constexpr int kNsf = 123;  // nodes of the Rez s-grid (stem3/scatfactsRez.h:23)

__device__ double seval_dev(const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ b,
                            const double* __restrict__ c, const double* __restrict__ d, double x0) {
  // stem3/stemutil.cpp:1069-1092: binary search, end intervals clamp
  int i;
  if (x0 <= x[0]) i = 0;
  else if (x0 >= x[kNsf - 2]) i = kNsf - 2;
  else {
    int lo = 0, hi = kNsf;
    while (hi - lo > 1) {
      const int k = (lo + hi) / 2;
      if (x0 < x[k]) hi = k; else lo = k;
    }
    i = lo;
  }
  const double z = x0 - x[i];
  return y[i] + (b[i] + (c[i] + d[i] * z) * z) * z;
}

// f(kx,kz) with the q_y sum (stemlib.cpp:1447-1480, :1652-1680); spl = [x, y, b, c, d,
// xq, yq] each kNsf doubles: the spline (x, y, b, c, d) evaluates the (kx, kz) term and
// its coefficients with the nodes (xq, yq) the q_y sum.  (xq, yq) = (x, y) for the
// neutral LUT; for the ion-offset LUT the reference passes the neutral table's nodes
// with the offset table's coefficients (quirk Q5, stemlib.cpp:1672).
__global__ void k_lut_fk(const double* __restrict__ spl, LutArgs g, double B, float2* __restrict__ out) {
  const int n = g.nx * g.nz;
  const double* sx = spl;
  const double* sy = spl + kNsf;
  const double* sb = spl + 2 * kNsf;
  const double* sc = spl + 3 * kNsf;
  const double* sd = spl + 4 * kNsf;
  const double* qx = spl + 5 * kNsf;
  const double* qy = spl + 6 * kNsf;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int iz = i / g.nx, ix = i % g.nx;
    const double kz = g.dkz * (iz < g.nz / 2 ? iz : iz - g.nz);
    const double kx = g.dkx * (ix < g.nx / 2 ? ix : ix - g.nx);
    const double s2 = kx * kx + kz * kz;
    float2 v = make_float2(0.f, 0.f);
    if (s2 < g.smax2) {
      double f = seval_dev(sx, sy, sb, sc, sd, sqrt(s2)) * exp(-s2 * B * 0.25);
      for (int iy = 1; iy < g.nx; iy++) {
        double s3 = g.dky * iy;
        s3 = s3 * s3 + s2;
        if (s3 < g.smax2) f += 2.0 * seval_dev(qx, qy, sb, sc, sd, sqrt(s3)) * exp(-s3 * B * 0.25);
        else break;
      }
      f *= g.dkx;
      const double phase = kz * g.zPos;
      double sn, cs;
      sincos(phase, &sn, &cs);
      v = make_float2((float)(f * cs), (float)(f * sn));
    }
    out[i] = v;
  }
}

// z-integration and scale (stemlib.cpp:1503-1522).  tmp is the real-space map [nz][nx].
__global__ void k_lut_extract(const float2* __restrict__ tmp, LutArgs g, float* __restrict__ lut) {
  const int Nr = g.nx / 2, Nz = g.nz / 2;
  const int izOff = (g.nzPerSlice - 1) / 2;
  const double scale = 47.8658 * g.dkx * g.dkz / g.nz * g.scale;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < Nr * Nz; i += gridDim.x * blockDim.x) {
    const int iz = i / Nr, ix = i % Nr;
    double zs = 0;
    for (int iiz = -izOff; iiz <= izOff; iiz++)
      if (iz + izOff + iiz < Nz) zs += tmp[ix + (size_t)(iz + izOff + iiz) * g.nx].x;
    if (zs < 0) zs = 0;
    lut[ix + (size_t)iz * Nr] = (float)(scale * zs);
  }
}

// One block per atom (stemlib.cpp:883-1213).  Float arithmetic follows the
// reference statement by statement where it decides sample indices.
// Per-atom set-up of getAtomPotential3D: the pixel box, the slices reached and the
// z offset into the LUT; false when the atom reaches no pixel or slice.
__device__ __forceinline__ bool atom_setup(const QsAtomDev& at, const ScatterArgs& s, AtomReach& r) {
  if (at.Znum == 0) return false;  // vacancy
  const float dzs = s.sliceThickness;
  const float c = dzs * s.slices;
  // :886-889, then :945 (the -dz and +dz are kept to reproduce float rounding)
  float atomZ = (float)((double)(at.z + s.czOffset) - (0.5 * (double)dzs * (double)(1 - s.centerSlices)));
  atomZ -= dzs;
  if (atomZ - s.atomRadius > c) return false;
  if (atomZ + s.atomRadius + dzs < 0) return false;
  atomZ += dzs;
  r.atomX = at.x - s.potOffsetX;
  r.atomY = at.y - s.potOffsetY;
  const int iAtomX = (int)floor(r.atomX / s.dx), iAtomY = (int)floor(r.atomY / s.dy);
  r.iAtomZ = (int)floor(atomZ / dzs + 0.5);
  if (!s.periodic) {
    r.iax0 = iAtomX - s.iRadX < 0 ? 0 : iAtomX - s.iRadX;
    r.iax1 = iAtomX + s.iRadX >= s.potNx ? s.potNx - 1 : iAtomX + s.iRadX;
    r.iay0 = iAtomY - s.iRadY < 0 ? 0 : iAtomY - s.iRadY;
    r.iay1 = iAtomY + s.iRadY >= s.potNy ? s.potNy - 1 : iAtomY + s.iRadY;
    if (!(r.iax0 < s.potNx && r.iax1 >= 0 && r.iay0 < s.potNy && r.iay1 >= 0)) return false;
  } else {  // :1106-1109, upper bounds exclusive
    r.iax0 = iAtomX - s.iRadX; r.iax1 = iAtomX + s.iRadX - 1;
    r.iay0 = iAtomY - s.iRadY; r.iay1 = iAtomY + s.iRadY - 1;
  }
  r.iaz0 = r.iAtomZ - s.iRadZ < 0 ? -r.iAtomZ : -s.iRadZ;
  r.iaz1 = r.iAtomZ + s.iRadZ >= s.slices ? s.slices - r.iAtomZ - 1 : s.iRadZ;
  if (!(r.iAtomZ + r.iaz0 < s.slices && r.iAtomZ + r.iaz1 >= 0)) return false;
  const float zrel = atomZ / dzs;
  const double dOffsZ = (double)((float)((float)(r.iAtomZ + r.iaz0) - zrel) * (float)s.nzSub);
  r.iOffsZ0 = (int)(dOffsZ + 0.5) * s.Nr;
  return true;
}

__device__ __forceinline__ int pmod(int a, int n) { return ((a % n) + n) % n; }

// LUT sample of pixel distance bin (ir, ddr) at z offset iOffsZ (:1170-1190).
__device__ __forceinline__ double lut_sample(const gfloat* lut, const ScatterArgs& s, int ir, double ddr, int iOffsZ) {
  const int limHi = s.Nr * (s.Nz - 1), limLo = -s.Nr * (s.Nz - 1);
  double pv = 0;
  if (iOffsZ < 0) {
    if (iOffsZ > limLo) pv = (1 - ddr) * lut[ir - iOffsZ + s.Nr] + ddr * lut[ir + 1 - iOffsZ + s.Nr];
  } else if (iOffsZ < limHi) {
    pv = (1 - ddr) * lut[ir + iOffsZ] + ddr * lut[ir + 1 + iOffsZ];
  }
  return pv;
}

// ---------------------------------------------------------------------------------
// Gather form of getAtomPotential3D's additions: atoms binned by the kPotTile x
// kPotTile pixel tiles their box reaches (wrapped for periodic boxes) and the blocks
// of kPotZc slices they reach; one thread per pixel of a tile sums, for the kPotZc
// slices of its block in registers, the LUT samples of every atom (every image of it
// inside a periodic box) and stores each voxel once.

// Set-up of every atom; reach[ia].iAtomZ = INT_MIN marks atoms that reach nothing.
__global__ __launch_bounds__(256) void k_atom_reach(const QsAtomDev* __restrict__ atoms, int natom,
                                                    const int* __restrict__ kind, ScatterArgs s,
                                                    AtomReach* __restrict__ reach) {
  const int ia = blockIdx.x * blockDim.x + threadIdx.x;
  if (ia >= natom) return;
  AtomReach r;
  if (!atom_setup(atoms[ia], s, r)) r.iAtomZ = INT_MIN;
  r.kind = kind[ia];
  r.q = atoms[ia].q;
  reach[ia] = r;
}

// Per bin (pixel tile, block of kPotZc slices), the number of atoms that reach it
// (fill = false), or the atom indices at the bin's next free list positions (fill =
// true).  Bin ((tx * tiles_y + ty) * zblocks + zb).
// Tiles of a grid of n pixels reached by the pixel box [lo, hi] (periodic boxes may
// leave [0, n) and are wrapped): up to two disjoint ranges t0[i]..t1[i]; returns how many.
__device__ __forceinline__ int tile_ranges(int lo, int hi, int n, int* t0, int* t1) {
  const int tiles = (n + kPotTile - 1) / kPotTile;
  if (hi - lo + 1 >= n) {
    t0[0] = 0; t1[0] = tiles - 1;
    return 1;
  }
  const int a = ((lo % n) + n) % n, b = ((hi % n) + n) % n;
  if (a <= b) {
    t0[0] = a / kPotTile; t1[0] = b / kPotTile;
    return 1;
  }
  if (a / kPotTile <= b / kPotTile) {  // the two wrapped pieces share a tile
    t0[0] = 0; t1[0] = tiles - 1;
    return 1;
  }
  t0[0] = 0; t1[0] = b / kPotTile;
  t0[1] = a / kPotTile; t1[1] = tiles - 1;
  return 2;
}

__global__ __launch_bounds__(256) void k_bin_atoms(const AtomReach* __restrict__ reach, int natom, ScatterArgs s,
                                                   int tiles_y, int zblocks, int* __restrict__ count_or_cursor,
                                                   int* __restrict__ list, bool fill) {
  const int ia = blockIdx.x * blockDim.x + threadIdx.x;
  if (ia >= natom) return;
  const AtomReach r = reach[ia];
  if (r.iAtomZ == INT_MIN) return;
  const int zb0 = (r.iAtomZ + r.iaz0) / kPotZc, zb1 = (r.iAtomZ + r.iaz1) / kPotZc;
  int x0[2], x1[2], y0[2], y1[2];
  const int nxr = tile_ranges(r.iax0, r.iax1, s.potNx, x0, x1), nyr = tile_ranges(r.iay0, r.iay1, s.potNy, y0, y1);
  for (int i = 0; i < nxr; i++)
    for (int tx = x0[i]; tx <= x1[i]; tx++)
      for (int j = 0; j < nyr; j++)
        for (int ty = y0[j]; ty <= y1[j]; ty++)
          for (int zb = zb0; zb <= zb1; zb++) {
            const int pos = atomicAdd(count_or_cursor + (tx * tiles_y + ty) * zblocks + zb, 1);
            if (fill) list[pos] = ia;
          }
}

template <bool PERIODIC>
__global__ __launch_bounds__(kPotTile * kPotTile) void k_gather(const AtomReach* __restrict__ reach,
                                                               const int* __restrict__ tstart,
                                                               const int* __restrict__ list,
                                                               const float* const* __restrict__ luts,
                                                               const float* const* __restrict__ oluts, ScatterArgs s,
                                                               int tiles_y, float* __restrict__ V) {
  const int tile = blockIdx.x;  // tile (tx, ty) = tx * tiles_y + ty
  const int ix = (tile / tiles_y) * kPotTile + threadIdx.x / kPotTile;
  const int iy = (tile % tiles_y) * kPotTile + threadIdx.x % kPotTile;
  const int z0 = blockIdx.y * kPotZc;
  const int bin = tile * gridDim.y + blockIdx.y;
  __shared__ AtomReach sh[kPotTile * kPotTile];  // the bin's atoms, a block-sized batch at a time
  const int step = s.nzSub * s.Nr;
  const double dr = (double)s.dr;
  float acc[kPotZc];
#pragma unroll
  for (int zz = 0; zz < kPotZc; zz++) acc[zz] = 0.f;
  const int b = tstart[bin], e = tstart[bin + 1];
  for (int k0 = b; k0 < e; k0 += kPotTile * kPotTile) {
    const int n = min(kPotTile * kPotTile, e - k0);
    __syncthreads();  // the previous batch is read
    if ((int)threadIdx.x < n) sh[threadIdx.x] = reach[list[k0 + threadIdx.x]];
    __syncthreads();
    for (int k = 0; k < n; k++) {
      const AtomReach r = sh[k];  // the same atom in every lane
      const gfloat* lut = (const gfloat*)luts[r.kind];
      // ion charge: q times the offset LUT's sample at the same (r, z) (stemlib.cpp:966-1006)
      const gfloat* olut = (r.q != 0.f && oluts) ? (const gfloat*)oluts[r.kind] : nullptr;
      // samples of box pixel (ax, ay) (unwrapped coordinates) into this pixel's sums
      auto add = [&](int ax, int ay) __attribute__((always_inline)) {
        double x2 = (float)(ax * s.dx - r.atomX);
        x2 *= x2;
        double y2 = (float)(ay * s.dy - r.atomY);
        y2 *= y2;
        const double rr = sqrt(x2 + y2);
        double ddr = rr / dr;
        const int ir = (int)floor(ddr);
        if (ir >= s.Nr - 1) return;
        ddr -= ir;
#pragma unroll
        for (int zz = 0; zz < kPotZc; zz++) {
          const int iaz = z0 + zz - r.iAtomZ;
          if (iaz < r.iaz0 || iaz > r.iaz1) continue;
          const int iOffsZ = r.iOffsZ0 + (iaz - r.iaz0) * step;
          double pv = lut_sample(lut, s, ir, ddr, iOffsZ);
          if (olut) pv += (double)r.q * lut_sample(olut, s, ir, ddr, iOffsZ);
          acc[zz] += (float)pv;
        }
      };
      if constexpr (!PERIODIC) {  // the pixel itself, when inside the box
        if (ix >= r.iax0 && ix <= r.iax1 && iy >= r.iay0 && iy <= r.iay1) add(ix, iy);
      } else {  // every image ix + i*potNx, iy + j*potNy inside the box
        for (int ax = r.iax0 + pmod(ix - r.iax0, s.potNx); ax <= r.iax1; ax += s.potNx)
          for (int ay = r.iay0 + pmod(iy - r.iay0, s.potNy); ay <= r.iay1; ay += s.potNy) add(ax, ay);
      }
    }
  }
  if (ix >= s.potNx || iy >= s.potNy) return;
  const size_t slice = (size_t)s.potNx * s.potNy;
#pragma unroll
  for (int zz = 0; zz < kPotZc; zz++)
    if (z0 + zz < s.slices) V[(size_t)(z0 + zz) * slice + (size_t)ix * s.potNy + iy] = acc[zz];
}

static unsigned gridn(size_t n) {
  size_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  return (unsigned)(g ? g : 1);
}

void launch_lut_fk(const double* spl, const LutArgs& g, double B, float2* out, hipStream_t st) {
  hipLaunchKernelGGL(k_lut_fk, dim3(gridn((size_t)g.nx * g.nz)), dim3(256), 0, st, spl, g, B, out);
}

void launch_lut_extract(const float2* tmp, const LutArgs& g, float* lut, hipStream_t st) {
  hipLaunchKernelGGL(k_lut_extract, dim3(gridn((size_t)(g.nx / 2) * (g.nz / 2))), dim3(256), 0, st, tmp, g, lut);
}

void launch_atom_reach(const QsAtomDev* atoms, int natom, const int* kind, const ScatterArgs& s, AtomReach* reach,
                       hipStream_t st) {
  if (natom > 0) hipLaunchKernelGGL(k_atom_reach, dim3((natom + 255) / 256), dim3(256), 0, st, atoms, natom, kind, s, reach);
}

void launch_bin_atoms(const AtomReach* reach, int natom, const ScatterArgs& s, int zblocks, int* count_or_cursor,
                      int* list, bool fill, hipStream_t st) {
  const int tiles_y = (s.potNy + kPotTile - 1) / kPotTile;
  if (natom > 0)
    hipLaunchKernelGGL(k_bin_atoms, dim3((natom + 255) / 256), dim3(256), 0, st, reach, natom, s, tiles_y, zblocks,
                       count_or_cursor, list, fill);
}

void launch_gather(const AtomReach* reach, const int* tstart, const int* list, const float* const* luts,
                   const float* const* oluts, const ScatterArgs& s, float* V, hipStream_t st) {
  const int tiles_x = (s.potNx + kPotTile - 1) / kPotTile, tiles_y = (s.potNy + kPotTile - 1) / kPotTile;
  const dim3 grid((unsigned)(tiles_x * tiles_y), (unsigned)((s.slices + kPotZc - 1) / kPotZc));
  if (s.periodic)
    hipLaunchKernelGGL(k_gather<true>, grid, dim3(kPotTile * kPotTile), 0, st, reach, tstart, list, luts, oluts, s,
                       tiles_y, V);
  else
    hipLaunchKernelGGL(k_gather<false>, grid, dim3(kPotTile * kPotTile), 0, st, reach, tstart, list, luts, oluts, s,
                       tiles_y, V);
}

}  // namespace qs
