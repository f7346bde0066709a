// qs_adjoint.hip -- gradient assembly kernels of the adjoint multislice.
//
// The forward model is the STEM path of runMulsSTEM (stem3/stemlib.cpp:2375-2534)
// with t_s = BL(exp(i sigma V_s)) (initSTEMSlices, :2280-2336), the probe of
// probe() (:1952-2191) and the potential of make3DSlices (:883-1213).  The
// backward sweeps live in qs_sweep.hpp; these kernels turn their outputs into
// dL/dV, dL/d(defocus, Cs) and dL/d(atom x, y).
#include <hip/hip_runtime.h>

#include "qs_internal.hpp"

namespace qs {

// dL/dV = sigma Im(g_u conj(u)), u = exp(i sigma V)   (dt = i sigma t dV)
__global__ void k_grad_v(const float2* __restrict__ gu, const float* __restrict__ V, float* __restrict__ gV, size_t n,
                         double sigma) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    double s, c;
    sincos((double)V[i] * sigma, &s, &c);
    const float2 g = gu[i];
    // Im(g * conj(u)) = g.y*c - g.x*s
    gV[i] = (float)(sigma * ((double)g.y * c - (double)g.x * s));
  }
}

// sum of target^2 over the ky columns outside the band (their intensity is 0)
__global__ void k_target_outside(const float* __restrict__ target, int nprobe, int nx, int ny, int band_a, int band_b,
                                  double* __restrict__ out) {
  __shared__ double sh[256];
  double acc = 0;
  const size_t n = (size_t)nprobe * nx * ny;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int iy = (int)(i % ny);
    if (iy > band_a && iy < ny - band_b) acc += (double)target[i] * target[i];
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, sh[0]);
}

// dL/d(df0), dL/dCs from G = FFT(sum_p g_psi0) (times the normalisation c):
// dL/dchi(k) = Im(conj(G) Psi0) with Psi0 = amp e^{-i chi};
// dchi/ddf = pi lambda k^2, dchi/dCs = pi lambda^3 k^4 / 2 (probe()'s small-angle chi).
__global__ void k_probe_grad(ProbeArgs a, const float2* __restrict__ G, double cscale, double* __restrict__ out) {
  __shared__ double s0[256], s1[256];
  double acc0 = 0, acc1 = 0;
  const int n = a.nx * a.ny;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int ix = i / a.ny, iy = i % a.ny;
    const double ky = (iy > a.ny / 2) ? (double)(iy - a.ny) : (double)iy;
    const double kx = (ix > a.nx / 2) ? (double)(ix - a.nx) : (double)ix;
    const double k2 = kx * kx * a.rx2 + ky * ky * a.ry2;
    double amp = 0.0;
    if (a.ismoth && fabs(k2 - a.k2max) <= a.pixel) amp = 0.5 * a.scale;
    else if (k2 <= a.k2max) amp = a.scale;
    if (amp == 0.0) continue;
    const double kt2 = k2 * a.wavlen * a.wavlen;
    const double kt = sqrt(kt2);
    const double phi = atan2(a.ry * ky, a.rx * kx);
    double chi = kt2 * (a.df_eff + a.astigMag * cos(2.0 * (phi - a.astigAngle))) / 2.0;
    double ktp = kt2 * kt;
    if (a.on3) chi += ktp * (a.a33 * cos(3.0 * (phi - a.phi33)) + a.a31 * cos(phi - a.phi31)) / 3.0;
    ktp *= kt;
    if (a.on4) chi += ktp * (a.a44 * cos(4.0 * (phi - a.phi44)) + a.a42 * cos(2.0 * (phi - a.phi42)) + a.Cs) / 4.0;
    ktp *= kt;
    if (a.on5)
      chi += ktp * (a.a55 * cos(5.0 * (phi - a.phi55)) + a.a53 * cos(3.0 * (phi - a.phi53)) + a.a51 * cos(phi - a.phi51)) / 5.0;
    ktp *= kt;
    if (a.on6)
      chi += ktp * (a.a66 * cos(6.0 * (phi - a.phi66)) + a.a64 * cos(4.0 * (phi - a.phi64)) + a.a62 * cos(2.0 * (phi - a.phi62)) + a.C5) / 6.0;
    chi *= 2.0 * M_PI / a.wavlen;
    chi -= 2.0 * M_PI * ((a.sx * kx / a.ax) + (a.sy * ky / a.by));
    double s, c;
    sincos(chi, &s, &c);
    const double pr = amp * c, pi = -amp * s;  // Psi0
    const double gr = cscale * G[i].x, gi = cscale * G[i].y;
    const double dchi = gr * pi - gi * pr;  // Im(conj(G) Psi0)
    acc0 += dchi * M_PI * a.wavlen * k2;
    acc1 += dchi * 0.5 * M_PI * a.wavlen * a.wavlen * a.wavlen * k2 * k2;
  }
  s0[threadIdx.x] = acc0;
  s1[threadIdx.x] = acc1;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      s0[threadIdx.x] += s0[threadIdx.x + st];
      s1[threadIdx.x] += s1[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(out, s0[0]);
    atomicAdd(out + 1, s1[0]);
  }
}

// dL/d(atom x, y): chain rule through the r-linear LUT lookup of the scatter
// (the z lookup is nearest-sample, piecewise constant, so dL/dz = 0).
__global__ __launch_bounds__(256) void k_atom_grad(const QsAtomDev* __restrict__ atoms, int natom,
                                                   const float* const* __restrict__ luts,
                                                   const float* const* __restrict__ oluts, const int* __restrict__ kind,
                                                   ScatterArgs s, const float* __restrict__ gV, float* __restrict__ grad) {
  __shared__ double shx[256], shy[256];
  const int ia = blockIdx.x;
  double ax = 0, ay = 0;
  const QsAtomDev at = atoms[ia];
  bool active = at.Znum != 0;
  const float dzs = s.sliceThickness;
  const float c = dzs * s.slices;
  float atomZ = (float)((double)(at.z + s.czOffset) - (0.5 * (double)dzs * (double)(1 - s.centerSlices)));
  atomZ -= dzs;
  if (atomZ - s.atomRadius > c || atomZ + s.atomRadius + dzs < 0) active = false;
  atomZ += dzs;
  const float atomX = at.x - s.potOffsetX, atomY = at.y - s.potOffsetY;
  const int iAtomX = (int)floor(atomX / s.dx), iAtomY = (int)floor(atomY / s.dy);
  const int iAtomZ = (int)floor(atomZ / dzs + 0.5);
  int iax0, iax1, iay0, iay1;
  if (!s.periodic) {
    iax0 = iAtomX - s.iRadX < 0 ? 0 : iAtomX - s.iRadX;
    iax1 = iAtomX + s.iRadX >= s.potNx ? s.potNx - 1 : iAtomX + s.iRadX;
    iay0 = iAtomY - s.iRadY < 0 ? 0 : iAtomY - s.iRadY;
    iay1 = iAtomY + s.iRadY >= s.potNy ? s.potNy - 1 : iAtomY + s.iRadY;
    if (!(iax0 < s.potNx && iax1 >= 0 && iay0 < s.potNy && iay1 >= 0)) active = false;
  } else {
    iax0 = iAtomX - s.iRadX; iax1 = iAtomX + s.iRadX - 1;
    iay0 = iAtomY - s.iRadY; iay1 = iAtomY + s.iRadY - 1;
  }
  const int iaz0 = iAtomZ - s.iRadZ < 0 ? -iAtomZ : -s.iRadZ;
  const int iaz1 = iAtomZ + s.iRadZ >= s.slices ? s.slices - iAtomZ - 1 : s.iRadZ;
  if (!(iAtomZ + iaz0 < s.slices && iAtomZ + iaz1 >= 0)) active = false;
  if (active) {
    const gfloat* lut = (const gfloat*)luts[kind[ia]];
    const gfloat* olut = (at.q != 0.f && oluts) ? (const gfloat*)oluts[kind[ia]] : nullptr;
    const double q = at.q;
    const float zrel = atomZ / dzs;
    const double dOffsZ = (double)((float)((float)(iAtomZ + iaz0) - zrel) * (float)s.nzSub);
    const int iOffsZ0 = (int)(dOffsZ + 0.5) * s.Nr;
    const int limHi = s.Nr * (s.Nz - 1), limLo = -s.Nr * (s.Nz - 1), step = s.nzSub * s.Nr;
    const int nbx = iax1 - iax0 + 1, nby = iay1 - iay0 + 1;
    const double dr = (double)s.dr;
    const size_t slice = (size_t)s.potNx * s.potNy;
    for (int pidx = threadIdx.x; pidx < nbx * nby; pidx += blockDim.x) {
      const int iax = iax0 + pidx / nby, iay = iay0 + pidx % nby;
      const double xd = (float)(iax * s.dx - atomX);
      const double yd = (float)(iay * s.dy - atomY);
      const double r = sqrt(xd * xd + yd * yd);
      const double ddr = r / dr;
      const int ir = (int)floor(ddr);
      if (ir >= s.Nr - 1 || r <= 0.0) continue;
      const int ixw = s.periodic ? (((iax % s.potNx) + s.potNx) % s.potNx) : iax;
      const int iyw = s.periodic ? (((iay % s.potNy) + s.potNy) % s.potNy) : iay;
      const float* g = gV + (size_t)(iAtomZ + iaz0) * slice + (size_t)ixw * s.potNy + iyw;
      int iOffsZ = iOffsZ0;
      double dVdr_g = 0;
      for (int iaz = iaz0; iaz <= iaz1; iaz++) {
        double slope = 0;
        if (iOffsZ < 0) {
          if (iOffsZ > limLo) {
            slope = (lut[ir + 1 - iOffsZ + s.Nr] - lut[ir - iOffsZ + s.Nr]) / dr;
            if (olut) slope += q * (olut[ir + 1 - iOffsZ + s.Nr] - olut[ir - iOffsZ + s.Nr]) / dr;
          }
        } else if (iOffsZ < limHi) {
          slope = (lut[ir + 1 + iOffsZ] - lut[ir + iOffsZ]) / dr;
          if (olut) slope += q * (olut[ir + 1 + iOffsZ] - olut[ir + iOffsZ]) / dr;
        }
        dVdr_g += slope * (double)(*g);
        g += slice;
        iOffsZ += step;
      }
      // r = |p - atom|: dr/dx_atom = -(p_x - atom_x)/r
      ax -= dVdr_g * xd / r;
      ay -= dVdr_g * yd / r;
    }
  }
  shx[threadIdx.x] = ax;
  shy[threadIdx.x] = ay;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      shx[threadIdx.x] += shx[threadIdx.x + st];
      shy[threadIdx.x] += shy[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    grad[3 * ia] = (float)shx[0];
    grad[3 * ia + 1] = (float)shy[0];
    grad[3 * ia + 2] = 0.f;
  }
}

__global__ void k_fill_zero_f(float* __restrict__ p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.f;
}

static unsigned gridn(size_t n) {
  size_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  return (unsigned)(g ? g : 1);
}

void launch_grad_v(const float2* gu, const float* V, float* gV, size_t n, double sigma, hipStream_t st) {
  hipLaunchKernelGGL(k_grad_v, dim3(gridn(n)), dim3(256), 0, st, gu, V, gV, n, sigma);
}

void launch_target_outside_band(const float* target, int nprobe, int nx, int ny, int band_a, int band_b,
                                double* out, hipStream_t st) {
  const size_t n = (size_t)nprobe * nx * ny;
  hipLaunchKernelGGL(k_target_outside, dim3(gridn(n) > 4096 ? 4096 : gridn(n)), dim3(256), 0, st, target, nprobe, nx,
                     ny, band_a, band_b, out);
}

void launch_probe_grad(const ProbeArgs& a, const float2* G, double cscale, double wavlen_p, double* out2,
                       hipStream_t st) {
  (void)wavlen_p;
  const size_t n = (size_t)a.nx * a.ny;
  hipLaunchKernelGGL(k_probe_grad, dim3(gridn(n) > 1024 ? 1024 : gridn(n)), dim3(256), 0, st, a, G, cscale, out2);
}

void launch_atom_grad(const QsAtomDev* atoms, int natom, const float* const* luts, const float* const* oluts,
                      const int* kind, const ScatterArgs& s, const float* gV, float* grad, hipStream_t st) {
  if (natom > 0)
    hipLaunchKernelGGL(k_atom_grad, dim3(natom), dim3(256), 0, st, atoms, natom, luts, oluts, kind, s, gV, grad);
}

void launch_cplx_fill_zero(float* p, size_t nfloats, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_zero_f, dim3(gridn(nfloats)), dim3(256), 0, st, p, nfloats);
}

}  // namespace qs
