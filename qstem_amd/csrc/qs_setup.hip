// qs_setup.hip -- set-up kernels (probe, transmission function, transposes),
// detector reduction, and size dispatch for the templated sweeps.
#include <hip/hip_runtime.h>

#include "qs_fft.hpp"
#include "qs_internal.hpp"

namespace qs {

constexpr int kBlockS = 256;

// Sum the per-tile partials (double) into det[bin][d][probe] and the total.
__global__ void k_reduce_partials(const float* __restrict__ partial, int nprobe, int ntiles, int ndet,
                                  float escale, float* __restrict__ det_bin, int det_stride,
                                  float* __restrict__ total, int probe0) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nprobe) return;
  for (int d = 0; d <= ndet; d++) {
    const int dd = d < ndet ? d : kMaxDet;
    double s = 0;
    for (int i = 0; i < ntiles; i++) s += partial[((size_t)p * ntiles + i) * (kMaxDet + 1) + dd];
    if (d < ndet) det_bin[(size_t)d * det_stride + probe0 + p] = (float)(s * escale);
    else if (total) total[probe0 + p] = (float)s;
  }
}

// Batched transpose [batch][a][b] -> [batch][b][a].
__global__ void k_transpose(const float2* __restrict__ src, float2* __restrict__ dst, int na, int nb) {
  __shared__ float2 tilebuf[32][33];
  const size_t off = (size_t)blockIdx.z * na * nb;
  const int b0 = blockIdx.x * 32, a0 = blockIdx.y * 32;
  for (int i = threadIdx.y; i < 32; i += 8) {
    const int ia = a0 + i, ib = b0 + threadIdx.x;
    if (ia < na && ib < nb) tilebuf[i][threadIdx.x] = src[off + (size_t)ia * nb + ib];
  }
  __syncthreads();
  for (int i = threadIdx.y; i < 32; i += 8) {
    const int ib = b0 + i, ia = a0 + threadIdx.x;
    if (ia < na && ib < nb) dst[off + (size_t)ib * na + ia] = tilebuf[threadIdx.x][i];
  }
}

// ============================================================================
// Transmission function (initSTEMSlices, stemlib.cpp:2280-2336).
__global__ void k_phase_grating(const float* __restrict__ V, float2* __restrict__ trans, size_t n, double scale) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double vz = (double)V[i] * scale;
    double s, c;
    sincos(vz, &s, &c);
    trans[i] = make_float2((float)c, (float)s);
  }
}

// Band limit in the transposed layout [slice][ky][kx]; scale 1/(potNx potNy).
__global__ void k_bandlimit_T(float2* __restrict__ data, int nslice, int nky, int nkx, const float* __restrict__ kx2,
                              const float* __restrict__ ky2, float k2max, float scale) {
  const size_t n = (size_t)nslice * nky * nkx;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int kx = (int)(i % nkx);
    const int ky = (int)((i / nkx) % nky);
    const float k2 = ky2[ky] + kx2[kx];
    data[i] = k2 < k2max ? cscale(data[i], scale) : make_float2(0.f, 0.f);
  }
}

// ============================================================================
// Probe (stemlib.cpp:1952-2191): chi(k) in FP64, aperture, in FFTW order.
__global__ void k_probe_kspace(ProbeArgs a, float2* __restrict__ out) {
  const int n = a.nx * a.ny;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int ix = i / a.ny, iy = i % a.ny;
    const double ky = (iy > a.ny / 2) ? (double)(iy - a.ny) : (double)iy;
    const double kx = (ix > a.nx / 2) ? (double)(ix - a.nx) : (double)ix;
    const double k2 = kx * kx * a.rx2 + ky * ky * a.ry2;
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
    double amp = 0.0;
    if (a.ismoth && fabs(k2 - a.k2max) <= a.pixel) amp = 0.5 * a.scale;
    else if (k2 <= a.k2max) amp = a.scale;
    double s, c;
    sincos(chi, &s, &c);
    out[i] = make_float2((float)(amp * c), (float)(-amp * s));
  }
}

// Gaussian and AIS real-space windows (stemlib.cpp:2118-2148), then |psi|^2 partial sums.
__global__ void k_probe_window(ProbeArgs a, float2* __restrict__ psi, double* __restrict__ partial) {
  __shared__ double sh[kBlockS];
  const int n = a.nx * a.ny;
  double acc = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int ix = i / a.ny, iy = i % a.ny;
    float2 v = psi[i];
    if (a.gaussFlag) {
      const double r = exp(-((double)((ix - a.nx / 2) * (ix - a.nx / 2) + (iy - a.ny / 2) * (iy - a.ny / 2))) /
                           ((double)a.nx * a.nx * a.gaussScale));
      v = make_float2((float)(v.x * r), (float)(v.y * r));
    }
    if (a.aAIS > 0) {
      const double x = ix * a.resX - a.sx, y = iy * a.resY - a.sy;
      const double r = sqrt(x * x + y * y);
      const double dd = r - 0.5 * a.aAIS + a.edge;
      if (dd > 0) v = make_float2(0.f, 0.f);
      else if (dd >= -a.edge) {
        const double sc = 0.5 * (1 - cos(M_PI * dd / a.edge));
        v = make_float2((float)(sc * v.x), (float)(sc * v.y));
      }
    }
    psi[i] = v;
    acc += (double)v.x * v.x + (double)v.y * v.y;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void k_probe_normalise(float2* __restrict__ psi, int n, const double* __restrict__ partial, int nparts) {
  __shared__ double total;
  if (threadIdx.x == 0) {
    double s = 0;
    for (int i = 0; i < nparts; i++) s += partial[i];
    total = sqrt((double)n / s);
  }
  __syncthreads();
  const double sc = total;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float2 v = psi[i];
    psi[i] = make_float2((float)(v.x * sc), (float)(v.y * sc));
  }
}

__global__ void k_fill_zero(float2* __restrict__ p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = make_float2(0.f, 0.f);
}


// ============================================================================
// Dispatch on the power-of-two sizes (definitions live in qs_sizes.hip).
#define QS_DECL(N)                                                                                    \
  void sweep_rows_##N(const SweepArgs&, int, hipStream_t);                                          \
  void sweep_cols_##N(const SweepArgs&, bool, bool, hipStream_t);                                   \
  hipError_t set_col_lds_##N();                                                                     \
  int col_tile_##N();                                                                               \
  void rows_fft_##N(float2*, long, int, const float2*, float, hipStream_t);                         \
  void bluestein_##N(float2*, long, int, const float2*, const float2*, const float2*, float, hipStream_t); \
  void adj_rows_##N(const AdjArgs&, bool, hipStream_t);                                             \
  void adj_cols_##N(const AdjArgs&, bool, hipStream_t);
QS_DECL(64)
QS_DECL(128)
QS_DECL(256)
QS_DECL(512)
QS_DECL(1024)
QS_DECL(2048)
QS_DECL(4096)
#undef QS_DECL

#define QS_SWITCH(n, CALL)      \
  switch (n) {                  \
    case 64: CALL(64); break;   \
    case 128: CALL(128); break; \
    case 256: CALL(256); break; \
    case 512: CALL(512); break; \
    case 1024: CALL(1024); break; \
    case 2048: CALL(2048); break; \
    case 4096: CALL(4096); break; \
    default: break;             \
  }

bool supported_size(int n) { return n >= 64 && n <= 4096 && (n & (n - 1)) == 0; }

int col_tile(int nx) {
  int r = 0;
#define C(N) r = col_tile_##N()
  QS_SWITCH(nx, C)
#undef C
  return r;
}

void launch_sweep_rows(const SweepArgs& a, int mode, hipStream_t st) {
#define C(N) sweep_rows_##N(a, mode, st)
  QS_SWITCH(a.ny, C)
#undef C
}

void launch_sweep_cols(const SweepArgs& a, bool rec, bool diff, hipStream_t st) {
#define C(N) sweep_cols_##N(a, rec, diff, st)
  QS_SWITCH(a.nx, C)
#undef C
}

hipError_t set_col_lds_limits() {
  hipError_t e = hipSuccess, r;
#define C(N) if ((r = set_col_lds_##N()) != hipSuccess) e = r
  C(64); C(128); C(256); C(512); C(1024); C(2048); C(4096);
#undef C
  return e;
}

void launch_rows_fft(float2* d, long nrows, int L, int dir, const float2* tw, float scale, hipStream_t st) {
#define C(N) rows_fft_##N(d, nrows, dir, tw, scale, st)
  QS_SWITCH(L, C)
#undef C
}

void launch_rows_bluestein(float2* d, long nrows, int n, int M, const float2* chirp, const float2* vhat,
                           const float2* tw, float scale, hipStream_t st) {
#define C(N) bluestein_##N(d, nrows, n, chirp, vhat, tw, scale, st)
  QS_SWITCH(M, C)
#undef C
}

void launch_adj_rows(const AdjArgs& a, bool first_slice, hipStream_t st) {
#define C(N) adj_rows_##N(a, first_slice, st)
  QS_SWITCH(a.f.ny, C)
#undef C
}

void launch_adj_cols(const AdjArgs& a, bool last_slice, hipStream_t st) {
#define C(N) adj_cols_##N(a, last_slice, st)
  QS_SWITCH(a.f.nx, C)
#undef C
}

void launch_reduce_partials(const float* partial, int nprobe, int ntiles, int ndet, float escale, float* det_bin,
                            int det_stride, float* total, int probe0, hipStream_t st) {
  hipLaunchKernelGGL(k_reduce_partials, dim3((nprobe + 127) / 128), dim3(128), 0, st, partial, nprobe, ntiles, ndet,
                     escale, det_bin, det_stride, total, probe0);
}

void launch_transpose(const float2* src, float2* dst, int batch, int na, int nb, hipStream_t st) {
  dim3 grid((nb + 31) / 32, (na + 31) / 32, batch);
  hipLaunchKernelGGL(k_transpose, grid, dim3(32, 8), 0, st, src, dst, na, nb);
}

static unsigned grid_for(size_t n) {
  size_t g = (n + kBlockS - 1) / kBlockS;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (unsigned)g;
}

void launch_phase_grating(const float* V, float2* trans, size_t n, double scale, hipStream_t st) {
  hipLaunchKernelGGL(k_phase_grating, dim3(grid_for(n)), dim3(kBlockS), 0, st, V, trans, n, scale);
}

void launch_bandlimit_T(float2* d, int nslice, int nky, int nkx, const float* kx2, const float* ky2, float k2max,
                        float scale, hipStream_t st) {
  const size_t n = (size_t)nslice * nky * nkx;
  hipLaunchKernelGGL(k_bandlimit_T, dim3(grid_for(n)), dim3(kBlockS), 0, st, d, nslice, nky, nkx, kx2, ky2, k2max, scale);
}

void launch_probe_kspace(const ProbeArgs& a, float2* out, hipStream_t st) {
  hipLaunchKernelGGL(k_probe_kspace, dim3(grid_for((size_t)a.nx * a.ny)), dim3(kBlockS), 0, st, a, out);
}

int probe_window_blocks(int n) { return (int)grid_for((size_t)n) > 1024 ? 1024 : (int)grid_for((size_t)n); }

void launch_probe_window(const ProbeArgs& a, float2* psi, double* partial, hipStream_t st) {
  const int nb = probe_window_blocks(a.nx * a.ny);
  hipLaunchKernelGGL(k_probe_window, dim3(nb), dim3(kBlockS), 0, st, a, psi, partial);
  hipLaunchKernelGGL(k_probe_normalise, dim3(nb), dim3(kBlockS), 0, st, psi, a.nx * a.ny, partial, nb);
}

void launch_fill_zero(float2* p, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(n)), dim3(kBlockS), 0, st, p, n);
}

}  // namespace qs
