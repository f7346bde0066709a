// qs_generic.hip -- multislice of wave grids that are not powers of two.
//
// The fused two-sweep kernels (qs_sweep.hpp) need power-of-two radix plans.  For
// any other nx x ny (up to 2048 per axis, e.g. the 400 x 400 of the reference's
// stem_STO_4x4.qsc) each slice of runMulsSTEM (stem3/stemlib.cpp:2405-2443) runs
// as separate passes over a full-size state [probe][nx][ny]:
//   transmit (k_gen_transmit), row FFTs of length ny, transpose, row FFTs of
//   length nx (Stockham or Bluestein row kernels), propagator and detector sums
//   on the transposed layout (k_gen_propagate), and the inverse the same way.
// The 1/(nx ny) of fft_normalize is folded into the x propagator factor, as in
// the fused sweeps.
#include <hip/hip_runtime.h>

#include "qs_fft.hpp"
#include "qs_internal.hpp"

namespace qs {

// psi[p][x][y] = (first ? probe[x][y] : psi[p][x][y]) * trans_cfg(p)[x + ipos.x][y + ipos.y]
// (transmit, stemlib.cpp:2880-2887).
__global__ __launch_bounds__(256) void k_gen_transmit(GenArgs a, int first) {
  const size_t n2 = (size_t)a.nx * a.ny;
  const size_t n = (size_t)a.nprobe * n2;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int p = (int)(i / n2);
    const size_t rem = i - (size_t)p * n2;
    const int x = (int)(rem / a.ny), y = (int)(rem % a.ny);
    const int pg = a.probe0 + p;
    const int2 ip = a.ipos[pg];
    const float2* tr = a.trans + (a.cfg ? (long long)a.cfg[pg] * a.cfg_stride : 0LL);
    const float2 t = tr[(size_t)(x + ip.x) * a.potNy + y + ip.y];
    const float2 v = first ? a.probe[rem] : a.state[i];
    a.state[i] = cmul(v, t);
  }
}

// On the transposed spectrum [p][ky][kx] of nprobe probes: x propagator x band limit
// (propagate_slow, stemlib.cpp:2831-2849), and when rec, detector sums of the
// propagated |Psi|^2 (collectIntensity, stemlib.cpp:2623-2659; slot kMaxDet is the
// total) per block of RPC ky rows, into partial[p][chunk][kMaxDet + 1]; when diff,
// the fftshifted |Psi|^2/sqrt(nx ny) of the unnormalised Psi into diff[probe0 + p].
constexpr int kGenRows = 4;  // ky rows per block
__global__ __launch_bounds__(256) void k_gen_propagate(GenPropArgs a) {
  const int p = blockIdx.y, chunk = blockIdx.x;
  const int ky0 = chunk * kGenRows;
  const int rows = min(kGenRows, a.ny - ky0);
  float dsum[kMaxDet + 1];
#pragma unroll
  for (int d = 0; d <= kMaxDet; d++) dsum[d] = 0.f;
  float2* data = a.data + (size_t)p * a.nx * a.ny;
  for (int idx = threadIdx.x; idx < rows * a.nx; idx += blockDim.x) {
    const int ky = ky0 + idx / a.nx, kx = idx % a.nx;
    const size_t k = (size_t)ky * a.nx + kx;
    float2 v = data[k];
    const float kx2 = a.kx2[kx];
    const bool in = kx2 < a.k2max && (kx2 + a.ky2[ky]) < a.k2max;
    v = in ? cmul(cmul(v, a.py[ky]), a.px[kx]) : make_float2(0.f, 0.f);
    data[k] = v;
    const float c = cnorm(v);
    if (a.rec) {
      dsum[kMaxDet] += c;
#pragma unroll
      for (int d = 0; d < kMaxDet; d++) {
        if (d >= a.ndet) break;
        int sx = (kx - a.det_sx[d]) % a.nx, sy = (ky - a.det_sy[d]) % a.ny;
        sx += sx < 0 ? a.nx : 0;
        sy += sy < 0 ? a.ny : 0;
        const float k2 = a.kx2[sx] + a.ky2[sy];
        dsum[d] += (k2 >= a.det_k2in[d] && k2 <= a.det_k2out[d]) ? c : 0.f;
      }
    }
    if (a.diff)
      a.diff[((size_t)(a.probe0 + p) * a.nx + (kx + a.nx / 2) % a.nx) * a.ny + (ky + a.ny / 2) % a.ny] =
          c * a.diff_scale;
  }
  if (!a.rec) return;
  __shared__ float red[kMaxDet + 1][4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int d = 0; d <= kMaxDet; d++) {
    float s = dsum[d];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) red[d][wv] = s;
  }
  __syncthreads();
  if (threadIdx.x <= kMaxDet)
    a.partial[((size_t)p * gridDim.x + chunk) * (kMaxDet + 1) + threadIdx.x] =
        red[threadIdx.x][0] + red[threadIdx.x][1] + red[threadIdx.x][2] + red[threadIdx.x][3];
}

static unsigned gen_grid(size_t n) {
  size_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  return (unsigned)(g ? g : 1);
}

void launch_gen_transmit(const GenArgs& a, bool first, hipStream_t st) {
  hipLaunchKernelGGL(k_gen_transmit, dim3(gen_grid((size_t)a.nprobe * a.nx * a.ny)), dim3(256), 0, st, a,
                     first ? 1 : 0);
}

int gen_chunks(int ny) { return (ny + kGenRows - 1) / kGenRows; }

void launch_gen_propagate(const GenPropArgs& a, int nprobe, hipStream_t st) {
  hipLaunchKernelGGL(k_gen_propagate, dim3((unsigned)gen_chunks(a.ny), (unsigned)nprobe), dim3(256), 0, st, a);
}

}  // namespace qs
