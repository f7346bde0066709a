// qs_sweep.hpp -- templated multislice sweep kernels and their launchers.
//
// Wave state layout in HBM between slices: state[p][x][c], complex64, where
// p is the probe, x the real-space row (0..nx-1) and c indexes only the ky
// columns that survive the 2/3 band limit of propagate_slow
// (stem3/stemlib.cpp:2831-2849); everything outside that band is zero after
// propagation, so it is neither stored nor transformed.
//
// One slice = two fused sweeps:
//   sweep A (per row x):     row IFFT (finishes slice s-1) -> transmit t_s
//                            (stemlib.cpp:2867-2888) -> row FFT (starts slice s)
//   sweep B (per ky column): column FFT (finishes the forward 2D FFT) ->
//                            propagate + band limit (:2760-2850) -> detector
//                            annuli (:2576-2687) -> column IFFT
// The 1/(nx*ny) of fft_normalize (:2890-2906) is folded into the propagator.
#pragma once
#include <hip/hip_runtime.h>

#include "qs_fft.hpp"
#include "qs_internal.hpp"

namespace qs {

constexpr int kBlock = 256;


// ============================================================================
// Sweep A: rows.
// mode 0: input is the shared real-space probe (first slice of a fresh run)
// mode 1: input is the band-compact state; row IFFT first
// mode 2: exit: row IFFT of the state, write the real-space wave, no transmit
template <int NY, int MODE>
__global__ __launch_bounds__(kBlock) void k_sweep_rows(SweepArgs a) {
  constexpr int E = Plan<NY>::E;
  constexpr int T = NY / E;
  constexpr int RPB = kBlock / T;  // rows per block
  __shared__ float2 lds[RPB * lds_slot<NY>()];
  const int slot_id = threadIdx.x / T;
  const int t = threadIdx.x % T;
  const long row = (long)blockIdx.x * RPB + slot_id;
  if (row >= (long)a.nprobe * a.nx) return;  // whole transforms only: T divides 64
  const int p = (int)(row / a.nx);
  const int x = (int)(row % a.nx);
  float2* slot = lds + slot_id * lds_slot<NY>();
  float2 v[E];
  const int pg = a.probe0 + p;  // global probe index (ipos, state)
  Twiddles<NY> tw;
  tw.init(a.twy, t);
  if constexpr (MODE == 0) {
    const float2* src = a.probe + (size_t)x * NY;
#pragma unroll
    for (int m = 0; m < E; m++) v[m] = src[t + T * m];
  } else {
    const float2* src = a.state + ((size_t)pg * a.nx + x) * a.kcp;
#pragma unroll
    for (int m = 0; m < E; m++) {
      const int iy = t + T * m;
      const int c = iy <= a.band_a ? iy : (iy >= NY - a.band_b ? iy - (NY - a.band_b) + a.band_a + 1 : -1);
      v[m] = c >= 0 ? src[c] : make_float2(0.f, 0.f);
    }
    fft_wave<NY, +1>(v, slot, tw, t);
  }
  if constexpr (MODE == 2) {
    float2* dst = a.out_wave + ((size_t)pg * a.nx + x) * NY;
#pragma unroll
    for (int m = 0; m < E; m++) dst[t + T * m] = v[m];
    return;
  } else {
    const int2 ip = a.ipos[pg];
    const float2* tr = a.trans + (size_t)(x + ip.x) * a.potNy + ip.y;
#pragma unroll
    for (int m = 0; m < E; m++) v[m] = cmul(v[m], tr[t + T * m]);
    fft_wave<NY, -1>(v, slot, tw, t);
    float2* dst = a.state + ((size_t)pg * a.nx + x) * a.kcp;
#pragma unroll
    for (int m = 0; m < E; m++) {
      const int iy = t + T * m;
      const int c = iy <= a.band_a ? iy : (iy >= NY - a.band_b ? iy - (NY - a.band_b) + a.band_a + 1 : -1);
      if (c >= 0) dst[c] = v[m];
    }
  }
}

// ============================================================================
// Sweep B: ky columns of the band, TC columns per block staged through LDS.
template <int NX> struct ColCfg {
  static constexpr int TC = NX <= 512 ? 8 : (NX == 1024 ? 8 : (NX == 2048 ? 4 : 2));
  static constexpr int STRIDE = ((lds_slot<NX>() + 31) / 32) * 32 + 2;  // float2 per LDS column
};

// Tile staging between the row-major state (TC contiguous complex per row) and
// the column-major LDS tile; 16-byte accesses (two columns per lane).
template <int NX>
__device__ __forceinline__ void stage_tile_in(float2* tile, const float2* __restrict__ base, int kcp) {
  constexpr int TC = ColCfg<NX>::TC, ST = ColCfg<NX>::STRIDE, H = TC / 2;
  for (int e = threadIdx.x; e < H * NX; e += kBlock) {
    const int c = 2 * (e % H), x = e / H;
    const float4 q = *reinterpret_cast<const float4*>(base + (size_t)x * kcp + c);
    tile[c * ST + lpad(x)] = make_float2(q.x, q.y);
    tile[(c + 1) * ST + lpad(x)] = make_float2(q.z, q.w);
  }
}

template <int NX>
__device__ __forceinline__ void stage_tile_out(const float2* tile, float2* __restrict__ base, int kcp) {
  constexpr int TC = ColCfg<NX>::TC, ST = ColCfg<NX>::STRIDE, H = TC / 2;
  for (int e = threadIdx.x; e < H * NX; e += kBlock) {
    const int c = 2 * (e % H), x = e / H;
    const float2 u = tile[c * ST + lpad(x)], w = tile[(c + 1) * ST + lpad(x)];
    *reinterpret_cast<float4*>(base + (size_t)x * kcp + c) = make_float4(u.x, u.y, w.x, w.y);
  }
}

template <int NX, bool REC, bool DIFF>
__global__ __launch_bounds__(kBlock) void k_sweep_cols(SweepArgs a) {
  constexpr int E = Plan<NX>::E;
  constexpr int T = NX / E;
  constexpr int TC = ColCfg<NX>::TC;
  constexpr int ST = ColCfg<NX>::STRIDE;
  constexpr int G = kBlock / T;  // transforms in flight per block
  constexpr int ROUNDS = (TC + G - 1) / G;
  extern __shared__ float2 tile[];  // TC * ST
  __shared__ float red[kMaxDet + 1];

  const int p = blockIdx.x / a.ntiles;
  const int tile_id = blockIdx.x % a.ntiles;
  const int pg = a.probe0 + p;
  const int c0 = tile_id * TC;
  float2* base = a.state + (size_t)pg * a.nx * a.kcp + c0;

  if constexpr (REC) {
    if (threadIdx.x <= kMaxDet) red[threadIdx.x] = 0.f;
  }
  stage_tile_in<NX>(tile, base, a.kcp);
  __syncthreads();

  const int slot_id = threadIdx.x / T;
  const int t = threadIdx.x % T;
  Twiddles<NX> tw;
  tw.init(a.twx, t);
  // per-lane propagator factors and k^2 (independent of the column)
  float2 pxr[E];
  float kx2r[E];
#pragma unroll
  for (int m = 0; m < E; m++) {
    pxr[m] = a.px[t + T * m];
    kx2r[m] = a.kx2[t + T * m];
  }
  for (int r = 0; r < ROUNDS; r++) {
    const int cl = r * G + slot_id;
    if (cl >= TC) break;  // only when G does not divide TC (uniform per transform)
    float2* col = tile + cl * ST;
    const int c = c0 + cl;
    const int iy = c <= a.band_a ? c : (c < a.kc ? c - (a.band_a + 1) + (a.ny - a.band_b) : -1);
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; m++) v[m] = col[lpad(t + T * m)];
    fft_wave<NX, -1>(v, col, tw, t);
    // propagate: P = px[ix]*py[iy] (1/(nx*ny) folded into px), zero outside k2max
    if (iy >= 0) {
      const float ky2 = a.ky2[iy];
      const float2 py = a.py[iy];
#pragma unroll
      for (int m = 0; m < E; m++) {
        const float kx2 = kx2r[m];
        if (kx2 < a.k2max && (kx2 + ky2) < a.k2max) v[m] = cmul(cmul(v[m], py), pxr[m]);
        else v[m] = make_float2(0.f, 0.f);
      }
      if constexpr (REC) {
        // collectIntensity: detector d sums |Psi|^2 over pixels whose
        // (shift-corrected) k^2 lies in [k2in, k2out]; slot kMaxDet is the total.
        float tot = 0.f;
#pragma unroll
        for (int m = 0; m < E; m++) tot += cnorm(v[m]);
        atomicAdd(&red[kMaxDet], tot);
        for (int d = 0; d < a.ndet; d++) {
          const int sx = a.det_sx[d], sy = a.det_sy[d];
          const float k2y = a.ky2[(iy - sy + 2 * a.ny) % a.ny];
          const float lo = a.det_k2in[d], hi = a.det_k2out[d];
          float s = 0.f;
#pragma unroll
          for (int m = 0; m < E; m++) {
            const int ix = t + T * m;
            const float k2 = a.kx2[(ix - sx + 2 * a.nx) % a.nx] + k2y;
            if (k2 >= lo && k2 <= hi) s += cnorm(v[m]);
          }
          atomicAdd(&red[d], s);
        }
      }
      if constexpr (DIFF) {  // fftshifted |Psi|^2/sqrt(nx*ny) of the unnormalised Psi
        float* dp = a.diff + (size_t)pg * a.nx * a.ny;
        const int sy = (iy + a.ny / 2) % a.ny;
#pragma unroll
        for (int m = 0; m < E; m++) {
          const int ix = t + T * m;
          dp[(size_t)((ix + a.nx / 2) % a.nx) * a.ny + sy] = cnorm(v[m]) * a.diff_scale;
        }
      }
    } else {
#pragma unroll
      for (int m = 0; m < E; m++) v[m] = make_float2(0.f, 0.f);
    }
    fft_wave<NX, +1>(v, col, tw, t);
    wave_lds_sync();
#pragma unroll
    for (int m = 0; m < E; m++) col[lpad(t + T * m)] = v[m];
  }
  __syncthreads();
  stage_tile_out<NX>(tile, base, a.kcp);
  if constexpr (REC) {
    if (threadIdx.x <= kMaxDet) a.partial[((size_t)p * a.ntiles + tile_id) * (kMaxDet + 1) + threadIdx.x] = red[threadIdx.x];
  }
}

// ============================================================================
// Generic batched row transforms of length L (power of two), full rows.
template <int L, int DIR>
__global__ __launch_bounds__(kBlock) void k_rows_fft(float2* __restrict__ data, long nrows, const float2* __restrict__ tw,
                                                     float scale) {
  constexpr int E = Plan<L>::E;
  constexpr int T = L / E;
  constexpr int RPB = kBlock / T;
  __shared__ float2 lds[RPB * lds_slot<L>()];
  const int slot_id = threadIdx.x / T, t = threadIdx.x % T;
  const long row = (long)blockIdx.x * RPB + slot_id;
  if (row >= nrows) return;
  float2* r = data + row * L;
  float2 v[E];
#pragma unroll
  for (int m = 0; m < E; m++) v[m] = r[t + T * m];
  Twiddles<L> twc;
  twc.init(tw, t);
  fft_wave<L, DIR>(v, lds + slot_id * lds_slot<L>(), twc, t);
#pragma unroll
  for (int m = 0; m < E; m++) r[t + T * m] = cscale(v[m], scale);
}

// Bluestein transform of rows of length n through a length-M power-of-two FFT.
// chirp[k] = exp(DIR*i*pi*k^2/n); vhat = FFT_M of the wrapped conj(chirp).
template <int M>
__global__ __launch_bounds__(kBlock) void k_rows_bluestein(float2* __restrict__ data, long nrows, int n,
                                                           const float2* __restrict__ chirp,
                                                           const float2* __restrict__ vhat,
                                                           const float2* __restrict__ tw, float scale) {
  constexpr int E = Plan<M>::E;
  constexpr int T = M / E;
  constexpr int RPB = kBlock / T;
  __shared__ float2 lds[RPB * lds_slot<M>()];
  const int slot_id = threadIdx.x / T, t = threadIdx.x % T;
  const long row = (long)blockIdx.x * RPB + slot_id;
  if (row >= nrows) return;
  float2* r = data + row * n;
  float2* slot = lds + slot_id * lds_slot<M>();
  float2 v[E];
#pragma unroll
  for (int m = 0; m < E; m++) {
    const int k = t + T * m;
    v[m] = k < n ? cmul(r[k], chirp[k]) : make_float2(0.f, 0.f);
  }
  Twiddles<M> twc;
  twc.init(tw, t);
  fft_wave<M, -1>(v, slot, twc, t);
#pragma unroll
  for (int m = 0; m < E; m++) v[m] = cmul(v[m], vhat[t + T * m]);
  fft_wave<M, +1>(v, slot, twc, t);
  const float s = scale / (float)M;
#pragma unroll
  for (int m = 0; m < E; m++) {
    const int k = t + T * m;
    if (k < n) r[k] = cscale(cmul(v[m], chirp[k]), s);
  }
}


// ============================================================================
// Launch templates, instantiated per size in qs_sizes.hip.
template <int N>
void launch_rows_mode(const SweepArgs& a, int mode, hipStream_t st) {
  constexpr int T = N / Plan<N>::E;
  constexpr int RPB = kBlock / T;
  const long rows = (long)a.nprobe * a.nx;
  dim3 grid((unsigned)((rows + RPB - 1) / RPB));
  if (mode == 0) hipLaunchKernelGGL((k_sweep_rows<N, 0>), grid, dim3(kBlock), 0, st, a);
  else if (mode == 1) hipLaunchKernelGGL((k_sweep_rows<N, 1>), grid, dim3(kBlock), 0, st, a);
  else hipLaunchKernelGGL((k_sweep_rows<N, 2>), grid, dim3(kBlock), 0, st, a);
}

template <int N>
constexpr int col_lds_bytes_t() { return (int)sizeof(float2) * ColCfg<N>::TC * ColCfg<N>::STRIDE; }

template <int N>
void launch_cols_t(const SweepArgs& a, bool rec, bool diff, hipStream_t st) {
  const size_t lds = col_lds_bytes_t<N>();
  dim3 grid((unsigned)(a.nprobe * a.ntiles));
  if (rec && diff) hipLaunchKernelGGL((k_sweep_cols<N, true, true>), grid, dim3(kBlock), lds, st, a);
  else if (rec) hipLaunchKernelGGL((k_sweep_cols<N, true, false>), grid, dim3(kBlock), lds, st, a);
  else if (diff) hipLaunchKernelGGL((k_sweep_cols<N, false, true>), grid, dim3(kBlock), lds, st, a);
  else hipLaunchKernelGGL((k_sweep_cols<N, false, false>), grid, dim3(kBlock), lds, st, a);
}

template <int N>
hipError_t set_col_lds_t() {
  const int b = col_lds_bytes_t<N>();
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sweep_cols<N, false, false>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, b);
  if (e == hipSuccess)
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sweep_cols<N, true, false>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, b);
  if (e == hipSuccess)
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sweep_cols<N, false, true>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, b);
  if (e == hipSuccess)
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sweep_cols<N, true, true>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, b);
  return e;
}

template <int L, int DIR>
void rows_fft_t(float2* d, long nrows, const float2* tw, float scale, hipStream_t st) {
  constexpr int RPB = kBlock / (L / Plan<L>::E);
  hipLaunchKernelGGL((k_rows_fft<L, DIR>), dim3((unsigned)((nrows + RPB - 1) / RPB)), dim3(kBlock), 0, st, d, nrows, tw,
                     scale);
}

template <int M>
void blue_t(float2* d, long nrows, int n, const float2* chirp, const float2* vhat, const float2* tw, float scale,
            hipStream_t st) {
  constexpr int RPB = kBlock / (M / Plan<M>::E);
  hipLaunchKernelGGL((k_rows_bluestein<M>), dim3((unsigned)((nrows + RPB - 1) / RPB)), dim3(kBlock), 0, st, d, nrows, n,
                     chirp, vhat, tw, scale);
}


// ============================================================================
// Adjoint sweeps (reverse mode of the two forward sweeps).
//
// Column sweep (slice s, reverse):  g_Phi = colFFT(g_X)  [last slice: 4 (I - T) Phi,
// Phi = colFFT(X_{S-1})/nx, I = |Phi|^2]  ->  g_Psi = conj(P) g_Phi  ->  g_R = colIFFT(g_Psi).
template <int NX, bool LAST>
__global__ __launch_bounds__(kBlock) void k_adj_cols(AdjArgs A) {
  const SweepArgs& a = A.f;
  constexpr int E = Plan<NX>::E;
  constexpr int T = NX / E;
  constexpr int TC = ColCfg<NX>::TC;
  constexpr int ST = ColCfg<NX>::STRIDE;
  constexpr int G = kBlock / T;
  constexpr int ROUNDS = (TC + G - 1) / G;
  extern __shared__ float2 tile[];
  __shared__ float red;
  const int p = blockIdx.x / a.ntiles;
  const int tile_id = blockIdx.x % a.ntiles;
  const int pg = a.probe0 + p;
  const int c0 = tile_id * TC;
  const size_t off = (size_t)pg * a.nx * a.kcp + c0;
  const float2* src = (LAST ? A.xprev : A.gx) + off;
  float2* dst = A.gx + off;
  if (threadIdx.x == 0) red = 0.f;
  stage_tile_in<NX>(tile, src, a.kcp);
  __syncthreads();
  const int slot_id = threadIdx.x / T;
  const int t = threadIdx.x % T;
  Twiddles<NX> tw;
  tw.init(a.twx, t);
  float lsum = 0.f;
  for (int r = 0; r < ROUNDS; r++) {
    const int cl = r * G + slot_id;
    if (cl >= TC) break;
    float2* col = tile + cl * ST;
    const int c = c0 + cl;
    const int iy = c <= a.band_a ? c : (c < a.kc ? c - (a.band_a + 1) + (a.ny - a.band_b) : -1);
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; m++) v[m] = col[lpad(t + T * m)];
    fft_wave<NX, -1>(v, col, tw, t);
    if (iy >= 0) {
      const float ky2 = a.ky2[iy];
      const float2 py = a.py[iy];
      const float inx = 1.0f / (float)a.nx;
#pragma unroll
      for (int m = 0; m < E; m++) {
        const int ix = t + T * m;
        const float kx2 = a.kx2[ix];
        const bool in = kx2 < a.k2max && (kx2 + ky2) < a.k2max;
        float2 g;
        if constexpr (LAST) {
          const float2 phi = in ? cscale(v[m], inx) : make_float2(0.f, 0.f);
          const float d = cnorm(phi) - A.target[((size_t)pg * a.nx + ix) * a.ny + iy];
          lsum += d * d;
          g = cscale(phi, 4.0f * d);
        } else {
          g = v[m];
        }
        v[m] = in ? cmulc(cmulc(g, py), a.px[ix]) : make_float2(0.f, 0.f);
      }
    } else {
#pragma unroll
      for (int m = 0; m < E; m++) v[m] = make_float2(0.f, 0.f);
    }
    fft_wave<NX, +1>(v, col, tw, t);
    wave_lds_sync();
#pragma unroll
    for (int m = 0; m < E; m++) col[lpad(t + T * m)] = v[m];
  }
  if constexpr (LAST) atomicAdd(&red, lsum);
  __syncthreads();
  stage_tile_out<NX>(tile, dst, a.kcp);
  if constexpr (LAST) {
    if (threadIdx.x == 0) A.loss_partial[(size_t)p * a.ntiles + tile_id] = red;
  }
}

// Row sweep (slice s, reverse):  g_a = rowIFFT(g_R);  psi_s = rowIFFT(X_{s-1}) (or the
// probe);  g_t += conj(psi_s) g_a;  g_psi = conj(t_s) g_a;  g_X_{s-1} = rowFFT(g_psi)
// (s > 0) or sum_p g_psi -> gprobe (s = 0).
template <int NY, bool FIRST>
__global__ __launch_bounds__(kBlock) void k_adj_rows(AdjArgs A) {
  const SweepArgs& a = A.f;
  constexpr int E = Plan<NY>::E;
  constexpr int T = NY / E;
  constexpr int RPB = kBlock / T;
  __shared__ float2 lds[RPB * lds_slot<NY>()];
  const int slot_id = threadIdx.x / T;
  const int t = threadIdx.x % T;
  const long row = (long)blockIdx.x * RPB + slot_id;
  if (row >= (long)a.nprobe * a.nx) return;
  const int p = (int)(row / a.nx);
  const int x = (int)(row % a.nx);
  const int pg = a.probe0 + p;
  float2* slot = lds + slot_id * lds_slot<NY>();
  Twiddles<NY> tw;
  tw.init(a.twy, t);
  float2* grow = A.gx + ((size_t)pg * a.nx + x) * a.kcp;
  float2 g[E], w[E];
#pragma unroll
  for (int m = 0; m < E; m++) {
    const int iy = t + T * m;
    const int c = iy <= a.band_a ? iy : (iy >= NY - a.band_b ? iy - (NY - a.band_b) + a.band_a + 1 : -1);
    g[m] = c >= 0 ? grow[c] : make_float2(0.f, 0.f);
  }
  fft_wave<NY, +1>(g, slot, tw, t);
  if constexpr (FIRST) {
    const float2* pr = a.probe + (size_t)x * NY;
#pragma unroll
    for (int m = 0; m < E; m++) w[m] = pr[t + T * m];
  } else {
    const float2* xr = A.xprev + ((size_t)pg * a.nx + x) * a.kcp;
#pragma unroll
    for (int m = 0; m < E; m++) {
      const int iy = t + T * m;
      const int c = iy <= a.band_a ? iy : (iy >= NY - a.band_b ? iy - (NY - a.band_b) + a.band_a + 1 : -1);
      w[m] = c >= 0 ? xr[c] : make_float2(0.f, 0.f);
    }
    fft_wave<NY, +1>(w, slot, tw, t);
  }
  const int2 ip = a.ipos[pg];
  const float2* tr = a.trans + (size_t)(x + ip.x) * a.potNy + ip.y;
  float* gt = A.gT + 2 * ((size_t)(x + ip.x) * a.potNy + ip.y);
#pragma unroll
  for (int m = 0; m < E; m++) {
    const int iy = t + T * m;
    const float2 gtv = cmulc(g[m], w[m]);  // conj(psi) * g_a
    atomicAdd(gt + 2 * iy, gtv.x);
    atomicAdd(gt + 2 * iy + 1, gtv.y);
    g[m] = cmulc(g[m], tr[iy]);  // conj(t) * g_a
  }
  if constexpr (FIRST) {
    float* gp = A.gprobe + 2 * (size_t)x * NY;
#pragma unroll
    for (int m = 0; m < E; m++) {
      atomicAdd(gp + 2 * (t + T * m), g[m].x);
      atomicAdd(gp + 2 * (t + T * m) + 1, g[m].y);
    }
  } else {
    fft_wave<NY, -1>(g, slot, tw, t);
#pragma unroll
    for (int m = 0; m < E; m++) {
      const int iy = t + T * m;
      const int c = iy <= a.band_a ? iy : (iy >= NY - a.band_b ? iy - (NY - a.band_b) + a.band_a + 1 : -1);
      if (c >= 0) grow[c] = g[m];
    }
  }
}

template <int N>
void launch_adj_rows_t(const AdjArgs& a, bool first, hipStream_t st) {
  constexpr int T = N / Plan<N>::E;
  constexpr int RPB = kBlock / T;
  const long rows = (long)a.f.nprobe * a.f.nx;
  dim3 grid((unsigned)((rows + RPB - 1) / RPB));
  if (first) hipLaunchKernelGGL((k_adj_rows<N, true>), grid, dim3(kBlock), 0, st, a);
  else hipLaunchKernelGGL((k_adj_rows<N, false>), grid, dim3(kBlock), 0, st, a);
}

template <int N>
void launch_adj_cols_t(const AdjArgs& a, bool last, hipStream_t st) {
  const size_t lds = col_lds_bytes_t<N>();
  dim3 grid((unsigned)(a.f.nprobe * a.f.ntiles));
  if (last) hipLaunchKernelGGL((k_adj_cols<N, true>), grid, dim3(kBlock), lds, st, a);
  else hipLaunchKernelGGL((k_adj_cols<N, false>), grid, dim3(kBlock), lds, st, a);
}

template <int N>
hipError_t set_adj_lds_t() {
  const int b = col_lds_bytes_t<N>();
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_adj_cols<N, true>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, b);
  if (e == hipSuccess)
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_adj_cols<N, false>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, b);
  return e;
}

}  // namespace qs
