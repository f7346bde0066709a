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

#include <algorithm>

#include "qs_dft.hpp"
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


// ============================================================================
// Mixed-radix path for wave grids whose sizes factor into 2, 3, 5 and 7 (400 =
// 4*4*5*5, 384, 480, 600, ...): the fused form of the power-of-two sweeps without
// their compile-time plans.  The state between slices is the y-spectrum
// [probe][x][ky] of the wave, stored only for the ky columns inside the 2/3 band
// (the others are zero after propagation and are neither read nor written):
//   k_gen_rows (per x row):       y-IFFT (finishes slice s-1) -> transmit ->
//                                 y-FFT (starts slice s); exit mode: y-IFFT only
//   k_gen_cols (per column tile): x-FFT -> propagator, band limit, detector sums,
//                                 diffraction pattern -> x-IFFT
// Transforms run in LDS as Stockham passes with a runtime radix list and a
// twiddle table W_n^m = exp(-2 pi i m/n), m < n.

namespace {

// Largest LDS footprint of one block in float2 (V vectors of gen_ld(n) each), and
// the block size.
constexpr int kGenLds = 8064, kGenBlock = 512;  // 63 KB of LDS, 1 KB left for static arrays

// LDS position of element i of a vector: one float2 of padding every 16, so that
// the strided writes of the first passes (element j*R + k) spread over the banks.
__device__ __forceinline__ int lp(int i) { return i + (i >> 4); }
// Vector stride in LDS for length n (odd, so that adjacent vectors start on other banks).
__host__ __device__ constexpr int gen_ld(int n) { return n + (n >> 4) + 1; }

// q = a / d, r = a % d for 0 <= a < 2^24, d >= 1 and a/d < 2^21, with rd = 1/d: the
// float quotient is then within one of the integer one (relative error below 2^-22),
// and one correction step makes it exact.
__device__ __forceinline__ void fdivmod(int a, int d, float rd, int& q, int& r) {
  q = (int)((float)a * rd);
  r = a - q * d;
  if (r >= d) {
    q++;
    r -= d;
  } else if (r < 0) {
    q--;
    r += d;
  }
}

// One Stockham pass of radix R over V vectors of length n in LDS (element i of
// vector v at x[v*ld + lp(i)]), all threads of the block; ns = product of the
// earlier radices.  All reads precede all writes (in place, two barriers).
template <int R, int DIR>
__device__ __forceinline__ void gen_pass(float2* x, int V, int n, int ld, int ns, const float2* __restrict__ tw,
                                         int tid) {
  constexpr int MAXB = (kGenLds / R + kGenBlock - 1) / kGenBlock;
  const int nb = n / R, total = V * nb, stride = n / (ns * R);
  const float rnb = 1.0f / (float)nb, rns = 1.0f / (float)ns;
  float2 a[MAXB][R];
  int dsto[MAXB];
#pragma unroll
  for (int i = 0; i < MAXB; i++) {
    const int b = tid + i * kGenBlock;
    if (b < total) {
      int v, j, q, jm;
      fdivmod(b, nb, rnb, v, j);
      fdivmod(j, ns, rns, q, jm);
      const float2* src = x + v * ld;
#pragma unroll
      for (int k = 0; k < R; k++) a[i][k] = src[lp(j + k * nb)];
      if (ns > 1) {
#pragma unroll
        for (int k = 1; k < R; k++) {
          const float2 w = tw[k * jm * stride];
          a[i][k] = DIR < 0 ? cmul(a[i][k], w) : cmulc(a[i][k], w);
        }
      }
      dft_any<R, DIR>(a[i]);
      dsto[i] = (v << 16) | (q * ns * R + jm);  // vector (< 16) and output element (< 2^16)
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < MAXB; i++) {
    const int b = tid + i * kGenBlock;
    if (b < total) {
      const int e = dsto[i] & 0xffff;
      float2* dst = x + (dsto[i] >> 16) * ld;
#pragma unroll
      for (int k = 0; k < R; k++) dst[lp(e + k * ns)] = a[i][k];
    }
  }
  __syncthreads();
}

template <int DIR>
__device__ void gen_fft(float2* x, int V, int ld, const GenPlan& pl, const float2* __restrict__ tw) {
  int ns = 1;
  for (int ip = 0; ip < pl.np; ip++) {
    const int R = pl.r[ip];
    // the thread index is made opaque per pass so that no pass's index arithmetic
    // is hoisted out of the loop (it would hold registers through all passes)
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    if (R == 2) gen_pass<2, DIR>(x, V, pl.n, ld, ns, tw, tid);
    else if (R == 3) gen_pass<3, DIR>(x, V, pl.n, ld, ns, tw, tid);
    else if (R == 4) gen_pass<4, DIR>(x, V, pl.n, ld, ns, tw, tid);
    else if (R == 5) gen_pass<5, DIR>(x, V, pl.n, ld, ns, tw, tid);
    else if (R == 7) gen_pass<7, DIR>(x, V, pl.n, ld, ns, tw, tid);
    else if (R == 8) gen_pass<8, DIR>(x, V, pl.n, ld, ns, tw, tid);
    else if (R == 9) gen_pass<9, DIR>(x, V, pl.n, ld, ns, tw, tid);
    else if (R == 16) gen_pass<16, DIR>(x, V, pl.n, ld, ns, tw, tid);
    else gen_pass<25, DIR>(x, V, pl.n, ld, ns, tw, tid);
    ns *= R;
  }
}

// Elements per thread when a block holds kGenLds values: all global loads of a
// phase are issued before the first LDS store (one batch of kGenPer per thread).
constexpr int kGenPer = (kGenLds + kGenBlock - 1) / kGenBlock;

}  // namespace

// Rows: a.rpb rows (probe, x) per block (a power of two); thread t works on row
// t / TPR of the block and its columns y = t % TPR + k TPR, TPR = kGenBlock / rpb.
// mode 0: the probe is the input (first slice of a fresh run); 1: the state's
// y-spectrum (y-IFFT first); 2: exit: y-IFFT of the state into a.exitw (real space),
// no transmit.  The transmission values are loaded with the input, before the y-IFFT.
__global__ __launch_bounds__(kGenBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_gen_rows(GenRowArgs a) {
  extern __shared__ float2 gsm[];
  const int ny = a.ny, ld = gen_ld(ny);
  const int nrows = a.nprobe * a.nx;
  const int r0 = blockIdx.x * a.rpb;
  const int V = min(a.rpb, nrows - r0);
  const int lt = 31 - __clz(kGenBlock / a.rpb);  // log2 TPR
  const int vr = threadIdx.x >> lt, yt = threadIdx.x & ((1 << lt) - 1);
  const bool rowok = vr < V;
  int p = 0, x = 0;
  if (rowok) fdivmod(r0 + vr, a.nx, 1.0f / (float)a.nx, p, x);
  const int pg = a.probe0 + p;
  const size_t rowe = ((size_t)pg * a.nx + x) * ny;  // state / exit row
  float2* lrow = gsm + vr * ld;
  const float2* trow = a.trans;
  if (rowok && a.mode != 2) {  // transmit (stemlib.cpp:2880-2887): the window's row
    const int2 ip = a.ipos[pg];
    trow = a.trans + (a.cfg ? (long long)a.cfg[pg] * a.cfg_stride : 0LL) + (size_t)(x + ip.x) * a.potNy + ip.y;
  }
  float2 sv[kGenPer], tr[kGenPer];
#pragma unroll
  for (int k = 0; k < kGenPer; k++) {
    const int y = yt + (k << lt);
    sv[k] = tr[k] = make_float2(0.f, 0.f);
    if (rowok && y < ny) {
      if (a.mode == 0) sv[k] = a.probe[(size_t)x * ny + y];
      else if (a.ky2[y] < a.k2max) sv[k] = a.state[rowe + y];
      if (a.mode != 2) tr[k] = trow[y];
    }
  }
#pragma unroll
  for (int k = 0; k < kGenPer; k++) {
    const int y = yt + (k << lt);
    if (rowok && y < ny) lrow[lp(y)] = sv[k];
  }
  __syncthreads();
  if (a.mode != 0) gen_fft<+1>(gsm, V, ld, a.plan, a.tw);
  if (a.mode == 2) {
#pragma unroll
    for (int k = 0; k < kGenPer; k++) {
      const int y = yt + (k << lt);
      if (rowok && y < ny) a.exitw[rowe + y] = lrow[lp(y)];
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < kGenPer; k++) {
    const int y = yt + (k << lt);
    if (rowok && y < ny) lrow[lp(y)] = cmul(lrow[lp(y)], tr[k]);
  }
  __syncthreads();
  gen_fft<-1>(gsm, V, ld, a.plan, a.tw);
#pragma unroll
  for (int k = 0; k < kGenPer; k++) {
    const int y = yt + (k << lt);
    if (rowok && y < ny && a.ky2[y] < a.k2max) a.state[rowe + y] = lrow[lp(y)];
  }
}

// Columns: a tile of a.tc adjacent ky columns (a power of two; a.tiles[blockIdx.x],
// only tiles with a column inside the band) of probe blockIdx.y, held column by
// column in LDS; thread t works on column t % tc and rows x = t / tc + k kGenBlock/tc.
// x-FFT, then propagate_slow's propagator and band limit (stemlib.cpp:2831-2849),
// collectIntensity's detector sums (:2623-2659) into partial[p][tile][kMaxDet + 1]
// (slot kMaxDet: the total) when rec, the fftshifted pattern into diff when set, then
// the x-IFFT.  Detector shifts arrive reduced to [0, nx) and [0, ny).
__global__ __launch_bounds__(kGenBlock) void k_gen_cols(GenColArgs a) {
  extern __shared__ float2 gsm[];
  const int nx = a.nx, ny = a.ny, ld = gen_ld(nx);
  const int p = blockIdx.y, tile = blockIdx.x;
  const int y0 = a.tiles[tile];
  const int V = min(a.tc, ny - y0);
  const int lc = 31 - __clz(a.tc);  // log2 tc
  const int c = threadIdx.x & (a.tc - 1), xt = threadIdx.x >> lc, xs = kGenBlock >> lc;
  const int ky = y0 + c;
  const bool colok = c < V;
  const bool inband = colok && a.ky2[ky] < a.k2max;
  float2* st = a.state + (size_t)(a.probe0 + p) * nx * ny + (colok ? ky : 0);
  float2* lcol = gsm + c * ld;
  float2 sv[kGenPer];
#pragma unroll
  for (int k = 0; k < kGenPer; k++) {
    const int x = xt + k * xs;
    sv[k] = make_float2(0.f, 0.f);
    if (inband && x < nx) sv[k] = st[(size_t)x * ny];
  }
#pragma unroll
  for (int k = 0; k < kGenPer; k++) {
    const int x = xt + k * xs;
    if (colok && x < nx) lcol[lp(x)] = sv[k];
  }
  __syncthreads();
  gen_fft<-1>(gsm, V, ld, a.plan, a.tw);
  float dsum[kMaxDet + 1];
#pragma unroll
  for (int d = 0; d <= kMaxDet; d++) dsum[d] = 0.f;
  const float2 pyk = colok ? a.py[ky] : make_float2(0.f, 0.f);
  const float ky2 = colok ? a.ky2[ky] : 0.f;
  for (int k = 0; k < kGenPer; k++) {
    const int kx = xt + k * xs;
    if (!colok || kx >= nx) continue;
    float2 v = lcol[lp(kx)];
    const float kx2 = a.kx2[kx];
    const bool in = kx2 < a.k2max && (kx2 + ky2) < a.k2max;
    v = in ? cmul(cmul(v, pyk), a.px[kx]) : make_float2(0.f, 0.f);
    lcol[lp(kx)] = v;
    const float e = cnorm(v);
    if (a.rec) {
      dsum[kMaxDet] += e;
      for (int d = 0; d < a.ndet; d++) {
        int sx = kx - a.det_sx[d], sy = ky - a.det_sy[d];
        sx += sx < 0 ? nx : 0;
        sy += sy < 0 ? ny : 0;
        const float k2 = a.kx2[sx] + a.ky2[sy];
        dsum[d] += (k2 >= a.det_k2in[d] && k2 <= a.det_k2out[d]) ? e : 0.f;
      }
    }
    if (a.diff) {
      const int hx = kx + nx / 2, hy = ky + ny / 2;
      a.diff[((size_t)(a.probe0 + p) * nx + (hx >= nx ? hx - nx : hx)) * ny + (hy >= ny ? hy - ny : hy)] =
          e * a.diff_scale;
    }
  }
  if (a.rec) {
    __shared__ float red[kMaxDet + 1][kGenBlock / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int d = 0; d <= kMaxDet; d++) {
      float s = dsum[d];
      for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
      if (lane == 0) red[d][wv] = s;
    }
    __syncthreads();
    if (threadIdx.x <= kMaxDet) {
      float s = 0.f;
      for (int w = 0; w < kGenBlock / 64; w++) s += red[threadIdx.x][w];
      a.partial[((size_t)p * gridDim.x + tile) * (kMaxDet + 1) + threadIdx.x] = s;
    }
  }
  __syncthreads();
  gen_fft<+1>(gsm, V, ld, a.plan, a.tw);
#pragma unroll
  for (int k = 0; k < kGenPer; k++) {
    const int x = xt + k * xs;
    if (inband && x < nx) st[(size_t)x * ny] = lcol[lp(x)];
  }
}

// Row transforms of rows of length n (contiguous, nrows of them) through the
// mixed-radix LDS passes: a.rpb rows per block (a power of two), thread t on row
// t / TPR and elements t % TPR + k TPR.
//   direct (a.M == 0): length-n transform in direction DIR (n's plan), times scale;
//   Bluestein (a.M > 0): chirp[k] x row, zero-padded to M, FFT_M, x vhat, IFFT_M,
//   x chirp x scale/M (a.plan is M's; chirp[k] = exp(DIR i pi k^2/n), vhat = FFT_M of
//   the wrapped conj(chirp)).
template <int DIR>
__global__ __launch_bounds__(kGenBlock) void k_gen_rowxform(GenRowXformArgs a) {
  extern __shared__ float2 gsm[];
  const int L = a.M > 0 ? a.M : a.n, ld = gen_ld(L);
  const long r0 = (long)blockIdx.x * a.rpb;
  const int V = (int)min((long)a.rpb, a.nrows - r0);
  const int lt = 31 - __clz(kGenBlock / a.rpb);
  const int vr = threadIdx.x >> lt, yt = threadIdx.x & ((1 << lt) - 1);
  const bool rowok = vr < V;
  float2* row = a.data + (r0 + (rowok ? vr : 0)) * a.n;
  float2* lrow = gsm + vr * ld;
  float2 sv[kGenPer];
#pragma unroll
  for (int k = 0; k < kGenPer; k++) {
    const int y = yt + (k << lt);
    sv[k] = make_float2(0.f, 0.f);
    if (rowok && y < a.n) {
      sv[k] = row[y];
      if (a.M > 0) sv[k] = cmul(sv[k], a.chirp[y]);
    }
  }
#pragma unroll
  for (int k = 0; k < kGenPer; k++) {
    const int y = yt + (k << lt);
    if (rowok && y < L) lrow[lp(y)] = sv[k];
  }
  __syncthreads();
  float s = a.scale;
  if (a.M > 0) {
    gen_fft<-1>(gsm, V, ld, a.plan, a.tw);
#pragma unroll
    for (int k = 0; k < kGenPer; k++) {
      const int y = yt + (k << lt);
      if (rowok && y < L) lrow[lp(y)] = cmul(lrow[lp(y)], a.vhat[y]);
    }
    __syncthreads();
    gen_fft<+1>(gsm, V, ld, a.plan, a.tw);
    s /= (float)a.M;
  } else {
    gen_fft<DIR>(gsm, V, ld, a.plan, a.tw);
  }
#pragma unroll
  for (int k = 0; k < kGenPer; k++) {
    const int y = yt + (k << lt);
    if (rowok && y < a.n) {
      float2 v = lrow[lp(y)];
      if (a.M > 0) v = cmul(v, a.chirp[y]);
      row[y] = cscale(v, s);
    }
  }
}

void launch_gen_rowxform(const GenRowXformArgs& a, int dir, hipStream_t st) {
  const int L = a.M > 0 ? a.M : a.n;
  const unsigned grid = (unsigned)((a.nrows + a.rpb - 1) / a.rpb);
  const size_t lds = sizeof(float2) * a.rpb * gen_ld(L);
  if (dir < 0) hipLaunchKernelGGL(k_gen_rowxform<-1>, dim3(grid), dim3(kGenBlock), lds, st, a);
  else hipLaunchKernelGGL(k_gen_rowxform<+1>, dim3(grid), dim3(kGenBlock), lds, st, a);
}

// ============================================================================
// Compiled-plan sweeps (qs_mix.hpp, one translation unit per length): dispatch.
#define QS_MIX_DECL(n)                                                   \
  void mix_rows_##n(const GenRowArgs& a, hipStream_t st);                \
  void mix_cols_##n(const GenColArgs& a, int nblocks, hipStream_t st);   \
  int mix_cpb_##n();
QS_MIX_SIZES(QS_MIX_DECL)
#undef QS_MIX_DECL

bool mix_compiled(int n) {
#define QS_MIX_CASE(m) if (n == m) return true;
  QS_MIX_SIZES(QS_MIX_CASE)
#undef QS_MIX_CASE
  return false;
}

int mix_cols_per_block(int n) {
#define QS_MIX_CASE(m) if (n == m) return mix_cpb_##m();
  QS_MIX_SIZES(QS_MIX_CASE)
#undef QS_MIX_CASE
  return 0;
}

void launch_mix_rows(const GenRowArgs& a, hipStream_t st) {
#define QS_MIX_CASE(m) if (a.ny == m) return mix_rows_##m(a, st);
  QS_MIX_SIZES(QS_MIX_CASE)
#undef QS_MIX_CASE
}

void launch_mix_cols(const GenColArgs& a, int nblocks, hipStream_t st) {
#define QS_MIX_CASE(m) if (a.nx == m) return mix_cols_##m(a, nblocks, st);
  QS_MIX_SIZES(QS_MIX_CASE)
#undef QS_MIX_CASE
}

bool gen_plan(int n, GenPlan* pl) { return gen_plan_variant(n, 0, pl); }

bool gen_plan_variant(int n, int variant, GenPlan* pl) {
  if (variant < 0 || variant >= kGenPlanVariants) return false;
  if (variant == kGenPlanCompiled) return mix_compiled(n) && gen_plan_variant(n, 0, pl);
  int m = n, e2 = 0;
  pl->n = n;
  pl->np = 0;
  while (m % 2 == 0) {
    m /= 2;
    e2++;
  }
  auto push = [&](int r) {
    if (pl->np >= kGenMaxPasses) return false;
    pl->r[pl->np++] = r;
    return true;
  };
  // variant 0: few passes: radix 16 for powers of two (then 8, 4 or 2 for the rest),
  // 25 and 9 for pairs of fives and threes, then 5, 3, 7; variant 1: radix 8 for powers
  // of two, 25 and 9; variant 2: radices 2-8 only (4 and 8 for powers of two)
  const int big = variant == 0 ? 4 : 3;  // log2 of the power-of-two radix
  if (variant == 2) {
    while (e2 >= 3 && (e2 % 2 == 1 || e2 >= 6)) {
      if (!push(8)) return false;
      e2 -= 3;
    }
    while (e2 >= 2) {
      if (!push(4)) return false;
      e2 -= 2;
    }
  } else {
    while (e2 >= big) {
      if (!push(1 << big)) return false;
      e2 -= big;
    }
  }
  if (e2 > 0 && !push(1 << e2)) return false;
  for (int r : {5, 3}) {
    while (variant != 2 && m % (r * r) == 0) {
      if (!push(r * r)) return false;
      m /= r * r;
    }
    while (m % r == 0) {
      if (!push(r)) return false;
      m /= r;
    }
  }
  while (m % 7 == 0) {
    if (!push(7)) return false;
    m /= 7;
  }
  return m == 1 && n >= 2 && (long)gen_ld(n) <= kGenLds;
}

// largest power of two <= min(16, kGenLds / gen_ld(n)) (at least 1)
static int gen_pow2_fit(int n) {
  const int m = std::max(1, std::min(16, kGenLds / gen_ld(n)));
  int r = 1;
  while (2 * r <= m) r *= 2;
  return r;
}
int gen_rows_per_block(int ny) { return gen_pow2_fit(ny); }
int gen_cols_per_tile(int nx) { return gen_pow2_fit(nx); }
int gen_max_vectors(int n) { return kGenLds / gen_ld(n); }

void launch_gen_rows(const GenRowArgs& a, hipStream_t st) {
  const int nrows = a.nprobe * a.nx;
  const unsigned grid = (unsigned)((nrows + a.rpb - 1) / a.rpb);
  hipLaunchKernelGGL(k_gen_rows, dim3(grid), dim3(kGenBlock), sizeof(float2) * a.rpb * gen_ld(a.ny), st, a);
}

void launch_gen_cols(const GenColArgs& a, int ntiles, hipStream_t st) {
  hipLaunchKernelGGL(k_gen_cols, dim3((unsigned)ntiles, (unsigned)a.nprobe), dim3(kGenBlock),
                     sizeof(float2) * a.tc * gen_ld(a.nx), st, a);
}

}  // namespace qs
