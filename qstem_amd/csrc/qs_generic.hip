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

template <int DIR>
__device__ __forceinline__ float2 rot_i(float2 v) {  // DIR * i * v
  return DIR < 0 ? make_float2(v.y, -v.x) : make_float2(-v.y, v.x);
}

template <int DIR>
__device__ __forceinline__ void dft3(float2* a) {
  const float h = 0.86602540378443865f;  // sin(2 pi/3)
  const float2 b = cadd(a[1], a[2]), d = csub(a[1], a[2]);
  const float2 t = make_float2(a[0].x - 0.5f * b.x, a[0].y - 0.5f * b.y);
  const float2 u = rot_i<DIR>(cscale(d, h));
  a[0] = cadd(a[0], b);
  a[1] = cadd(t, u);
  a[2] = csub(t, u);
}

template <int DIR>
__device__ __forceinline__ void dft5(float2* a) {
  const float c1 = 0.30901699437494742f, c2 = -0.80901699437494742f;  // cos(2 pi/5), cos(4 pi/5)
  const float s1 = 0.95105651629515357f, s2 = 0.58778525229247313f;   // sin(2 pi/5), sin(4 pi/5)
  const float2 b1 = cadd(a[1], a[4]), b2 = cadd(a[2], a[3]);
  const float2 d1 = csub(a[1], a[4]), d2 = csub(a[2], a[3]);
  const float2 a0 = a[0];
  const float2 r1 = make_float2(a0.x + c1 * b1.x + c2 * b2.x, a0.y + c1 * b1.y + c2 * b2.y);
  const float2 r2 = make_float2(a0.x + c2 * b1.x + c1 * b2.x, a0.y + c2 * b1.y + c1 * b2.y);
  const float2 i1 = rot_i<DIR>(make_float2(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y));
  const float2 i2 = rot_i<DIR>(make_float2(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y));
  a[0] = cadd(a0, cadd(b1, b2));
  a[1] = cadd(r1, i1);
  a[4] = csub(r1, i1);
  a[2] = cadd(r2, i2);
  a[3] = csub(r2, i2);
}

template <int DIR>
__device__ __forceinline__ void dft7(float2* a) {
  const float c1 = 0.62348980185873353f, c2 = -0.22252093395631440f, c3 = -0.90096886790241913f;
  const float s1 = 0.78183148246802981f, s2 = 0.97492791218182361f, s3 = 0.43388373911755812f;
  const float2 b1 = cadd(a[1], a[6]), b2 = cadd(a[2], a[5]), b3 = cadd(a[3], a[4]);
  const float2 d1 = csub(a[1], a[6]), d2 = csub(a[2], a[5]), d3 = csub(a[3], a[4]);
  const float2 a0 = a[0];
  // X_k = a0 + sum_j cos(2 pi jk/7) b_j +- DIR i sum_j sin(2 pi jk/7) d_j, k = 1, 2, 3
  auto re = [&](float x1, float x2, float x3) {
    return make_float2(a0.x + x1 * b1.x + x2 * b2.x + x3 * b3.x, a0.y + x1 * b1.y + x2 * b2.y + x3 * b3.y);
  };
  auto im = [&](float y1, float y2, float y3) {
    return rot_i<DIR>(make_float2(y1 * d1.x + y2 * d2.x + y3 * d3.x, y1 * d1.y + y2 * d2.y + y3 * d3.y));
  };
  const float2 r1 = re(c1, c2, c3), r2 = re(c2, c3, c1), r3 = re(c3, c1, c2);
  const float2 i1 = im(s1, s2, s3), i2 = im(s2, -s3, -s1), i3 = im(s3, -s1, s2);
  a[0] = cadd(a0, cadd(b1, cadd(b2, b3)));
  a[1] = cadd(r1, i1);
  a[6] = csub(r1, i1);
  a[2] = cadd(r2, i2);
  a[5] = csub(r2, i2);
  a[3] = cadd(r3, i3);
  a[4] = csub(r3, i3);
}

template <int R1, int R2, int DIR>
__device__ __forceinline__ void dft_2stage(float2* a);

template <int R, int DIR>
__device__ __forceinline__ void dft_any(float2* a) {
  if constexpr (R == 3) dft3<DIR>(a);
  else if constexpr (R == 5) dft5<DIR>(a);
  else if constexpr (R == 7) dft7<DIR>(a);
  else if constexpr (R == 9) dft_2stage<3, 3, DIR>(a);
  else if constexpr (R == 20) dft_2stage<4, 5, DIR>(a);
  else if constexpr (R == 25) dft_2stage<5, 5, DIR>(a);
  else dft<R, DIR>(a);
}

// Length R1*R2 DFT in registers (natural order in and out) by one Cooley-Tukey step:
// input n = R2 n1 + n2, output k = k1 + R1 k2; R1-point DFTs over n1, twiddles
// W_N^(n2 k1) (constants: twiddle_at of compile-time indices), R2-point DFTs over n2.
template <int R1, int R2, int DIR>
__device__ __forceinline__ void dft_2stage(float2* a) {
  constexpr int N = R1 * R2;
  float2 t[N];
#pragma unroll
  for (int n2 = 0; n2 < R2; n2++) {
    float2 u[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; n1++) u[n1] = a[R2 * n1 + n2];
    dft_any<R1, DIR>(u);
#pragma unroll
    for (int k1 = 0; k1 < R1; k1++) {
      float2 v = u[k1];
      if ((n2 * k1) % N != 0) {
        const float2 w = twiddle_at<N>((n2 * k1) % N);
        v = DIR < 0 ? cmul(v, w) : cmulc(v, w);
      }
      t[n2 * R1 + k1] = v;
    }
  }
#pragma unroll
  for (int k1 = 0; k1 < R1; k1++) {
    float2 u[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; n2++) u[n2] = t[n2 * R1 + k1];
    dft_any<R2, DIR>(u);
#pragma unroll
    for (int k2 = 0; k2 < R2; k2++) a[k1 + R1 * k2] = u[k2];
  }
}

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
// Wavefront-resident mixed-radix sweeps (compiled plans).  The runtime-planned
// kernels above share every Stockham pass among all threads of a block: two block
// barriers per pass, and threads idle wherever the radix leaves the butterfly
// count short of the block (half of them in the radix-25 pass of 400).  For a
// length with a compiled two-pass plan N = R0 x R1, a transform is owned by T = N/E
// lanes of one wavefront, each holding E values in registers (element t + T m in
// v[m], the layout of coalesced row loads); the one exchange between the passes
// goes through the transform's own LDS slot and needs wavefront-level ordering
// only.  64/T transforms share a wavefront; lanes past them mirror the last
// transform's lanes (same inputs, same values written to the same LDS words) and
// store nothing.  State layout, modes and outputs are those of k_gen_rows and
// k_gen_cols; the column kernel reads its column straight from the state rows
// (the three columns of a wavefront are adjacent, so each row access is one
// 24-byte piece).

// E values per lane, passes of radix R0 then R1; LD: float2 per transform slot in LDS
template <int N> struct MixPlan { static constexpr int E = 0; };
template <> struct MixPlan<400> { static constexpr int E = 20, R0 = 20, R1 = 20, LD = 420; };

namespace {

constexpr int kMixWaves = 4;  // wavefronts per block
template <int N> constexpr int mix_slots() { return 64 / (N / MixPlan<N>::E); }  // transforms per wavefront
template <int N> constexpr int mix_ld() { return MixPlan<N>::LD; }
// LDS position of the value between the passes with Stockham index i = j R0 + k
// (butterfly j of the first pass, its output k): k (N/R0 + 1) + j, the first pass's
// outputs transposed with one float2 of padding per row.  At 400 (20 x 20, three
// transforms per wavefront, slot stride 420 = 4 mod 32) the first pass's writes (lanes
// of a transform on consecutive j) and the second pass's reads (stride 21, odd) fall on
// distinct eight-byte bank pairs in every 16-lane write group and 32-lane read group,
// across the three transforms of a wavefront too; four blocks of twelve slots fit the
// CU's 160 KB.
template <int N>
__device__ __forceinline__ int mix_pos(int i) {
  constexpr int R0 = MixPlan<N>::R0;
  return (i % R0) * (N / R0 + 1) + i / R0;
}

// Buffer-resource access (MUBUF): a per-lane byte offset plus a wavefront-uniform
// one, so that the E strided column elements of a lane need one address register.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t mix_rsrc(const void* base, long bytes) {
  const int n = bytes < 0x7fffffffL ? (int)bytes : 0x7fffffff;
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, n, 0x00020000);
}
__device__ __forceinline__ float2 mix_load(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ void mix_store(float2 v, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  using U2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0u, 0, 0));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v), r, voff, soff, 0);
}

// Length-N transform of v (element t + T m in v[m], in and out) in two Stockham
// passes: radix R0 over registers into the LDS slot (butterfly j = t + T b reads
// elements j + k N/R0 and writes j R0 + k), then radix R1 from the slot (reads
// j + k N/R1, twiddles W_N^(k (j mod R0)) from the table tw = W_N^m, m < N (k (j mod
// R0) < N), writes j + k R0 = t + T (b + k NB1)).
template <int N, int DIR>
__device__ __forceinline__ void mix_fft(float2* v, float2* slot, const float2* __restrict__ tw, int t) {
  using P = MixPlan<N>;
  constexpr int E = P::E, T = N / E, R0 = P::R0, R1 = P::R1, NB0 = E / R0, NB1 = E / R1;
  static_assert(R0 * R1 == N && E % R0 == 0 && E % R1 == 0 && R0 == T * NB1, "two-pass plan");
  float2 a[NB0][R0];
#pragma unroll
  for (int b = 0; b < NB0; b++) {
#pragma unroll
    for (int k = 0; k < R0; k++) a[b][k] = v[b + k * NB0];
    dft_any<R0, DIR>(a[b]);
  }
  wave_lds_sync();  // the slot's earlier reads are done
#pragma unroll
  for (int b = 0; b < NB0; b++) {
    const int j = t + T * b;
#pragma unroll
    for (int k = 0; k < R0; k++) slot[mix_pos<N>(j * R0 + k)] = a[b][k];
  }
  wave_lds_sync();
#pragma unroll
  for (int b = 0; b < NB1; b++) {
    const int j = t + T * b;
    // opaque per call: the twiddle loads are neither hoisted out of the callers'
    // loops nor shared between the two transforms of a kernel (registers held)
    int jm = j % R0;
    asm volatile("" : "+v"(jm));
    float2 c[R1];
#pragma unroll
    for (int k = 0; k < R1; k++) c[k] = slot[mix_pos<N>(j + k * (N / R1))];
    // w^k loaded for k < 4 and k a multiple of 4, the others as w^(k & ~3) w^(k & 3)
    // (7 loads instead of 19 for radix 20: fewer registers in flight)
    float2 pw[R1];
#pragma unroll
    for (int k = 1; k < R1; k++)
      if (k < 4 || (k & 3) == 0) pw[k] = tw[k * jm];
#pragma unroll
    for (int k = 5; k < R1; k++)
      if (k & 3) pw[k] = cmul(pw[k & ~3], pw[k & 3]);
#pragma unroll
    for (int k = 1; k < R1; k++) c[k] = DIR < 0 ? cmul(c[k], pw[k]) : cmulc(c[k], pw[k]);
    dft_any<R1, DIR>(c);
#pragma unroll
    for (int k = 0; k < R1; k++) v[b + k * NB1] = c[k];
  }
}

}  // namespace

// Rows (length ny = N): transform slot g of the launch advances rows g*a.rpb ..
// g*a.rpb + a.rpb - 1 (row = probe * nx + x).  Modes as k_gen_rows: 0 probe input,
// 1 state input (y-IFFT first), 2 exit (y-IFFT into a.exitw, no transmit).
template <int N>
__global__ __launch_bounds__(64 * kMixWaves) __attribute__((amdgpu_waves_per_eu(4))) void k_mix_rows(GenRowArgs a) {
  using P = MixPlan<N>;
  constexpr int E = P::E, T = N / E, S = mix_slots<N>(), LD = mix_ld<N>();
  __shared__ float2 lds[kMixWaves * S * LD];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int t = lane % T;
  const int sub = min(lane / T, S - 1);  // lanes past S*T mirror transform S-1
  const bool act = lane < S * T;
  float2* slot = lds + (wv * S + sub) * LD;
  unsigned band = 0;  // bit m: element t + T m is a ky column inside the band
#pragma unroll
  for (int m = 0; m < E; m++) band |= (a.ky2[t + T * m] < a.k2max ? 1u : 0u) << m;
  const int nrows = a.nprobe * a.nx;
  const int g = (blockIdx.x * kMixWaves + wv) * S + sub;
  for (int i = 0; i < a.rpb; i++) {
    const int r0 = g * a.rpb + i;
    const bool ok = act && r0 < nrows;
    const int r = min(r0, nrows - 1);
    const int p = r / a.nx, x = r - p * a.nx, pg = a.probe0 + p;
    const size_t rowe = ((size_t)pg * a.nx + x) * N;
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; m++) {
      const int y = t + T * m;
      if (a.mode == 0) v[m] = a.probe[(size_t)x * N + y];
      else v[m] = ((band >> m) & 1) ? a.state[rowe + y] : make_float2(0.f, 0.f);
    }
    if (a.mode != 0) mix_fft<N, +1>(v, slot, a.tw, t);
    if (a.mode == 2) {
      if (ok) {
#pragma unroll
        for (int m = 0; m < E; m++) a.exitw[rowe + t + T * m] = v[m];
      }
      continue;
    }
    // transmit (stemlib.cpp:2880-2887): the window's row, loaded after the IFFT (its
    // registers are not held through the transform)
    const int2 ip = a.ipos[pg];
    const float2* trow =
        a.trans + (a.cfg ? (long long)a.cfg[pg] * a.cfg_stride : 0LL) + (size_t)(x + ip.x) * a.potNy + ip.y;
#pragma unroll
    for (int m = 0; m < E; m++) v[m] = cmul(v[m], trow[t + T * m]);
    mix_fft<N, -1>(v, slot, a.tw, t);
    if (ok) {
#pragma unroll
      for (int m = 0; m < E; m++)
        if ((band >> m) & 1) a.state[rowe + t + T * m] = v[m];
    }
  }
}

// Columns (length nx = N): block b of probe p (a one-dimensional grid of nb blocks
// per probe) has transform slots ci = 12 b .. 12 b + 11 and slot ci takes band column
// a.tiles[ci] (a.tc of them, ascending).  The block's columns are staged through the
// slots in LDS: each state row's piece of the twelve columns is read, and written back,
// by twelve consecutive lanes (one contiguous piece per row and instruction; read
// straight by the transforms' lanes, every lane would touch twenty rows per access and
// the rows' lines would be fetched and written back in 24-byte pieces).  Twelve columns
// do not fill the rows' 128-byte lines, so the blocks of kMixGroup consecutive b (48
// columns: three lines) are placed on one XCD back to back, where they share its L2
// (blocks are dispatched round-robin over the 8 XCDs).  x-FFT, propagator, band limit,
// detector sums into partial[p][b][kMaxDet + 1] and the diffraction pattern as
// k_gen_cols, x-IFFT.  ND: most detectors (a.ndet <= ND), so that only ND + 1 sums are
// held in registers.
constexpr int kMixGroup = 4;
template <int N, int ND>
__global__ __launch_bounds__(64 * kMixWaves) __attribute__((amdgpu_waves_per_eu(ND <= 4 ? 4 : 3))) void k_mix_cols(GenColArgs a) {
  using P = MixPlan<N>;
  constexpr int E = P::E, T = N / E, S = mix_slots<N>(), LD = mix_ld<N>();
  // register slots m whose kx = t + T m lies, for every lane, strictly inside
  // (N/3, N - N/3): outside the 2/3 band, zero after the propagator
  constexpr int A3 = N / 3;
  __shared__ float2 lds[kMixWaves * S * LD];
  __shared__ float red[ND + 1][kMixWaves];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int t = lane % T;
  const int sub = min(lane / T, S - 1);
  const bool act = lane < S * T;
  float2* slot = lds + (wv * S + sub) * LD;
  // logical block: groups of kMixGroup consecutive ones per XCD (the last partial
  // round of 8 kMixGroup blocks keeps the dispatch order)
  const int nblk = (int)gridDim.x, blk = (int)blockIdx.x;
  int lb = blk;
  if (blk < (nblk / (8 * kMixGroup)) * (8 * kMixGroup)) {
    const int xcd = blk & 7, slot8 = blk >> 3;
    lb = (slot8 / kMixGroup) * (8 * kMixGroup) + xcd * kMixGroup + slot8 % kMixGroup;
  }
  const int nb = (a.tc + kMixWaves * S - 1) / (kMixWaves * S);  // blocks per probe
  const int p = lb / nb, b = lb - p * nb, ny = a.ny;
  const int ci = (b * kMixWaves + wv) * S + sub;
  const bool ok = act && ci < a.tc;
  const int ky = a.tiles[min(ci, a.tc - 1)];
  // this probe's state; element x of column ky at byte (x ny + ky) 8 (below 2^31: nx, ny <= 2048)
  const auto rs = mix_rsrc(a.state + (size_t)(a.probe0 + p) * N * ny, (long)N * ny * 8);
  // staging: lane (sc, sx) moves block column sc (slot sc) of rows sx + SR i
  constexpr int CPB = kMixWaves * S, SR = 64 * kMixWaves / CPB;
  const int sc = threadIdx.x % CPB, sx = threadIdx.x / CPB;
  const int sci = b * CPB + sc;
  const bool sok = sx < SR && sci < a.tc;
  const unsigned soff = sok ? (unsigned)(sx * ny + a.tiles[sci]) * 8u : 0u, sstep = (unsigned)(SR * ny) * 8u;
  float2* const sslot = lds + sc * LD;
  if (sok) {
    for (int i = 0; i * SR + sx < N; i++) sslot[sx + SR * i] = mix_load(rs, soff, i * sstep);
  }
  __syncthreads();
  float2 v[E];
#pragma unroll
  for (int m = 0; m < E; m++) v[m] = slot[t + T * m];
  mix_fft<N, -1>(v, slot, a.tw, t);
  const float2 pyk = a.py[ky];
  const float ky2 = a.ky2[ky];
  float dsum[ND + 1];  // detector sums of this lane's elements; dsum[ND]: the total
#pragma unroll
  for (int d = 0; d <= ND; d++) dsum[d] = 0.f;
#pragma unroll
  for (int m = 0; m < E; m++) {
    if (T * m > A3 && T * m + T - 1 < N - A3) {  // in the band gap for every lane
      v[m] = make_float2(0.f, 0.f);
      continue;
    }
    const int kx = t + T * m;
    const float kx2 = a.kx2[kx];
    const bool in = kx2 < a.k2max && (kx2 + ky2) < a.k2max;
    v[m] = in ? cmul(cmul(v[m], pyk), a.px[kx]) : make_float2(0.f, 0.f);
    const float e = cnorm(v[m]);
    if (a.rec) {
      dsum[ND] += e;
#pragma unroll
      for (int d = 0; d < ND; d++) {
        if (d >= a.ndet) break;
        float k2 = kx2 + ky2;
        if (a.det_sx[d] != 0 || a.det_sy[d] != 0) {
          int sx = kx - a.det_sx[d], sy = ky - a.det_sy[d];
          sx += sx < 0 ? N : 0;
          sy += sy < 0 ? ny : 0;
          k2 = a.kx2[sx] + a.ky2[sy];
        }
        dsum[d] += (k2 >= a.det_k2in[d] && k2 <= a.det_k2out[d]) ? e : 0.f;
      }
    }
    if (a.diff && ok) {
      const int hx = kx + N / 2, hy = ky + ny / 2;
      a.diff[((size_t)(a.probe0 + p) * N + (hx >= N ? hx - N : hx)) * ny + (hy >= ny ? hy - ny : hy)] =
          e * a.diff_scale;
    }
  }
  if (a.rec) {
#pragma unroll
    for (int d = 0; d <= ND; d++) {
      if (d < ND && d >= a.ndet) continue;
      float s = ok ? dsum[d] : 0.f;  // mirrored and surplus slots add nothing
      for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
      if (lane == 0) red[d][wv] = s;
    }
    __syncthreads();
    if (threadIdx.x < a.ndet || threadIdx.x == ND) {
      float s = 0.f;
#pragma unroll
      for (int w = 0; w < kMixWaves; w++) s += red[threadIdx.x][w];
      a.partial[((size_t)p * nb + b) * (kMaxDet + 1) + (threadIdx.x == ND ? kMaxDet : threadIdx.x)] = s;
    }
  }
  mix_fft<N, +1>(v, slot, a.tw, t);
  wave_lds_sync();  // the transform's last reads of its slot are done
#pragma unroll
  for (int m = 0; m < E; m++) slot[t + T * m] = v[m];
  __syncthreads();
  if (sok) {
    for (int i = 0; i * SR + sx < N; i++) mix_store(sslot[sx + SR * i], rs, soff, i * sstep);
  }
}

bool mix_compiled(int n) { return MixPlan<400>::E > 0 && n == 400; }
int mix_cols_per_block(int n) { return mix_compiled(n) ? kMixWaves * mix_slots<400>() : 0; }

void launch_mix_rows(const GenRowArgs& a, hipStream_t st) {
  const long nrows = (long)a.nprobe * a.nx;
  if (a.ny == 400) {
    const long per = (long)kMixWaves * mix_slots<400>() * a.rpb;
    hipLaunchKernelGGL(k_mix_rows<400>, dim3((unsigned)((nrows + per - 1) / per)), dim3(64 * kMixWaves), 0, st, a);
  }
}

void launch_mix_cols(const GenColArgs& a, int nblocks, hipStream_t st) {
  const dim3 grid((unsigned)(nblocks * a.nprobe)), block(64 * kMixWaves);
  if (a.nx == 400) {
    if (a.ndet <= 4) hipLaunchKernelGGL((k_mix_cols<400, 4>), grid, block, 0, st, a);
    else hipLaunchKernelGGL((k_mix_cols<400, kMaxDet>), grid, block, 0, st, a);
  }
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
