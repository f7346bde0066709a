// qs_fft.hpp -- wavefront-resident Stockham FFT building blocks for gfx950.
//
// A length-N transform is owned by T = N/E consecutive threads: lanes of one
// 64-wide wavefront up to N = 1024, a whole 256-thread block beyond; each
// thread holds E complex values in registers.  In the natural layout thread t
// holds elements t + T*m, m = 0..E-1, which is also the layout coalesced
// global loads and stores produce, so the first pass takes its input straight
// from registers and the last pass leaves its output there.  Between passes
// the values are exchanged through a per-transform LDS slot.
//
// Pass structure (Stockham autosort): for radix R and span Ns, butterfly
// j = t + T*b (b < E/R) reads x[j + k*N/R], applies twiddles
// W_{Ns R}^{k (j mod Ns)} and writes y_k to out[(j/Ns)*Ns*R + j%Ns + k*Ns].
// Directions follow FFTW/QSTEM: DIR = -1 forward, +1 backward, both
// unnormalised (libs/data_containers.cpp:50-51).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

namespace qs {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float cnorm(float2 a) { return a.x * a.x + a.y * a.y; }
// multiply by -i (DIR=-1) or +i (DIR=+1)
template <int DIR>
__device__ __forceinline__ float2 mul_i(float2 a) {
  return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

// ---- in-register DFTs ------------------------------------------------------
template <int DIR>
__device__ __forceinline__ void dft2(float2& a, float2& b) {
  float2 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <int DIR>
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
  float2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
  float2 t2 = cadd(a1, a3), t3 = mul_i<DIR>(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

// dft4 with inputs known to be zero (bit j of Z: input j is zero); the
// additions with a zero operand are left out.
__device__ __forceinline__ float2 cneg(float2 a) { return make_float2(-a.x, -a.y); }
template <int DIR, int Z>
__device__ __forceinline__ void dft4z(float2& a0, float2& a1, float2& a2, float2& a3) {
  if constexpr (Z == 0) {
    dft4<DIR>(a0, a1, a2, a3);
  } else {
    float2 t0, t1, t2, t3;
    if constexpr ((Z & 4) != 0) { t0 = a0; t1 = a0; }
    else if constexpr ((Z & 1) != 0) { t0 = a2; t1 = cneg(a2); }
    else { t0 = cadd(a0, a2); t1 = csub(a0, a2); }
    if constexpr ((Z & 8) != 0) { t2 = a1; t3 = mul_i<DIR>(a1); }
    else if constexpr ((Z & 2) != 0) { t2 = a3; t3 = mul_i<DIR>(cneg(a3)); }
    else { t2 = cadd(a1, a3); t3 = mul_i<DIR>(csub(a1, a3)); }
    a0 = cadd(t0, t2);
    a2 = csub(t0, t2);
    a1 = cadd(t1, t3);
    a3 = csub(t1, t3);
  }
}

// W8^k for k = 1,3 with sign DIR: e^{DIR*i*pi*k/4}
template <int DIR>
__device__ __forceinline__ float2 w8_1(float2 a) {  // a * e^{DIR i pi/4}
  const float h = 0.70710678118654752440f;
  return DIR < 0 ? make_float2((a.x + a.y) * h, (a.y - a.x) * h) : make_float2((a.x - a.y) * h, (a.y + a.x) * h);
}
template <int DIR>
__device__ __forceinline__ float2 w8_3(float2 a) {  // a * e^{DIR i 3pi/4}
  const float h = 0.70710678118654752440f;
  return DIR < 0 ? make_float2((a.y - a.x) * h, -(a.x + a.y) * h) : make_float2(-(a.x + a.y) * h, (a.x - a.y) * h);
}

// Z: bit j set when input j is zero; O: bit k set when output k is needed
// (the others are left unspecified).
template <int DIR, int Z = 0, int O = 0xff>
__device__ __forceinline__ void dft8(float2* a) {
  // decimation in time: even/odd halves, then combine with W8^k
  constexpr int ZE = ((Z >> 0) & 1) | (((Z >> 2) & 1) << 1) | (((Z >> 4) & 1) << 2) | (((Z >> 6) & 1) << 3);
  constexpr int ZO = ((Z >> 1) & 1) | (((Z >> 3) & 1) << 1) | (((Z >> 5) & 1) << 2) | (((Z >> 7) & 1) << 3);
  float2 e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
  float2 o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
  dft4z<DIR, ZE>(e0, e1, e2, e3);
  dft4z<DIR, ZO>(o0, o1, o2, o3);
  o1 = w8_1<DIR>(o1);
  o2 = mul_i<DIR>(o2);
  o3 = w8_3<DIR>(o3);
  if constexpr ((O & 0x01) != 0) a[0] = cadd(e0, o0);
  if constexpr ((O & 0x10) != 0) a[4] = csub(e0, o0);
  if constexpr ((O & 0x02) != 0) a[1] = cadd(e1, o1);
  if constexpr ((O & 0x20) != 0) a[5] = csub(e1, o1);
  if constexpr ((O & 0x04) != 0) a[2] = cadd(e2, o2);
  if constexpr ((O & 0x40) != 0) a[6] = csub(e2, o2);
  if constexpr ((O & 0x08) != 0) a[3] = cadd(e3, o3);
  if constexpr ((O & 0x80) != 0) a[7] = csub(e3, o3);
}

template <int DIR, int Z = 0>
__device__ __forceinline__ void dft16(float2* a) {
  // 16 = 4 x 4: columns of stride 4, twiddles W16^{n1 k2}, rows of length 4
  const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f;
  const float h = 0.70710678118654752440f;
  float2 b[16];
  auto colz = [](int n1) constexpr {
    return ((Z >> n1) & 1) | (((Z >> (n1 + 4)) & 1) << 1) | (((Z >> (n1 + 8)) & 1) << 2) | (((Z >> (n1 + 12)) & 1) << 3);
  };
  auto col = [&](auto n1c) {
    constexpr int n1 = decltype(n1c)::value;
    float2 x0 = a[n1], x1 = a[n1 + 4], x2 = a[n1 + 8], x3 = a[n1 + 12];
    dft4z<DIR, colz(n1)>(x0, x1, x2, x3);
    b[n1 * 4 + 0] = x0; b[n1 * 4 + 1] = x1; b[n1 * 4 + 2] = x2; b[n1 * 4 + 3] = x3;
  };
  col(std::integral_constant<int, 0>{});
  col(std::integral_constant<int, 1>{});
  col(std::integral_constant<int, 2>{});
  col(std::integral_constant<int, 3>{});
  // twiddle b[n1*4+k2] *= W16^{n1*k2}
  const float sg = (float)DIR;
  const float2 w1 = make_float2(c1, sg * s1), w2 = make_float2(h, sg * h), w3 = make_float2(s1, sg * c1);
  const float2 w6 = make_float2(-h, sg * h), w9 = make_float2(-c1, -sg * s1);
  b[1 * 4 + 1] = cmul(b[5], w1);
  b[1 * 4 + 2] = cmul(b[6], w2);
  b[1 * 4 + 3] = cmul(b[7], w3);
  b[2 * 4 + 1] = cmul(b[9], w2);
  b[2 * 4 + 2] = mul_i<DIR>(b[10]);
  b[2 * 4 + 3] = cmul(b[11], w6);
  b[3 * 4 + 1] = cmul(b[13], w3);
  b[3 * 4 + 2] = cmul(b[14], w6);
  b[3 * 4 + 3] = cmul(b[15], w9);
#pragma unroll
  for (int k2 = 0; k2 < 4; k2++) {
    float2 x0 = b[k2], x1 = b[4 + k2], x2 = b[8 + k2], x3 = b[12 + k2];
    dft4<DIR>(x0, x1, x2, x3);
    a[k2] = x0; a[k2 + 4] = x1; a[k2 + 8] = x2; a[k2 + 12] = x3;
  }
}

// Z, O: zero inputs and needed outputs (dft8 prunes both, dft16 the inputs).
template <int R, int DIR, int Z = 0, int O = -1>
__device__ __forceinline__ void dft(float2* a) {
  if constexpr (R == 2) dft2<DIR>(a[0], a[1]);
  else if constexpr (R == 4) dft4z<DIR, Z & 15>(a[0], a[1], a[2], a[3]);
  else if constexpr (R == 8) dft8<DIR, Z & 0xff, O & 0xff>(a);
  else if constexpr (R == 16) dft16<DIR, Z & 0xffff>(a);
}

// ---- radix plans -------------------------------------------------------------
// E: values per thread; T = N/E threads per transform (one wavefront up to
// N = 1024, the whole 256-thread block for 2048 and 4096); NP passes of radix
// R0..R3 (each divides E).  At 4096, {8,8,8,8} over 512 threads (about half
// the registers per thread, twice the waves per CU) measured slower than
// {16,16,16} over 256: forward rows by 27 %, reverse columns by 57 %, forward
// columns 3 % faster.
template <int N> struct Plan;
template <> struct Plan<64>   { static constexpr int E = 4,  NP = 3, R0 = 4,  R1 = 4,  R2 = 4,  R3 = 1; };
template <> struct Plan<128>  { static constexpr int E = 8,  NP = 3, R0 = 8,  R1 = 8,  R2 = 2,  R3 = 1; };
template <> struct Plan<256>  { static constexpr int E = 8,  NP = 3, R0 = 8,  R1 = 8,  R2 = 4,  R3 = 1; };
template <> struct Plan<512>  { static constexpr int E = 8,  NP = 3, R0 = 8,  R1 = 8,  R2 = 8,  R3 = 1; };
template <> struct Plan<1024> { static constexpr int E = 16, NP = 3, R0 = 16, R1 = 8,  R2 = 8,  R3 = 1; };
template <> struct Plan<2048> { static constexpr int E = 8,  NP = 4, R0 = 8,  R1 = 8,  R2 = 8,  R3 = 4; };
template <> struct Plan<4096> { static constexpr int E = 16, NP = 3, R0 = 16, R1 = 16, R2 = 16, R3 = 1; };

// LDS placement of element i of a length-N transform, in one of two layouts.
// On gfx950 a ds_read_b64 serves 32 lanes at a time over 32 eight-byte bank
// pairs and a ds_write_b64 16 lanes over 16.
//  SWZ (default up to N = 512): the low four index bits are XORed with a
//    function of bits 4.. (i ^ f_N(i >> 4)); every read and write pattern of
//    fft_wave below then hits distinct bank pairs (sizes 64 and 128 also need
//    the slot strides of lds_slot, as several transforms share a wavefront).
//    The write addresses are no longer immediate offsets from one base, so
//    the compiler keeps one address register per access.
//  PAD: one float2 of padding every 16 (i + (i >> 4)): immediate offsets from
//    a per-thread base, at a residual 1.3-1.6-way conflict rate; used beyond
//    512 and wherever registers are the scarcer resource.
template <int N> constexpr bool kSwz = N <= 512;
template <int N, bool SWZ = kSwz<N>>
__device__ __forceinline__ int lidx(int i) {
  if constexpr (!SWZ) {
    return i + (i >> 4);
  } else {
    const int j = i >> 4;
    int f;
    if constexpr (N == 64) f = ((j & 1) ? 9 : 0) ^ (((j ^ (j >> 1)) & 1) ? 6 : 0);
    else f = (j & 7) | (((j >> 2) & 1) << 3);
    return i ^ f;
  }
}
template <int N, bool SWZ = kSwz<N>>
constexpr int lds_slot() { return SWZ ? (N == 64 ? 80 : (N == 128 ? 144 : N)) : N + N / 16; }

// Orders this lane's LDS reads before later LDS writes of other lanes of the
// same wavefront and vice versa (all lanes of a transform live in one wave).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS ordering between the threads of one length-N transform: a wavefront
// barrier when the transform lives in one wavefront, else a block barrier
// (the block is the transform; every thread of it reaches every call).
template <int N>
__device__ __forceinline__ void xform_sync() {
  if constexpr (N / Plan<N>::E <= 64) wave_lds_sync();
  else __syncthreads();
}

// W_N^m = exp(-2 pi i m/N), 0 <= m < N, computed instead of loaded: m/N turns is
// reduced exactly to the nearest quarter turn q/4 and a remainder a with |a| <= pi/4
// (a = (4m - qN) * pi/(2N): an exact integer times one constant, one rounding),
// whose sine and cosine come from their Taylor series (degrees 9 and 10, truncation
// below 2e-9); the quarter turn permutes and negates them.
template <int N>
__device__ __forceinline__ float2 twiddle_at(int m) {
  const int q = (4 * m + N / 2) / N;
  const float a = (float)(4 * m - q * N) * (float)(3.14159265358979323846 / (2.0 * N));
  const float a2 = a * a;
  const float sa = a + a * a2 * fmaf(a2, fmaf(a2, fmaf(a2, 1.0f / 362880.0f, -1.0f / 5040.0f), 1.0f / 120.0f),
                                     -1.0f / 6.0f);
  const float ca = fmaf(a2, fmaf(a2, fmaf(a2, fmaf(a2, fmaf(a2, -1.0f / 3628800.0f, 1.0f / 40320.0f),
                                                   -1.0f / 720.0f), 1.0f / 24.0f), -0.5f), 1.0f);
  // (cos, sin) of q pi/2 + a
  const int qq = q & 3;
  const float c = qq == 0 ? ca : (qq == 1 ? -sa : (qq == 2 ? -ca : sa));
  const float sn = qq == 0 ? sa : (qq == 1 ? ca : (qq == 2 ? -sa : -ca));
  return make_float2(c, -sn);
}

// Per-thread twiddles of the twiddled passes.  For small three-pass plans they
// are loaded once into registers (init) and reused by every transform the
// thread performs; larger plans compute them on every pass (twiddle_at: no
// table loads and their latency inside the passes, and no registers held).
template <int N>
struct Twiddles {
  using P = Plan<N>;
  static constexpr int T = N / P::E;
  static constexpr int NS1 = P::R0, NS2 = P::R0 * P::R1;
  static constexpr int NB1 = P::E / P::R1, NB2 = P::E / P::R2;
  static constexpr int C1 = NB1 * (P::R1 - 1), C2 = NB2 * (P::R2 - 1);
  static constexpr bool REG = P::NP == 3 && (C1 + C2) <= 16;
  float2 w1[REG ? C1 : 1], w2[REG ? C2 : 1];
  const float2* tab;  // the table W_N^m, m < N
  // REG: both twiddled passes' factors from the table, held in registers;
  // otherwise each pass computes its own (twiddle_at) or loads them from tab
  __device__ __forceinline__ void init(const float2* __restrict__ tw, int t) {
    tab = tw;
    if constexpr (REG) {
#pragma unroll
      for (int b = 0; b < NB1; b++) {
        const int jm = (t + T * b) % NS1;
#pragma unroll
        for (int k = 1; k < P::R1; k++) w1[b * (P::R1 - 1) + k - 1] = tw[k * jm * (N / (NS1 * P::R1))];
      }
#pragma unroll
      for (int b = 0; b < NB2; b++) {
        const int jm = (t + T * b) % NS2;
#pragma unroll
        for (int k = 1; k < P::R2; k++) w2[b * (P::R2 - 1) + k - 1] = tw[k * jm * (N / (NS2 * P::R2))];
      }
    }
  }
};

// One Stockham pass.  FIRST: input from registers (natural layout).  LAST:
// output to registers (natural layout).  Otherwise through the LDS slot.
// wr: this pass's register twiddles, or nullptr: the pass computes them (TWC)
// or loads them from tab, the table W_N^m = exp(-2 pi i m/N), m < N.
// ZIN (first pass) / NEED (last pass): masks over the natural-layout register
// slots m (element t + T*m) that are zero on input / needed on output; used when
// one butterfly covers all E slots (R == E), ignored otherwise.
template <int N, int R, int NS, int DIR, bool FIRST, bool LAST, bool SWZ, int ZIN = 0, int NEED = -1, bool TWC = true>
__device__ __forceinline__ void stockham_pass(float2* v, float2* slot, const float2* wr,
                                              const float2* __restrict__ tab, int t) {
  constexpr int E = Plan<N>::E;
  constexpr int T = N / E;
  constexpr int NB = E / R;
  float2 a[NB][R];
  if constexpr (FIRST) {
#pragma unroll
    for (int b = 0; b < NB; b++)
#pragma unroll
      for (int k = 0; k < R; k++) a[b][k] = v[b + k * NB];
  } else {
#pragma unroll
    for (int b = 0; b < NB; b++)
#pragma unroll
      for (int k = 0; k < R; k++) a[b][k] = slot[lidx<N, SWZ>(t + T * b + k * (N / R))];
  }
  if constexpr (NS > 1) {
    // computed or table twiddles are formed where they are used: the empty asm
    // makes the index offset opaque, so they are not hoisted out of the callers'
    // loops into registers held across the whole kernel (and the table stays in
    // the global address space: global loads, not flat ones that also hold up
    // the LDS waits)
    int z = 0;
    if (!wr) asm volatile("" : "+s"(z));
#pragma unroll
    for (int b = 0; b < NB; b++) {
      const int jm = (t + T * b) % NS;
      // w^k of w = W_N^{jm N/(NS R)}: computed or loaded for k < 4 and k a
      // multiple of 4, the others as w^(k & ~3) * w^(k & 3) (6 evaluations or
      // loads for a radix-16 pass)
      float2 pw[R];
      if (!wr) {
#pragma unroll
        for (int k = 1; k < R; k++)
          if ((k & 3) == 0 || k < 4) {
            const int m = z + k * jm * (N / (NS * R));
            pw[k] = TWC ? twiddle_at<N>(m) : tab[m];
          }
#pragma unroll
        for (int k = 5; k < R; k++)
          if (k & 3) pw[k] = cmul(pw[k & ~3], pw[k & 3]);
      }
#pragma unroll
      for (int k = 1; k < R; k++) {
        const float2 w = wr ? wr[b * (R - 1) + k - 1] : pw[k];
        a[b][k] = DIR < 0 ? cmul(a[b][k], w) : cmulc(a[b][k], w);
      }
    }
  }
#pragma unroll
  for (int b = 0; b < NB; b++) {
    if constexpr (NB == 1) dft<R, DIR, FIRST ? ZIN : 0, LAST ? NEED : -1>(a[b]);
    else dft<R, DIR>(a[b]);
  }
  if constexpr (LAST) {
#pragma unroll
    for (int b = 0; b < NB; b++)
#pragma unroll
      for (int k = 0; k < R; k++) v[b + k * NB] = a[b][k];
  } else {
    xform_sync<N>();  // all earlier reads of the slot precede these writes
#pragma unroll
    for (int b = 0; b < NB; b++) {
      const int j = t + T * b;
      const int base = (j / NS) * NS * R + (j % NS);
#pragma unroll
      for (int k = 0; k < R; k++) slot[lidx<N, SWZ>(base + k * NS)] = a[b][k];
    }
    xform_sync<N>();
  }
}

// Register slots m (element t + T*m) that lie, for every thread, strictly
// inside the gap (N/3, N - N/3) of the 2/3 band limit of propagate_slow
// (stem3/stemlib.cpp:2815-2849): the band keeps |index| <= floor(N/3) on each
// axis, so these slots are zero in k-space after the propagator.
template <int N>
constexpr int gap_mask() {
  constexpr int E = Plan<N>::E, T = N / E, A = N / 3;
  int m = 0;
  for (int j = 0; j < E && j < 31; j++)
    if (T * j > A && T * j + T - 1 < N - A) m |= 1 << j;
  return m;
}

// Full length-N transform of the thread-distributed vector v (natural layout
// in and out).  slot: this transform's LDS region of lds_slot<N, SWZ>() float2.
// ZIN: slots known to be zero on input; NEED: slots needed on output.
template <int N, int DIR, bool SWZ = kSwz<N>, int ZIN = 0, int NEED = -1, bool TWC = true>
__device__ __forceinline__ void fft_wave(float2* v, float2* slot, const Twiddles<N>& tw, int t) {
  using P = Plan<N>;
  constexpr int S1 = P::R0;
  constexpr int S2 = P::R0 * P::R1;
  constexpr bool REG = Twiddles<N>::REG;
  stockham_pass<N, P::R0, 1, DIR, true, false, SWZ, ZIN, -1, TWC>(v, slot, nullptr, tw.tab, t);
  stockham_pass<N, P::R1, S1, DIR, false, false, SWZ, 0, -1, TWC>(v, slot, REG ? tw.w1 : nullptr, tw.tab, t);
  if constexpr (P::NP == 3) {
    stockham_pass<N, P::R2, S2, DIR, false, true, SWZ, 0, NEED, TWC>(v, slot, REG ? tw.w2 : nullptr, tw.tab, t);
  } else {
    stockham_pass<N, P::R2, S2, DIR, false, false, SWZ, 0, -1, TWC>(v, slot, nullptr, tw.tab, t);
    stockham_pass<N, P::R3, S2 * P::R2, DIR, false, true, SWZ, 0, NEED, TWC>(v, slot, nullptr, tw.tab, t);
  }
}

}  // namespace qs
