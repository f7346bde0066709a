// qs_dft.hpp -- small DFTs in registers (natural order in and out) for the
// mixed-radix Stockham passes: radix 3, 5, 7 directly, composite radices as two
// register stages with compile-time twiddles, powers of two from qs_fft.hpp.
// Direction as in qs_fft.hpp: DIR = -1 forward (exp(-2 pi i jk/R)), +1 backward.
#pragma once
#include "qs_fft.hpp"

namespace qs {

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

// Composite radices as R1 x R2 (the first stage over R1-point DFTs).
template <int R, int DIR>
__device__ __forceinline__ void dft_any(float2* a) {
  if constexpr (R == 3) dft3<DIR>(a);
  else if constexpr (R == 5) dft5<DIR>(a);
  else if constexpr (R == 7) dft7<DIR>(a);
  else if constexpr (R == 6) dft_2stage<2, 3, DIR>(a);
  else if constexpr (R == 9) dft_2stage<3, 3, DIR>(a);
  else if constexpr (R == 10) dft_2stage<2, 5, DIR>(a);
  else if constexpr (R == 12) dft_2stage<4, 3, DIR>(a);
  else if constexpr (R == 14) dft_2stage<2, 7, DIR>(a);
  else if constexpr (R == 15) dft_2stage<3, 5, DIR>(a);
  else if constexpr (R == 18) dft_2stage<2, 9, DIR>(a);
  else if constexpr (R == 20) dft_2stage<4, 5, DIR>(a);
  else if constexpr (R == 21) dft_2stage<3, 7, DIR>(a);
  else if constexpr (R == 24) dft_2stage<8, 3, DIR>(a);
  else if constexpr (R == 25) dft_2stage<5, 5, DIR>(a);
  else if constexpr (R == 28) dft_2stage<4, 7, DIR>(a);
  else if constexpr (R == 30) dft_2stage<5, 6, DIR>(a);
  else {
    static_assert(R == 2 || R == 4 || R == 8 || R == 16, "radix without a DFT");
    dft<R, DIR>(a);
  }
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


}  // namespace qs
