// qs_mix.hpp -- wavefront-resident mixed-radix sweeps with compiled radix plans.
//
// The runtime-planned kernels of qs_generic.hip share every Stockham pass among all
// threads of a block: two block barriers per pass, and threads idle wherever the
// radix leaves the butterfly count short of the block (half of them in the radix-25
// pass of 400).  For a length with a compiled plan of two to four passes, a
// transform is owned by T = N/E lanes of one wavefront, each holding E values in
// registers (element t + T m in v[m], the layout of coalesced row loads); the
// exchanges between the passes go through the transform's own LDS slot and need
// wavefront-level ordering only.  64/T transforms share a wavefront; lanes past them
// mirror the last transform's lanes (same inputs, same values written to the same LDS
// words) and store nothing.  State layout ([probe][x][ny], in-band ky columns only),
// modes and outputs are those of k_gen_rows and k_gen_cols (runMulsSTEM's slice step,
// stem3/stemlib.cpp:2405-2443: transmit, FFT, propagate_slow with its band limit,
// collectIntensity's detector sums, inverse FFT).
#pragma once
#include <hip/hip_runtime.h>

#include "qs_dft.hpp"
#include "qs_fft.hpp"
#include "qs_internal.hpp"

namespace qs {

// Compiled plans: E values per lane, NP passes of radix R0, R1 (, R2, R3), each
// dividing E (so every lane owns E/R butterflies of a pass); LD float2 per transform
// slot.
template <int N> struct MixPlan { static constexpr int E = 0; };
template <> struct MixPlan<288> { static constexpr int E = 24, NP = 3, R0 = 24, R1 = 6, R2 = 2, R3 = 1, LD = 297; };
template <> struct MixPlan<300> { static constexpr int E = 30, NP = 3, R0 = 30, R1 = 5, R2 = 2, R3 = 1, LD = 309; };
template <> struct MixPlan<320> { static constexpr int E = 20, NP = 3, R0 = 20, R1 = 4, R2 = 4, R3 = 1, LD = 331; };
template <> struct MixPlan<360> { static constexpr int E = 30, NP = 3, R0 = 30, R1 = 6, R2 = 2, R3 = 1, LD = 371; };
template <> struct MixPlan<384> { static constexpr int E = 24, NP = 3, R0 = 24, R1 = 8, R2 = 2, R3 = 1, LD = 397; };
template <> struct MixPlan<400> { static constexpr int E = 20, NP = 2, R0 = 20, R1 = 20, R2 = 1, R3 = 1, LD = 420; };
template <> struct MixPlan<441> { static constexpr int E = 21, NP = 3, R0 = 21, R1 = 7, R2 = 3, R3 = 1, LD = 455; };
template <> struct MixPlan<448> { static constexpr int E = 28, NP = 3, R0 = 28, R1 = 4, R2 = 4, R3 = 1, LD = 463; };
template <> struct MixPlan<500> { static constexpr int E = 10, NP = 3, R0 = 10, R1 = 10, R2 = 5, R3 = 1, LD = 517; };
template <> struct MixPlan<540> { static constexpr int E = 30, NP = 3, R0 = 30, R1 = 6, R2 = 3, R3 = 1, LD = 557; };
template <> struct MixPlan<600> { static constexpr int E = 30, NP = 3, R0 = 30, R1 = 10, R2 = 2, R3 = 1, LD = 619; };
template <> struct MixPlan<640> { static constexpr int E = 20, NP = 4, R0 = 20, R1 = 4, R2 = 4, R3 = 2, LD = 661; };
template <> struct MixPlan<768> { static constexpr int E = 24, NP = 3, R0 = 24, R1 = 8, R2 = 4, R3 = 1, LD = 793; };
template <> struct MixPlan<900> { static constexpr int E = 30, NP = 3, R0 = 30, R1 = 15, R2 = 2, R3 = 1, LD = 929; };
template <> struct MixPlan<1280> { static constexpr int E = 20, NP = 4, R0 = 20, R1 = 4, R2 = 4, R3 = 4, LD = 1321; };
template <> struct MixPlan<1350> { static constexpr int E = 30, NP = 3, R0 = 30, R1 = 15, R2 = 3, R3 = 1, LD = 1393; };
template <> struct MixPlan<1536> { static constexpr int E = 24, NP = 3, R0 = 24, R1 = 8, R2 = 8, R3 = 1, LD = 1585; };
template <> struct MixPlan<1800> { static constexpr int E = 30, NP = 3, R0 = 30, R1 = 30, R2 = 2, R3 = 1, LD = 1857; };

constexpr int kMixWaves = 4;  // wavefronts per block
template <int N> constexpr int mix_slots() { return 64 / (N / MixPlan<N>::E); }  // transforms per wavefront
template <int N> constexpr int mix_cpb() { return kMixWaves * mix_slots<N>(); }  // columns per column block

// LDS position of the value with Stockham index i between two passes.
//  Two passes (N = R0 R1): k (N/R0 + 1) + j for i = j R0 + k (butterfly j of the first
//    pass, its output k): the first pass's outputs transposed with one float2 of
//    padding per row.  At 400 (20 x 20, three transforms per wavefront, slot stride
//    420 = 4 mod 32) the first pass's writes (lanes of a transform on consecutive j)
//    and the second pass's reads (stride 21, odd) fall on distinct eight-byte bank
//    pairs in every 16-lane write and 32-lane read group, across the three
//    transforms of a wavefront too, and four blocks of twelve slots fit the CU's LDS.
//  Three or four passes: i + i/32 (one float2 of padding every 32).
template <int N>
__device__ __forceinline__ int mix_pos(int i) {
  using P = MixPlan<N>;
  if constexpr (P::NP == 2) return (i % P::R0) * (N / P::R0 + 1) + i / P::R0;
  else return i + (i >> 5);
}

// One Stockham pass of radix R over span NS (product of the earlier radices): lane t's
// butterflies j = t + T b (b < E/R) read elements j + k N/R (FIRST: from v, where
// element t + T m is v[m]), multiply by W_N^(k (j mod NS) N/(NS R)) (tw = W_N^m, m < N),
// and write (j / NS) NS R + j mod NS + k NS (LAST: into v, same layout).
template <int N, int R, int NS, int DIR, bool FIRST, bool LAST>
__device__ __forceinline__ void mix_pass(float2* v, float2* slot, const float2* __restrict__ tw, int t) {
  using P = MixPlan<N>;
  constexpr int E = P::E, T = N / E, NB = E / R;
  static_assert(E % R == 0, "every lane owns whole butterflies");
  float2 a[NB][R];
#pragma unroll
  for (int b = 0; b < NB; b++) {
#pragma unroll
    for (int k = 0; k < R; k++) {
      if constexpr (FIRST) a[b][k] = v[b + k * NB];
      else a[b][k] = slot[mix_pos<N>(t + T * b + k * (N / R))];
    }
  }
  if constexpr (NS > 1) {
#pragma unroll
    for (int b = 0; b < NB; b++) {
      // opaque per call: the twiddle loads are neither hoisted out of the callers'
      // loops nor shared between transforms (registers held)
      int e = ((t + T * b) % NS) * (N / (NS * R));
      asm volatile("" : "+v"(e));
      // w^k loaded for k < 4 and k a multiple of 4, the others as w^(k & ~3) w^(k & 3)
      float2 pw[R];
#pragma unroll
      for (int k = 1; k < R; k++)
        if (k < 4 || (k & 3) == 0) pw[k] = tw[k * e];
#pragma unroll
      for (int k = 5; k < R; k++)
        if (k & 3) pw[k] = cmul(pw[k & ~3], pw[k & 3]);
#pragma unroll
      for (int k = 1; k < R; k++) a[b][k] = DIR < 0 ? cmul(a[b][k], pw[k]) : cmulc(a[b][k], pw[k]);
    }
  }
#pragma unroll
  for (int b = 0; b < NB; b++) dft_any<R, DIR>(a[b]);
  if constexpr (LAST) {
    static_assert(N / R == T * NB, "the last pass's outputs are the natural layout");
#pragma unroll
    for (int b = 0; b < NB; b++)
#pragma unroll
      for (int k = 0; k < R; k++) v[b + k * NB] = a[b][k];
  } else {
    wave_lds_sync();  // the slot's earlier reads are done
#pragma unroll
    for (int b = 0; b < NB; b++) {
      const int j = t + T * b;
      const int base = (j / NS) * NS * R + j % NS;
#pragma unroll
      for (int k = 0; k < R; k++) slot[mix_pos<N>(base + k * NS)] = a[b][k];
    }
    wave_lds_sync();
  }
}

// Length-N transform of v (element t + T m in v[m], in and out) through the slot.
template <int N, int DIR>
__device__ __forceinline__ void mix_fft(float2* v, float2* slot, const float2* __restrict__ tw, int t) {
  using P = MixPlan<N>;
  if constexpr (P::NP == 2) {
    static_assert(P::R0 * P::R1 == N, "two-pass plan");
    mix_pass<N, P::R0, 1, DIR, true, false>(v, slot, tw, t);
    mix_pass<N, P::R1, P::R0, DIR, false, true>(v, slot, tw, t);
  } else if constexpr (P::NP == 3) {
    static_assert(P::R0 * P::R1 * P::R2 == N, "three-pass plan");
    mix_pass<N, P::R0, 1, DIR, true, false>(v, slot, tw, t);
    mix_pass<N, P::R1, P::R0, DIR, false, false>(v, slot, tw, t);
    mix_pass<N, P::R2, P::R0 * P::R1, DIR, false, true>(v, slot, tw, t);
  } else {
    static_assert(P::NP == 4 && P::R0 * P::R1 * P::R2 * P::R3 == N, "four-pass plan");
    mix_pass<N, P::R0, 1, DIR, true, false>(v, slot, tw, t);
    mix_pass<N, P::R1, P::R0, DIR, false, false>(v, slot, tw, t);
    mix_pass<N, P::R2, P::R0 * P::R1, DIR, false, false>(v, slot, tw, t);
    mix_pass<N, P::R3, P::R0 * P::R1 * P::R2, DIR, false, true>(v, slot, tw, t);
  }
}

// Buffer-resource access (MUBUF): a per-lane byte offset plus a wavefront-uniform one.
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

// Rows (length ny = N): transform slot g of the launch advances rows g*a.rpb ..
// g*a.rpb + a.rpb - 1 (row = probe * nx + x).  Modes as k_gen_rows: 0 probe input,
// 1 state input (y-IFFT first), 2 exit (y-IFFT into a.exitw, no transmit).
template <int N>
__global__ __launch_bounds__(64 * kMixWaves) __attribute__((amdgpu_waves_per_eu(4))) void k_mix_rows(GenRowArgs a) {
  using P = MixPlan<N>;
  constexpr int E = P::E, T = N / E, S = mix_slots<N>(), LD = P::LD;
  static_assert(E <= 32, "the band mask holds one bit per element");
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
// per probe) has transform slots ci = CPB b .. CPB b + CPB - 1 (CPB = mix_cpb<N>) and
// slot ci takes band column a.tiles[ci] (a.tc of them, ascending).  The block's
// columns are staged through the slots in LDS: each state row's piece of the block's
// columns is read, and written back, by CPB consecutive lanes (one contiguous piece
// per row and instruction; read straight by the transforms' lanes, every lane would
// touch T rows per access and the rows' lines would be fetched and written back in
// pieces of a few columns).  Blocks of kMixGroup consecutive b are placed on one XCD
// back to back, where they share its L2 the rows' 128-byte lines that their columns
// split (blocks are dispatched round-robin over the 8 XCDs).  x-FFT, propagator, band
// limit, detector sums into partial[p][b][kMaxDet + 1] and the diffraction pattern as
// k_gen_cols, x-IFFT.  ND: most detectors (a.ndet <= ND), so that only ND + 1 sums are
// held in registers.
constexpr int kMixGroup = 4;
template <int N, int ND>
__global__ __launch_bounds__(64 * kMixWaves) __attribute__((amdgpu_waves_per_eu(ND <= 4 ? 4 : 3))) void k_mix_cols(GenColArgs a) {
  using P = MixPlan<N>;
  constexpr int E = P::E, T = N / E, S = mix_slots<N>(), LD = P::LD, CPB = mix_cpb<N>();
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
  const int nb = (a.tc + CPB - 1) / CPB;  // blocks per probe
  const int p = lb / nb, b = lb - p * nb, ny = a.ny;
  const int ci = (b * kMixWaves + wv) * S + sub;
  const bool ok = act && ci < a.tc;
  const int ky = a.tiles[min(ci, a.tc - 1)];
  // this probe's state; element x of column ky at byte (x ny + ky) 8 (below 2^31: nx, ny <= 2048)
  const auto rs = mix_rsrc(a.state + (size_t)(a.probe0 + p) * N * ny, (long)N * ny * 8);
  // staging: lane (sc, sx) moves block column sc (slot sc) of rows sx + SR i
  constexpr int SR = 64 * kMixWaves / CPB;
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
          int sx2 = kx - a.det_sx[d], sy2 = ky - a.det_sy[d];
          sx2 += sx2 < 0 ? N : 0;
          sy2 += sy2 < 0 ? ny : 0;
          k2 = a.kx2[sx2] + a.ky2[sy2];
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

template <int N>
void launch_mix_rows_t(const GenRowArgs& a, hipStream_t st) {
  const long nrows = (long)a.nprobe * a.nx;
  const long per = (long)kMixWaves * mix_slots<N>() * a.rpb;
  hipLaunchKernelGGL(k_mix_rows<N>, dim3((unsigned)((nrows + per - 1) / per)), dim3(64 * kMixWaves), 0, st, a);
}

template <int N>
void launch_mix_cols_t(const GenColArgs& a, int nblocks, hipStream_t st) {
  const dim3 grid((unsigned)(nblocks * a.nprobe)), block(64 * kMixWaves);
  if (a.ndet <= 4) hipLaunchKernelGGL((k_mix_cols<N, 4>), grid, block, 0, st, a);
  else hipLaunchKernelGGL((k_mix_cols<N, kMaxDet>), grid, block, 0, st, a);
}

}  // namespace qs
