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
// Threads per block of the row-type kernels for transform length N: 256, or
// the transform's own thread count when larger (one transform per block).
template <int N> constexpr int row_block() { return (N / Plan<N>::E) > kBlock ? N / Plan<N>::E : kBlock; }


// ============================================================================
// Buffer-resource (MUBUF) access with hardware range checking: a load at an
// offset at or past the resource's byte size returns zero and a store there is
// dropped.  kOut is past every resource built here (sizes stay below 2^31).
using rsrc_t = __amdgpu_buffer_rsrc_t;
constexpr unsigned kOut = 0x80000000u;

__device__ __forceinline__ rsrc_t make_rsrc(const void* base, long bytes) {
  const int n = bytes < 0x7fffffffL ? (int)bytes : 0x7fffffff;
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, n, 0x00020000);
}
__device__ __forceinline__ float2 buf_load_c64(rsrc_t r, unsigned off) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ void buf_store_c64(float2 v, rsrc_t r, unsigned off) {
  using U2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0u, 0, 0));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v), r, off, 0, 0);
}
// 16-byte store with the whole offset in the VGPR offset and a zero soffset.  A MUBUF
// store of more than 8 bytes reads its data VGPRs after issue, and the compiler puts
// wait states before a later VALU write of those VGPRs only when the soffset is not an
// SGPR; with an SGPR soffset it scheduled such writes directly after the store, and on
// gfx950 the store then sometimes wrote the new values (LDS addresses in place of the
// wave data, a few elements per launch, differing between runs).
__device__ __forceinline__ void buf_store_f4(float4 v, rsrc_t r, unsigned off) {
  using U4 = decltype(__builtin_amdgcn_raw_buffer_load_b128(r, 0u, 0, 0));
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(U4, v), r, off, 0, 0);
}

// Byte offset of ky column iy inside a band-compact row, or kOut when iy lies in
// the zero gap (band_a, ny-band_b) between the two halves of the band;
// gap = ny - band_b - band_a - 1.
__device__ __forceinline__ unsigned band_off(int iy, int band_a, int gap) {
  const int c = iy > band_a ? iy - gap : iy;
  return (unsigned)(iy - band_a - 1) < (unsigned)gap ? kOut : (unsigned)c * 8u;
}

__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// ============================================================================
// sin and cos of a phase x (|x| well below 2^16 rad): Cody-Waite reduction by
// pi/2 in two FMA steps (pi/2 = H + L, H = float(pi/2); x - kH is exact under the
// fused multiply-add), then the Taylor series of sin and cos on |r| <= pi/4 to
// degrees 9 and 10 (truncation below 2e-9) and the quadrant k mod 4.
__device__ __forceinline__ void sincos_phase(float x, float& s, float& c) {
  const float k = rintf(x * 0.63661977236758134f);  // 2/pi
  float r = fmaf(-k, 1.57079637050628662f, x);       // H
  r = fmaf(-k, -4.37113900018624283e-8f, r);         // L = pi/2 - H
  const float r2 = r * r;
  const float ps = r + r * r2 * fmaf(r2, fmaf(r2, fmaf(r2, 1.0f / 362880.0f, -1.0f / 5040.0f), 1.0f / 120.0f),
                                     -1.0f / 6.0f);
  const float pc = fmaf(r2, fmaf(r2, fmaf(r2, fmaf(r2, fmaf(r2, -1.0f / 3628800.0f, 1.0f / 40320.0f),
                                                   -1.0f / 720.0f), 1.0f / 24.0f), -0.5f), 1.0f);
  const int q = (int)k & 3;
  const float s0 = (q & 1) ? pc : ps, c0 = (q & 1) ? ps : pc;
  s = (q & 2) ? -s0 : s0;
  c = ((q + 1) & 2) ? -c0 : c0;
}

// Columns per tile of the column sweeps at length N: at 512 one whole 128-byte line of
// each state row (16 columns).
template <int N> constexpr int col_tc() { return N == 512 ? 16 : (N <= 1024 ? 8 : (N == 2048 ? 4 : 2)); }

// Largest band-compact row pitch (complex values) at length N: the 2/3 band keeps at
// most 2 floor(N/3) + 1 columns, padded to whole column tiles and then to 16.
template <int N>
constexpr int band_pitch_max() {
  constexpr int tc = col_tc<N>();
  constexpr int kc = 2 * (N / 3) + 1;
  return ((kc + tc - 1) / tc * tc + 15) / 16 * 16;
}

// Sweep A: rows.
// mode 0: input is the shared real-space probe (first slice of a fresh run)
// mode 1: input is the band-compact state (a.state_in if set, else a.state); row IFFT first
// mode 2: exit: row IFFT of the state, write the real-space wave, no transmit
// mode 3: input is a full real-space row of each probe (a.probe + probe*nx*NY);
//         row FFT only, no transmit (first sweep of a band limit)
// mode 4: as mode 2, but the row g is not stored: dL/dV = sigma Im(g conj(u)),
//         u = exp(i sigma V), is written to a.gv (last sweep of the band-limited
//         potential gradient)
// Each wavefront owns a chunk of TPW*a.rw consecutive rows of one probe
// (TPW = 64/T transforms side by side, a.rw steps; the chunk divides nx), so
// probe, window offset and row bases are wave-uniform scalars.  Twiddles and
// band offsets are set up once per wavefront; all loads of a step are issued
// before its first transform.
template <int NY, int MODE>
__global__ __launch_bounds__(row_block<NY>(), Plan<NY>::E <= 8 && NY / Plan<NY>::E <= 64 ? 4 : (NY == 4096 && MODE >= 3 ? 3 : 1)) void k_sweep_rows(SweepArgs a) {
  constexpr int E = Plan<NY>::E;
  constexpr int T = NY / E;
  constexpr int BLK = row_block<NY>();
  static_assert(T <= 64 || T == BLK, "a transform is a wavefront part or the whole block");
  constexpr int TPW = T <= 64 ? 64 / T : 1;   // transforms side by side per unit
  constexpr int UPB = T <= 64 ? BLK / 64 : 1;  // units (wavefronts, or the block) per block
  constexpr int SPB = BLK / T;       // transform slots per block
  constexpr bool BOFF = E <= 16;     // keep the band offsets in registers
  constexpr int GAP = gap_mask<NY>();  // slots in the band gap: never loaded, stored or computed
  __shared__ float2 lds[SPB * lds_slot<NY>()];
  const int lane = T <= 64 ? (threadIdx.x & 63) : threadIdx.x;
  const int sub = lane / T;
  const int t = lane % T;
  const int chunk = TPW * a.rw;
  const int unit = T <= 64 ? (threadIdx.x >> 6) : 0;
  const int row0 = wave_uniform((blockIdx.x * UPB + unit) * chunk);
  const int nrows = a.nprobe * a.nx;
  if (row0 >= nrows) return;         // whole wavefronts only (chunk divides nprobe*nx)
  const int p = row0 / a.nx;
  const int x0 = row0 - p * a.nx;    // first row of the chunk
  const int pg = a.probe0 + p;       // global probe index (ipos, state)
  float2* slot = lds + (threadIdx.x / T) * lds_slot<NY>();
  const size_t srow0 = ((size_t)pg * a.nx + x0) * a.kcp;
  const rsrc_t rs = make_rsrc(a.state + srow0, (long)(nrows - row0) * a.kcp * 8);
  // loads from a.state_in when given (the state of slice s-1 kept elsewhere)
  const rsrc_t rsi = a.state_in ? make_rsrc(a.state_in + srow0, (long)(nrows - row0) * a.kcp * 8) : rs;
  const unsigned rowb = (unsigned)a.kcp * 8u;
  const int gap = NY - a.band_b - a.band_a - 1;
  // per-lane byte offset of element m within step 0's rows (kOut outside the band);
  // step i adds the wave-uniform i*TPW*rowb as the scalar offset
  unsigned boff[BOFF ? E : 1];
  auto bo = [&](int m) -> unsigned {
    const unsigned o = band_off(t + T * m, a.band_a, gap);
    return o == kOut ? kOut : o + (unsigned)sub * rowb;
  };
  if constexpr (BOFF) {
#pragma unroll
    for (int m = 0; m < E; m++) boff[m] = bo(m);
  }
  // MODE 1 up to 512: the state rows of the next step are copied straight into this
  // wavefront's LDS staging area (buffer loads with an LDS destination, no registers)
  // while the current step is transformed; only the first step waits for its state.
  // A step's TPW band rows hold at most 384 complex values (three 1 KB copies): the
  // band keeps at most 2 floor(NY/3) + 1 columns, padded to whole tiles and 16.
  constexpr bool DMA = MODE == 1 && NY <= 512;
  constexpr int kStg = 384;
  static_assert(!DMA || TPW * band_pitch_max<NY>() <= kStg, "a step's band rows exceed the staging area");
  __shared__ float2 stg[DMA ? (BLK / 64) * kStg : 1];
  float2* const wstg = stg + (DMA ? wave_uniform(threadIdx.x >> 6) * kStg : 0);
  auto prefetch = [&](int i) {
    if constexpr (DMA) {
      // the whole offset in the VGPR offset, so that the resource's range check covers
      // it: the copy's tail past the last row reads zeros, and past the chunk's last
      // step (kOut) nothing is read
      const unsigned vo = (i < a.rw ? (unsigned)(i * TPW) * rowb : kOut) + 16u * lane;
#pragma unroll
      for (int j = 0; j < 3; j++)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rsi, (__attribute__((address_space(3))) void*)(wstg + 128 * j), 16,
                                                 vo + 1024u * j, 0, 0, 0);
    }
  };
  int2 ip = make_int2(0, 0);
  const float2* trc = a.trans;  // this probe's configuration, slice s
  if constexpr (MODE <= 1) {
    ip = a.ipos[pg];
    if (a.cfg) trc += (size_t)a.cfg[pg] * (size_t)a.cfg_stride;
  }
  const float2* tr0 = trc + (size_t)(x0 + ip.x) * a.potNy + ip.y;

  auto load = [&](int i, float2* dv, float2* dt) {
    const int xr = i * TPW + sub;  // row within the chunk
    if constexpr (MODE <= 1) {
      const float2* tr = tr0 + (size_t)xr * a.potNy;
#pragma unroll
      for (int m = 0; m < E; m++) dt[m] = tr[t + T * m];
    }
    if constexpr (MODE == 0 || MODE == 3) {
      const float2* src = a.probe + ((MODE == 3 ? (size_t)pg * a.nx : (size_t)0) + x0 + xr) * NY;
#pragma unroll
      for (int m = 0; m < E; m++) dv[m] = src[t + T * m];
    } else if constexpr (DMA) {
#pragma unroll
      for (int m = 0; m < E; m++) {
        if ((GAP >> m) & 1) {
          dv[m] = make_float2(0.f, 0.f);
          continue;
        }
        const unsigned o = BOFF ? boff[m] : bo(m);
        const float2 w = wstg[min(o >> 3, (unsigned)kStg - 1u)];  // index kept inside the area
        dv[m] = o == kOut ? make_float2(0.f, 0.f) : w;
      }
    } else {
      const unsigned so = (unsigned)(i * TPW) * rowb;
#pragma unroll
      for (int m = 0; m < E; m++) {
        if ((GAP >> m) & 1) dv[m] = make_float2(0.f, 0.f);  // inside the band gap for every thread
        else dv[m] = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rsi, BOFF ? boff[m] : bo(m), so, 0));
      }
    }
  };
  Twiddles<NY> tw;
  tw.init(a.twy, t);
  // one step: finish row i (v: state or probe row, tv: its transmission row)
  auto step = [&](int i, float2* v, const float2* tv) {
    if constexpr (MODE == 1 || MODE == 2 || MODE == 4) fft_wave<NY, +1, kSwz<NY>, GAP>(v, slot, tw, t);
    const int xr = i * TPW + sub;
    if constexpr (MODE == 2) {
      float2* dst = a.out_wave + ((size_t)pg * a.nx + x0 + xr) * NY;
#pragma unroll
      for (int m = 0; m < E; m++) dst[t + T * m] = v[m];
    } else if constexpr (MODE == 4) {
      const size_t r = ((size_t)pg * a.nx + x0 + xr) * NY;
      const float* vp = a.vpot + r;
      float* gp = a.gv + r;
#pragma unroll
      for (int m = 0; m < E; m++) {
        float sn, cs;
        sincos_phase(vp[t + T * m] * a.gv_sigma, sn, cs);
        gp[t + T * m] = a.gv_sigma * (v[m].y * cs - v[m].x * sn);
      }
    } else {
      if constexpr (MODE <= 1) {
#pragma unroll
        for (int m = 0; m < E; m++) v[m] = cmul(v[m], tv[m]);
      }
      fft_wave<NY, -1, kSwz<NY>, 0, ~GAP>(v, slot, tw, t);
      const unsigned so = (unsigned)(i * TPW) * rowb;
      using U2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(rs, 0u, 0, 0));
#pragma unroll
      for (int m = 0; m < E; m++)
        if (!((GAP >> m) & 1))
          __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v[m]), rs, BOFF ? boff[m] : bo(m), so, 0);
    }
  };
  // one register set: at 4 wavefronts per SIMD the other wavefronts cover the
  // load latency (measured equal to a two-set software pipeline at 3 per SIMD); at
  // 512 the state rows come through the LDS copies (the transmission rows of the next
  // step in registers as well measured slower: 121 VGPRs)
  float2 va[E], ta[E];
  prefetch(0);
  for (int i = 0; i < a.rw; i++) {
    load(i, va, ta);
    if constexpr (DMA) {
      __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): this wavefront's staging reads are done
      prefetch(i + 1);  // unconditional: one path into the step, so its waits count these copies
    }
    step(i, va, ta);
  }
  // vmcnt(0): the last (empty) copy has landed before the workgroup's LDS can be released
  if constexpr (DMA) __builtin_amdgcn_s_waitcnt(0x0F70);
}

// ============================================================================
// Sweep B: ky columns of the band, TC columns per block staged through LDS.
template <int NX> struct ColCfg {
  // TC columns per tile; BLK threads per block.  At 512 a tile is 16 columns =
  // one whole 128-byte line of each state row, staged by a 512-thread block, so
  // no line is fetched by two blocks.
  static constexpr int TC = col_tc<NX>();
  static constexpr int BLK = NX == 512 ? 512 : row_block<NX>();
  // tiles grouped on one XCD: those sharing a 128-byte line (16 columns); pairs
  // of whole-line tiles at 512 (measured faster than no grouping)
  static constexpr int GS = TC >= 16 ? 2 : 16 / TC;
  // large-size twiddles computed in the pass, or loaded from the table at 1024
  // (measured: columns 147 vs 128 us per 32-probe launch)
  static constexpr bool TWC = NX != 1024;
  // column slots use the PAD layout (registers are the scarcer resource here);
  // float2 per LDS column, rounded so the tile staging is at most 2-way
  static constexpr int STRIDE = ((lds_slot<NX, false>() + 31) / 32) * 32 + 2;
};

// Tile staging between the row-major state (TC contiguous complex per row) and
// the column-major LDS tile (PAD layout); 16-byte accesses (two columns per
// thread).  Thread (c, x0) moves columns c, c+1 of rows x0 + XS*i: one buffer
// offset and one LDS base per thread, the row step as the scalar offset and
// as immediate LDS offsets (XS is a multiple of 16 and x0 < XS, so
// lidx(x0 + XS*i) = lidx(x0) + XS*i + XS*i/16).  Global loads are issued in
// batches of up to 8 per thread before their LDS writes.
template <int NX>
struct TileStage {
  static constexpr int TC = ColCfg<NX>::TC, ST = ColCfg<NX>::STRIDE, H = TC / 2;
  static constexpr int BLK = ColCfg<NX>::BLK;
  static constexpr int IT = H * NX / BLK, IB = IT < 8 ? IT : 8, XS = BLK / H;
  static_assert(H * NX % BLK == 0 && BLK % H == 0 && IT % IB == 0 && XS % 16 == 0,
                "tile staging assumes whole rounds");
  static constexpr int lstep(int i) { return XS * i + XS * i / 16; }
  rsrc_t rs;          // the probe's band-compact state
  unsigned voff;      // byte offset of (row x0, column c) relative to the tile's first column
  int lbase;          // LDS index of (column c, row x0)
  unsigned rstep;     // bytes per XS rows
  __device__ __forceinline__ TileStage(const float2* pstate, int nx, int kcp) {
    const int c = 2 * (threadIdx.x % H), x0 = threadIdx.x / H;
    rs = make_rsrc(pstate, (long)nx * kcp * 8);
    voff = (unsigned)(x0 * kcp + c) * 8u;
    lbase = c * ST + lidx<NX, false>(x0);
    rstep = (unsigned)(XS * kcp) * 8u;
  }
  __device__ __forceinline__ void in(float2* tile, int c0) const {
    const unsigned co = (unsigned)c0 * 8u;
#pragma unroll
    for (int i0 = 0; i0 < IT; i0 += IB) {
      if (i0 > 0) __builtin_amdgcn_sched_barrier(0);  // keep the batches apart (registers)
      float4 q[IB];
#pragma unroll
      for (int i = 0; i < IB; i++)
        q[i] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rs, voff + co, (i0 + i) * rstep, 0));
#pragma unroll
      for (int i = 0; i < IB; i++) {
        tile[lbase + lstep(i0 + i)] = make_float2(q[i].x, q[i].y);
        tile[lbase + ST + lstep(i0 + i)] = make_float2(q[i].z, q[i].w);
      }
    }
  }
  // the same rows as in(), into registers: q[i] = columns c, c+1 of row x0 + XS*i
  __device__ __forceinline__ void in_regs(float4* q, int c0) const {
    const unsigned co = (unsigned)c0 * 8u;
#pragma unroll
    for (int i = 0; i < IT; i++)
      q[i] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rs, voff + co, i * rstep, 0));
  }
  // columns c, c+1 of row x0 + XS*i from u[i], w[i]
  __device__ __forceinline__ void out_regs(const float2* u, const float2* w, int c0) const {
    const unsigned co = (unsigned)c0 * 8u;
#pragma unroll
    for (int i = 0; i < IT; i++) buf_store_f4(make_float4(u[i].x, u[i].y, w[i].x, w[i].y), rs, voff + co + i * rstep);
  }
  __device__ __forceinline__ void out(const float2* tile, int c0) const {
    const unsigned co = (unsigned)c0 * 8u;
#pragma unroll
    for (int i0 = 0; i0 < IT; i0 += IB) {
      if (i0 > 0) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int i = i0; i < i0 + IB; i++) {
        const float2 u = tile[lbase + lstep(i)], w = tile[lbase + ST + lstep(i)];
        buf_store_f4(make_float4(u.x, u.y, w.x, w.y), rs, voff + co + i * rstep);
      }
    }
  }
};

// Column sweeps whose tile staging hands each thread exactly the elements t + T*m
// of both tile columns (two columns per tile, one transform per block: 4096): the
// loads go straight into the transform's registers and the results straight back
// to memory, and the block needs a single LDS column slot for the exchanges (used
// for the launches without detector sums or diffraction patterns).
template <int NX>
constexpr bool col_direct() {
  using S = TileStage<NX>;
  constexpr int T = NX / Plan<NX>::E;
  return S::H == 1 && ColCfg<NX>::TC == 2 && ColCfg<NX>::BLK == T && S::XS == T && S::IT == Plan<NX>::E;
}

// Work item of block blockIdx.x such that work items GS*j ... GS*j+GS-1 (the
// tiles that share the state's 128-byte lines when a block stages one tile)
// run back to back on the same XCD: blocks are dispatched round-robin over the
// 8 XCDs, each with its own L2, so blocks b, b+8, ..., b+8*(GS-1) share an L2.
// The last partial group of 8*GS blocks keeps the identity mapping.
template <int GS>
__device__ __forceinline__ int xcd_group_block(int blk, int nblk) {
  if (blk >= (nblk & ~(8 * GS - 1))) return blk;
  const int xcd = blk & 7, slot = blk >> 3;
  return (slot / GS) * (8 * GS) + xcd * GS + (slot % GS);
}

// Sum over the 64 lanes of a wavefront (result valid in every lane).
__device__ __forceinline__ float wave_sum(float s) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  return s;
}

// A block advances a.tpb consecutive tiles of one probe; tiles that share
// 128-byte state lines run on one XCD at about the same time, so the line is
// fetched from HBM once and the others hit L2.
template <int NX, bool REC, bool DIFF>
__global__ __launch_bounds__(ColCfg<NX>::BLK) void k_sweep_cols(SweepArgs a) {
  constexpr int E = Plan<NX>::E;
  constexpr int T = NX / E;
  constexpr int TC = ColCfg<NX>::TC;
  constexpr int ST = ColCfg<NX>::STRIDE;
  constexpr int G = ColCfg<NX>::BLK / T;  // transforms in flight per block
  constexpr int ROUNDS = (TC + G - 1) / G;
  constexpr bool PXR = E <= 8;   // per-thread propagator factors and k^2 kept in registers
  constexpr int GAP = gap_mask<NX>();  // kx slots in the band gap: zero after the propagator
  extern __shared__ float2 tile[];  // TC * ST
  __shared__ float red[kMaxDet + 1];

  const int bpp = (a.ntiles + a.tpb - 1) / a.tpb;  // blocks per probe
  const int b = xcd_group_block<ColCfg<NX>::GS>(blockIdx.x, gridDim.x);
  const int p = b / bpp;
  const int tile_first = (b % bpp) * a.tpb;
  const int tile_end = min(tile_first + a.tpb, a.ntiles);
  const int pg = a.probe0 + p;
  const TileStage<NX> stage(a.state + (size_t)pg * a.nx * a.kcp, a.nx, a.kcp);

  // one transform per wavefront when T == 64: the column is wave-uniform
  const int slot_id = T == 64 ? wave_uniform(threadIdx.x >> 6) : threadIdx.x / T;
  const int t = threadIdx.x % T;
  Twiddles<NX> tw;
  tw.init(a.twx, t);
  // per-thread propagator factors and k^2 (independent of the column)
  float2 pxr[PXR ? E : 1];
  float kx2r[PXR ? E : 1];
  if constexpr (PXR) {
#pragma unroll
    for (int m = 0; m < E; m++) {
      pxr[m] = a.px[t + T * m];
      kx2r[m] = a.kx2[t + T * m];
    }
  }
  auto kx2_of = [&](int m) -> float { return PXR ? kx2r[m] : a.kx2[t + T * m]; };
  auto px_of = [&](int m) -> float2 { return PXR ? pxr[m] : a.px[t + T * m]; };
  if constexpr (col_direct<NX>() && !REC && !DIFF) {
    // column c of the tile in v (element t + T*m in v[m]), LDS slot `tile` for the
    // exchanges: column FFT, propagator and band mask, column IFFT
    auto column = [&](float2 (&v)[E], int c) __attribute__((always_inline)) {
      const int iy = c <= a.band_a ? c : (c < a.kc ? c - (a.band_a + 1) + (a.ny - a.band_b) : -1);
      fft_wave<NX, -1, false, 0, ~GAP, ColCfg<NX>::TWC>(v, tile, tw, t);
      if (iy >= 0) {
        const float ky2 = a.ky2[iy];
        const float2 py = a.py[iy];
#pragma unroll
        for (int m = 0; m < E; m++) {
          const float kx2 = kx2_of(m);
          if (!((GAP >> m) & 1) && kx2 < a.k2max && (kx2 + ky2) < a.k2max) v[m] = cmul(cmul(v[m], py), px_of(m));
          else v[m] = make_float2(0.f, 0.f);
        }
      } else {
#pragma unroll
        for (int m = 0; m < E; m++) v[m] = make_float2(0.f, 0.f);
      }
      fft_wave<NX, +1, false, GAP, -1, ColCfg<NX>::TWC>(v, tile, tw, t);
    };
    for (int tl = tile_first; tl < tile_end; tl++) {
      if (tl > tile_first) __syncthreads();  // the previous tile's last slot reads are done
      const int c0 = tl * TC;
      float4 q[E];
      stage.in_regs(q, c0);
      float2 u[E], w[E];
#pragma unroll
      for (int m = 0; m < E; m++) {
        u[m] = make_float2(q[m].x, q[m].y);
        w[m] = make_float2(q[m].z, q[m].w);
      }
      column(u, c0);
      xform_sync<NX>();  // the second column reuses the slot
      column(w, c0 + 1);
      stage.out_regs(u, w, c0);
    }
    return;
  }
  for (int tl = tile_first; tl < tile_end; tl++) {
    if (tl > tile_first) __syncthreads();  // previous tile's LDS reads (stage-out, red) are done
    stage.in(tile, tl * TC);
    if constexpr (REC) {
      if (threadIdx.x <= kMaxDet) red[threadIdx.x] = 0.f;
    }
    __syncthreads();
    const int c0 = tl * TC;
    float dsum[REC ? kMaxDet + 1 : 1];
    if constexpr (REC) {
#pragma unroll
      for (int d = 0; d <= kMaxDet; d++) dsum[d] = 0.f;
    }
#pragma unroll 1
    for (int r = 0; r < ROUNDS; r++) {
      const int cl = r * G + slot_id;
      if (cl >= TC) break;  // only when G does not divide TC (uniform per transform)
      float2* col = tile + cl * ST;
      const int c = c0 + cl;
      const int iy = c <= a.band_a ? c : (c < a.kc ? c - (a.band_a + 1) + (a.ny - a.band_b) : -1);
      float2 v[E];
#pragma unroll
      for (int m = 0; m < E; m++) v[m] = col[lidx<NX, false>(t + T * m)];
      fft_wave<NX, -1, false, 0, ~GAP, ColCfg<NX>::TWC>(v, col, tw, t);
      // propagate: P = px[ix]*py[iy] (1/(nx*ny) folded into px), zero outside k2max
      if (iy >= 0) {
        const float ky2 = a.ky2[iy];
        const float2 py = a.py[iy];
#pragma unroll
        for (int m = 0; m < E; m++) {
          if ((GAP >> m) & 1) {  // kx in the band gap for every thread
            v[m] = make_float2(0.f, 0.f);
            continue;
          }
          const float kx2 = kx2_of(m);
          if (kx2 < a.k2max && (kx2 + ky2) < a.k2max) v[m] = cmul(cmul(v[m], py), px_of(m));
          else v[m] = make_float2(0.f, 0.f);
        }
        if constexpr (REC) {
          // collectIntensity: detector d sums |Psi|^2 over pixels whose
          // (shift-corrected) k^2 lies in [k2in, k2out]; slot kMaxDet is the total.
#pragma unroll
          for (int m = 0; m < E; m++) dsum[kMaxDet] += cnorm(v[m]);
#pragma unroll
          for (int d = 0; d < kMaxDet; d++) {  // unrolled: dsum stays in registers
            if (d >= a.ndet) break;
            const int sx = a.det_sx[d];
            const float k2y = a.ky2[(iy - a.det_sy[d]) & (a.ny - 1)];
            const float lo = a.det_k2in[d], hi = a.det_k2out[d];
            float s = 0.f;
            if (sx == 0) {
#pragma unroll
              for (int m = 0; m < E; m++) {
                const float k2 = kx2_of(m) + k2y;
                s += (k2 >= lo && k2 <= hi) ? cnorm(v[m]) : 0.f;
              }
            } else {
#pragma unroll
              for (int m = 0; m < E; m++) {
                const float k2 = a.kx2[(t + T * m - sx) & (a.nx - 1)] + k2y;
                s += (k2 >= lo && k2 <= hi) ? cnorm(v[m]) : 0.f;
              }
            }
            dsum[d] += s;
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
      fft_wave<NX, +1, false, GAP, -1, ColCfg<NX>::TWC>(v, col, tw, t);
      xform_sync<NX>();
#pragma unroll
      for (int m = 0; m < E; m++) col[lidx<NX, false>(t + T * m)] = v[m];
    }
    if constexpr (REC) {  // one LDS atomic per wavefront and sum
#pragma unroll
      for (int d = 0; d < kMaxDet; d++) {
        if (d >= a.ndet) break;
        const float s = wave_sum(dsum[d]);
        if ((threadIdx.x & 63) == 0) atomicAdd(&red[d], s);
      }
      const float s = wave_sum(dsum[kMaxDet]);
      if ((threadIdx.x & 63) == 0) atomicAdd(&red[kMaxDet], s);
    }
    __syncthreads();
    stage.out(tile, c0);
    if constexpr (REC) {
      if (threadIdx.x <= kMaxDet) a.partial[((size_t)p * a.ntiles + tl) * (kMaxDet + 1) + threadIdx.x] = red[threadIdx.x];
    }
  }
}

// ============================================================================
// Generic batched row transforms of length L (power of two), full rows.
template <int L, int DIR>
__global__ __launch_bounds__(row_block<L>()) void k_rows_fft(float2* __restrict__ data, long nrows,
                                                             const float2* __restrict__ tw, float scale) {
  constexpr int E = Plan<L>::E;
  constexpr int T = L / E;
  constexpr int RPB = row_block<L>() / T;
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
__global__ __launch_bounds__(row_block<M>()) void k_rows_bluestein(float2* __restrict__ data, long nrows, int n,
                                                           const float2* __restrict__ chirp,
                                                           const float2* __restrict__ vhat,
                                                           const float2* __restrict__ tw, float scale) {
  constexpr int E = Plan<M>::E;
  constexpr int T = M / E;
  constexpr int RPB = row_block<M>() / T;
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
// Steps per unit (a wavefront, or the block for block-wide transforms): enough
// units to fill every CU several times over, each looping over up to a.rw
// steps (a power of two dividing nx/TPW).
template <int N>
void launch_rows_mode(const SweepArgs& a0, int mode, hipStream_t st) {
  constexpr int T = N / Plan<N>::E;
  constexpr int TPW = T <= 64 ? 64 / T : 1;
  constexpr int UPB = T <= 64 ? row_block<N>() / 64 : 1;
  SweepArgs a = a0;
  const long rows = (long)a.nprobe * a.nx;
  const int max_rw = a.rw > 0 ? a.rw : 4;  // caller's cap (0: default, measured best at 512^2)
  const long target_units = 256L * 8 * UPB;
  int rw = 1;
  while (rw < max_rw && rows / ((long)TPW * rw * 2) >= target_units && (a.nx / TPW) % (rw * 2) == 0) rw *= 2;
  a.rw = rw;
  const long units = rows / ((long)TPW * rw);
  dim3 grid((unsigned)((units + UPB - 1) / UPB));
  if (mode == 0) hipLaunchKernelGGL((k_sweep_rows<N, 0>), grid, dim3(row_block<N>()), 0, st, a);
  else if (mode == 1) hipLaunchKernelGGL((k_sweep_rows<N, 1>), grid, dim3(row_block<N>()), 0, st, a);
  else if (mode == 2) hipLaunchKernelGGL((k_sweep_rows<N, 2>), grid, dim3(row_block<N>()), 0, st, a);
  else if (mode == 3) hipLaunchKernelGGL((k_sweep_rows<N, 3>), grid, dim3(row_block<N>()), 0, st, a);
  else hipLaunchKernelGGL((k_sweep_rows<N, 4>), grid, dim3(row_block<N>()), 0, st, a);
}

template <int N>
constexpr int col_lds_bytes_t() { return (int)sizeof(float2) * ColCfg<N>::TC * ColCfg<N>::STRIDE; }

// Tiles per block: a.tpb from the caller when > 0, else 1 (measured best at
// 512^2: more tiles per block only lengthen each block's serial phases).
template <int N>
void launch_cols_t(const SweepArgs& a0, bool rec, bool diff, hipStream_t st) {
  const size_t lds = col_direct<N>() && !rec && !diff ? sizeof(float2) * ColCfg<N>::STRIDE : col_lds_bytes_t<N>();
  SweepArgs a = a0;
  if (a.tpb <= 0) a.tpb = 1;
  dim3 grid((unsigned)(a.nprobe * ((a.ntiles + a.tpb - 1) / a.tpb)));
  if (rec && diff) hipLaunchKernelGGL((k_sweep_cols<N, true, true>), grid, dim3(ColCfg<N>::BLK), lds, st, a);
  else if (rec) hipLaunchKernelGGL((k_sweep_cols<N, true, false>), grid, dim3(ColCfg<N>::BLK), lds, st, a);
  else if (diff) hipLaunchKernelGGL((k_sweep_cols<N, false, true>), grid, dim3(ColCfg<N>::BLK), lds, st, a);
  else hipLaunchKernelGGL((k_sweep_cols<N, false, false>), grid, dim3(ColCfg<N>::BLK), lds, st, a);
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
  constexpr int RPB = row_block<L>() / (L / Plan<L>::E);
  hipLaunchKernelGGL((k_rows_fft<L, DIR>), dim3((unsigned)((nrows + RPB - 1) / RPB)), dim3(row_block<L>()), 0, st, d, nrows, tw,
                     scale);
}

template <int M>
void blue_t(float2* d, long nrows, int n, const float2* chirp, const float2* vhat, const float2* tw, float scale,
            hipStream_t st) {
  constexpr int RPB = row_block<M>() / (M / Plan<M>::E);
  hipLaunchKernelGGL((k_rows_bluestein<M>), dim3((unsigned)((nrows + RPB - 1) / RPB)), dim3(row_block<M>()), 0, st, d, nrows, n,
                     chirp, vhat, tw, scale);
}


// ============================================================================
// Adjoint sweeps (reverse mode of the two forward sweeps).
//
// Column sweep (slice s, reverse):  g_Phi = colFFT(g_X)  [last slice: 4 (I - T) Phi,
// Phi = colFFT(X_{S-1})/nx, I = |Phi|^2]  ->  g_Psi = conj(P) g_Phi  ->  g_R = colIFFT(g_Psi).
template <int NX, bool LAST>
__global__ __launch_bounds__(ColCfg<NX>::BLK) void k_adj_cols(AdjArgs A) {
  const SweepArgs& a = A.f;
  constexpr int E = Plan<NX>::E;
  constexpr int T = NX / E;
  constexpr int TC = ColCfg<NX>::TC;
  constexpr int ST = ColCfg<NX>::STRIDE;
  constexpr int G = ColCfg<NX>::BLK / T;
  constexpr int ROUNDS = (TC + G - 1) / G;
  extern __shared__ float2 tile[];
  __shared__ float red;
  const int b = xcd_group_block<ColCfg<NX>::GS>(blockIdx.x, gridDim.x);
  const int p = b / a.ntiles;
  const int tile_id = b % a.ntiles;
  const int pg = a.probe0 + p;
  const int c0 = tile_id * TC;
  const size_t off = (size_t)pg * a.nx * a.kcp;
  if (threadIdx.x == 0) red = 0.f;
  TileStage<NX>((LAST ? A.xprev : A.gx) + off, a.nx, a.kcp).in(tile, c0);
  __syncthreads();
  const int slot_id = threadIdx.x / T;
  const int t = threadIdx.x % T;
  Twiddles<NX> tw;
  tw.init(a.twx, t);
  float lsum = 0.f;
#pragma unroll 1
  for (int r = 0; r < ROUNDS; r++) {
    const int cl = r * G + slot_id;
    if (cl >= TC) break;
    float2* col = tile + cl * ST;
    const int c = c0 + cl;
    const int iy = c <= a.band_a ? c : (c < a.kc ? c - (a.band_a + 1) + (a.ny - a.band_b) : -1);
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; m++) v[m] = col[lidx<NX, false>(t + T * m)];
    fft_wave<NX, -1, false, 0, -1, ColCfg<NX>::TWC>(v, col, tw, t);
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
    fft_wave<NX, +1, false, 0, -1, ColCfg<NX>::TWC>(v, col, tw, t);
    xform_sync<NX>();
#pragma unroll
    for (int m = 0; m < E; m++) col[lidx<NX, false>(t + T * m)] = v[m];
  }
  if constexpr (LAST) atomicAdd(&red, lsum);
  __syncthreads();
  TileStage<NX>(A.gx + off, a.nx, a.kcp).out(tile, c0);
  if constexpr (LAST) {
    if (threadIdx.x == 0) A.loss_partial[(size_t)p * a.ntiles + tile_id] = red;
  }
}

// Row sweep (slice s, reverse):  g_a = rowIFFT(g_R);  psi_s = rowIFFT(X_{s-1}) (or the
// probe);  g_t += conj(psi_s) g_a;  g_psi = conj(t_s) g_a;  g_X_{s-1} = rowFFT(g_psi)
// (s > 0) or sum_p g_psi -> gprobe (s = 0).  BLF (the window is the whole slice): the
// row of g_t is complete here, so its row FFT, band-compact in the potential grid's
// band, goes to A.blstate (the band limit's first pass) instead of the row to A.gT.
template <int NY, bool FIRST, bool BLF>
__global__ __launch_bounds__(row_block<NY>()) void k_adj_rows(AdjArgs A) {
  const SweepArgs& a = A.f;
  constexpr int E = Plan<NY>::E;
  constexpr int T = NY / E;
  constexpr int RPB = row_block<NY>() / T;
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
  // band columns through buffer resources: offsets in the band gap are kOut, so
  // their loads return 0 and their stores are dropped (no per-element branches)
  const int gap = NY - a.band_b - a.band_a - 1;
  const size_t roff = ((size_t)pg * a.nx + x) * a.kcp;
  const rsrc_t rg = make_rsrc(A.gx + roff, (long)a.kcp * 8);
  unsigned bo[E];
#pragma unroll
  for (int m = 0; m < E; m++) bo[m] = band_off(t + T * m, a.band_a, gap);
  float2 g[E], w[E];
#pragma unroll
  for (int m = 0; m < E; m++) g[m] = buf_load_c64(rg, bo[m]);
  fft_wave<NY, +1>(g, slot, tw, t);
  if constexpr (FIRST) {
    const float2* pr = a.probe + (size_t)x * NY;
#pragma unroll
    for (int m = 0; m < E; m++) w[m] = pr[t + T * m];
  } else {
    const rsrc_t rx = make_rsrc(A.xprev + roff, (long)a.kcp * 8);
#pragma unroll
    for (int m = 0; m < E; m++) w[m] = buf_load_c64(rx, bo[m]);
    fft_wave<NY, +1>(w, slot, tw, t);
  }
  const int2 ip = a.ipos[pg];
  const float2* tr = a.trans + (size_t)(x + ip.x) * a.potNy + ip.y;
  float* gt = BLF ? nullptr : A.gT + 2 * ((size_t)(x + ip.x) * a.potNy + ip.y);  // BLF: no dL/dt buffer
#pragma unroll
  for (int m = 0; m < E; m++) {
    const int iy = t + T * m;
    const float2 gtv = cmulc(g[m], w[m]);  // conj(psi) * g_a
    if constexpr (BLF) {
      w[m] = gtv;  // kept for the row FFT below
    } else if (A.single_writer == 2) {  // the window is the whole slice: this is the pixel's only term
      *reinterpret_cast<float2*>(gt + 2 * iy) = gtv;
    } else if (A.single_writer) {  // one probe per launch: each transmission pixel has one writer
      float2* p2 = reinterpret_cast<float2*>(gt + 2 * iy);
      const float2 o = *p2;
      *p2 = make_float2(o.x + gtv.x, o.y + gtv.y);
    } else {
      atomicAdd(gt + 2 * iy, gtv.x);
      atomicAdd(gt + 2 * iy + 1, gtv.y);
    }
    g[m] = cmulc(g[m], tr[iy]);  // conj(t) * g_a
  }
  if constexpr (BLF) {
    // the band limit's first pass (k_sweep_rows mode 3) on the row x of g_t: row FFT,
    // band columns of the potential grid stored (offsets in its band gap are kOut)
    fft_wave<NY, -1>(w, slot, tw, t);
    const rsrc_t rb = make_rsrc(A.blstate + (size_t)(x + ip.x) * A.bl_kcp, (long)A.bl_kcp * 8);
#pragma unroll
    for (int m = 0; m < E; m++) buf_store_c64(w[m], rb, band_off(t + T * m, A.bl_band_a, A.bl_gap));
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
    for (int m = 0; m < E; m++) buf_store_c64(g[m], rg, bo[m]);
  }
}

template <int N>
void launch_adj_rows_t(const AdjArgs& a, bool first, hipStream_t st) {
  constexpr int T = N / Plan<N>::E;
  constexpr int RPB = row_block<N>() / T;
  const long rows = (long)a.f.nprobe * a.f.nx;
  dim3 grid((unsigned)((rows + RPB - 1) / RPB));
  if (a.blstate) {
    if (first) hipLaunchKernelGGL((k_adj_rows<N, true, true>), grid, dim3(row_block<N>()), 0, st, a);
    else hipLaunchKernelGGL((k_adj_rows<N, false, true>), grid, dim3(row_block<N>()), 0, st, a);
  } else {
    if (first) hipLaunchKernelGGL((k_adj_rows<N, true, false>), grid, dim3(row_block<N>()), 0, st, a);
    else hipLaunchKernelGGL((k_adj_rows<N, false, false>), grid, dim3(row_block<N>()), 0, st, a);
  }
}

template <int N>
void launch_adj_cols_t(const AdjArgs& a, bool last, hipStream_t st) {
  const size_t lds = col_lds_bytes_t<N>();
  dim3 grid((unsigned)(a.f.nprobe * a.f.ntiles));
  if (last) hipLaunchKernelGGL((k_adj_cols<N, true>), grid, dim3(ColCfg<N>::BLK), lds, st, a);
  else hipLaunchKernelGGL((k_adj_cols<N, false>), grid, dim3(ColCfg<N>::BLK), lds, st, a);
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
