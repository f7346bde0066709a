// qs_capi.cpp -- C ABI over the gfx950 multislice kernels.
//
// Host responsibilities: validate parameters, derive the small
// single-precision lookup tables exactly as the reference derives them
// (k^2 tables and band limit of propagate_slow, detector k^2 limits of
// readFile), allocate device buffers, and sequence the sweeps.  All
// per-pixel work runs in the kernels: the slice sweeps and their reverse
// (qs_sweep.hpp, instantiated per size by qs_sizes.hip), set-up and probe
// (qs_setup.hip), potential builder (qs_potential.hip) and adjoint helpers
// (qs_adjoint.hip).
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <algorithm>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <unistd.h>
#include <vector>

#include "../../include/qstem.h"
#include "qs_internal.hpp"

using qs::ProbeArgs;
using qs::SweepArgs;

namespace {

thread_local std::string g_create_error;

// Device memory held by the library (all contexts): current and peak bytes.
static std::atomic<long long> g_dev_bytes{0}, g_dev_peak{0};
static void dev_account(long long delta) {
  const long long now = g_dev_bytes.fetch_add(delta) + delta;
  long long peak = g_dev_peak.load();
  while (now > peak && !g_dev_peak.compare_exchange_weak(peak, now)) {
  }
}

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t n) {
    if (n <= bytes && p) return hipSuccess;
    release();
    hipError_t e = hipMalloc(&p, n > 0 ? n : 16);
    if (e == hipSuccess) {
      bytes = n;
      dev_account((long long)n);
    } else {
      p = nullptr;
    }
    return e;
  }
  void release() {
    if (p) {
      (void)hipFree(p);
      dev_account(-(long long)bytes);
    }
    p = nullptr;
    bytes = 0;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

// Small double-precision FFT for table set-up (Bluestein kernels' vhat).
void host_fft_pow2(std::vector<std::complex<double>>& a, int sign) {
  const int n = (int)a.size();
  for (int i = 1, j = 0; i < n; i++) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (int len = 2; len <= n; len <<= 1) {
    for (int k = 0; k < len / 2; k++) {
      const double ang = sign * 2.0 * M_PI * k / len;
      const std::complex<double> w(cos(ang), sin(ang));
      for (int i = 0; i < n; i += len) {
        auto u = a[i + k], v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
      }
    }
  }
}

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

// Unnormalised DFT of any length whose prime factors are small (recursive
// decimation in time by the smallest prime factor), for host set-up tables.
void host_fft_any(std::vector<std::complex<double>>& a, int sign) {
  const int n = (int)a.size();
  if (n <= 1) return;
  int p = 2;
  while (n % p) p++;
  const int m = n / p;
  std::vector<std::vector<std::complex<double>>> sub(p, std::vector<std::complex<double>>(m));
  for (int r = 0; r < p; r++)
    for (int k = 0; k < m; k++) sub[r][k] = a[r + (size_t)p * k];
  for (auto& v : sub) host_fft_any(v, sign);
  for (int k = 0; k < n; k++) {
    std::complex<double> acc = 0.0;
    for (int r = 0; r < p; r++) {
      const double ang = sign * 2.0 * M_PI * (double)(((long long)r * k) % n) / n;
      acc += sub[r][k % m] * std::complex<double>(cos(ang), sin(ang));
    }
    a[k] = acc;
  }
}

double wavelength(double kev) {  // stem3/stemutil.cpp:714-725
  const double emass = 510.99906, hc = 12.3984244;
  return hc / sqrt(kev * (2.0 * emass + kev));
}

std::vector<float2> twiddles(int n) {
  std::vector<float2> t(n);
  for (int m = 0; m < n; m++) {
    const double a = -2.0 * M_PI * m / n;
    t[m] = make_float2((float)cos(a), (float)sin(a));
  }
  return t;
}

// Row transform of length n for the transmission band limit: either a direct
// power-of-two Stockham transform or Bluestein through M = 2^k >= 2n-1.
struct RowXform {
  int n = 0, M = 0;
  bool gen = false;   // through the mixed-radix LDS passes (direct when M == 0, else Bluestein)
  qs::GenPlan plan{};
  int rpb = 0;
  DevBuf tw, chirp_f, vhat_f, chirp_b, vhat_b;
};

struct LutArgsHost {
  int nx = 0, nz = 0, nzp = 0;
  double dkx = 0, dkz = 0;
};

}  // namespace

struct qs_ctx {
  qs_params p{};
  int device = 0;  // HIP device current at qs_create
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // second stream of qs_run_stem / qs_run_groups: consecutive probe groups (or the
  // contexts of qs_run_groups) alternate between the two streams so that one group's
  // launches fill the other's ramp-up and drain; each call joins it back into stream
  int nstreams = 2;
  hipStream_t stream2 = nullptr;
  hipEvent_t evfork = nullptr, evjoin = nullptr;
  float last_ms = 0.f;
  std::string err;
  int batch = 0;
  int rows_per_wave = 0;  // cap on row-sweep steps per wavefront (0: kernel default)
  int tiles_per_block = 0;  // column-sweep tiles per block (0: launcher default)
  // frozen-phonon configurations: nconfig transmission functions of trans_slices
  // slices each, back to back in trans; the potential setters fill cur_config
  int nconfig = 1;
  int cur_config = 0;
  qs_ctx* trans_src = nullptr;  // non-owning: the context whose transmission functions this one uses
  DevBuf cfgidx;                // per-probe configuration of the current run
  // HIP graph of the qs_run_stem slice loop, reused while the call shape matches
  bool use_graph = false;
  hipGraphExec_t gexec = nullptr, gexec2 = nullptr;  // one graph per stream
  std::vector<long long> gkey;
  // derived tables
  int band_a = 0, band_b = 0, kc = 0, kcp = 0, ntiles = 0, tc = 0;
  float k2max = 0.f;
  std::vector<float> kx2, ky2;
  float det_k2in[QS_MAX_DET]{}, det_k2out[QS_MAX_DET]{};
  // device
  DevBuf twx, twy, px, py, dkx2, dky2;
  DevBuf probe, trans, scratch, V;
  DevBuf state, ipos, partial, partial2, det, diff, exitw, total;
  DevBuf pkx2, pky2, probe_partial;
  RowXform rx, ry;  // potNx and potNy row transforms
  // wave grids that are not powers of two (or outside 64..4096): the unfused path of
  // qs_generic.hip with row transforms of lengths nx and ny
  bool generic = false;
  RowXform wx, wy;
  DevBuf gscratch;
  // generic sizes whose prime factors are 2, 3, 5 and 7: the fused mixed-radix path
  // (k_gen_rows / k_gen_cols): radix plans, column tiles holding in-band ky columns,
  // rows per block of the row kernel
  bool mixed = false;
  qs::GenPlan gpx{}, gpy{};
  int gvx = 0, gvy = 0;  // radix plan variants of gpx, gpy
  DevBuf gtiles;
  int gntiles = 0, gtc = 0, grpb = 0;
  int gkc = 0;  // entries of gtiles (tiles, or in-band columns for the compiled column kernel)
  // band limit of the transmission functions as band-compact sweeps (potNx and
  // potNy powers of two): row FFT into the ky band, column FFT / disk mask with
  // 1/(potNx potNy) / column IFFT, row IFFT; else row transforms and transposes
  struct BandSweeps {
    bool on = false;
    int band_a = 0, band_b = 0, kc = 0, kcp = 0, ntiles = 0;
  } bl;
  DevBuf blpx, blpy, blstate;
  // potential builder
  std::vector<double> scat;       // 93 x 123 Rez table
  std::vector<double> scat_offs;  // 93 x 123 ion-offset table (scatParOffs)
  DevBuf atoms, kinds, lutptrs, olutptrs, spl, luttmp, lutscr;
  std::vector<DevBuf> oluts;      // ion-offset LUT per kind (empty buffer: no charged atom of the kind)
  // what the last qs_build_potential built (qs_atom_gradients reuses its LUTs and kinds)
  struct BuiltPot {
    bool valid = false;
    std::vector<int> Z;
    std::vector<float> q;
    qs_pot_params pp{};
  } built;
  DevBuf reach, tcount, tstart, tlist;  // gather form: per-atom reach, atoms binned by pixel tile
  std::vector<DevBuf> luts;
  RowXform lx, lz;  // LUT transform lengths
  int trans_slices = 0;
  bool v_valid = false;     // c->V holds the potential of the current transmission slices
  double probe_norm = 1.0;  // c of psi0 = c * IFFT(Psi0)
  bool custom_wave = false;  // incident wave set by qs_set_incident_wave
  ProbeArgs probe_args{};
  bool have_probe = false;
  // adjoint
  std::vector<DevBuf> ckpt;
  DevBuf gx, target, lossp, gT, gprobe, gV, dsum, adj_ipos;
  // band limit of dL/dt on the second stream: a second band-compact buffer, events
  // "row sweep wrote buffer b" and "band limit done with buffer b"
  DevBuf blstate2;
  hipEvent_t evrow[2] = {nullptr, nullptr}, evbl[2] = {nullptr, nullptr};
  // forward pass kept by qs_forward_save for qs_adjoint_saved (checkpoints, the
  // adjoint seed in gx, the loss); cleared by every change of potential, probe or
  // optics (with the graph) and by the reverse pass itself
  bool adj_saved = false;
  int adj_nprobe = 0, adj_K = 0;
  float adj_loss = 0.f;
  int state_probes = 0;
  bool state_valid = false;
  // optional per-launch timing of the two sweeps (HIP events on the context stream)
  bool profiling = false;
  std::vector<hipEvent_t> evpool;
  double ms_rows = 0, ms_cols = 0;
  long n_rows = 0, n_cols = 0;
};

static int fail(qs_ctx* c, int code, const std::string& m) {
  if (c) c->err = m;
  return code;
}

#define HIPCHK(ctx, expr)                                                                   \
  do {                                                                                      \
    hipError_t e__ = (expr);                                                                \
    if (e__ != hipSuccess)                                                                  \
      return fail(ctx, QS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e__));     \
  } while (0)

template <class T>
static hipError_t upload(DevBuf& b, const std::vector<T>& v, hipStream_t st) {
  hipError_t e = b.ensure(sizeof(T) * v.size());
  if (e != hipSuccess) return e;
  return hipMemcpyAsync(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, st);
}

// Launch shape of the mixed-radix path: rows per block of k_gen_rows, columns per
// tile of k_gen_cols (powers of two; 0 = the default, the most that fit the block's
// LDS, at most 16) and the radix plan variants of x and y.  Variant
// kGenPlanCompiled selects the wavefront-resident kernels of a compiled plan for
// that axis (rows: rpb = rows per transform slot, a power of two up to 16, default
// 2; columns: tc unused).  Rebuilds the column work list: tiles of tc columns that
// hold an in-band ky column, or, for the compiled column kernel, the in-band ky
// columns themselves.
static int set_mixed_shape(qs_ctx* c, int rpb, int tc, int vx, int vy) {
  const int nx = c->p.nx, ny = c->p.ny;
  const bool cx = vx == qs::kGenPlanCompiled, cy = vy == qs::kGenPlanCompiled;
  if (rpb == 0) rpb = cy ? 2 : qs::gen_rows_per_block(ny);
  if (tc == 0) tc = qs::gen_cols_per_tile(nx);
  qs::GenPlan px{}, py{};
  if ((rpb & (rpb - 1)) || (tc & (tc - 1)) || rpb < 1 || tc < 1 || rpb > 16 || tc > 16 ||
      (!cy && rpb > qs::gen_max_vectors(ny)) || tc > qs::gen_max_vectors(nx) || !qs::gen_plan_variant(nx, vx, &px) ||
      !qs::gen_plan_variant(ny, vy, &py))
    return fail(c, QS_ERR_ARG, "mixed-radix launch shape: rows per block and columns per tile must be powers of "
                               "two up to 16 that fit the block's LDS, plan variants 0-2 (3: compiled plans only)");
  const int w = cx ? 1 : tc;
  std::vector<int> tiles;  // tiles of w columns with a ky column inside the band
  for (int y0 = 0; y0 < ny; y0 += w) {
    bool any = false;
    for (int y = y0; y < std::min(ny, y0 + w); y++) any = any || c->ky2[y] < c->k2max;
    if (any) tiles.push_back(y0);
  }
  if (tiles.empty()) return fail(c, QS_ERR_ARG, "no ky column inside the band");
  HIPCHK(c, upload(c->gtiles, tiles, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->gkc = (int)tiles.size();
  c->gntiles = cx ? (c->gkc + qs::mix_cols_per_block(nx) - 1) / qs::mix_cols_per_block(nx) : c->gkc;
  c->grpb = rpb;
  c->gtc = tc;
  c->gpx = px;
  c->gpy = py;
  c->gvx = vx;
  c->gvy = vy;
  return QS_OK;
}

static int setup_rowxform(qs_ctx* c, RowXform& r, int n) {
  r.n = n;
  r.gen = false;
  if (is_pow2(n) && n >= 64 && n <= 4096) {
    r.M = 0;
    HIPCHK(c, upload(r.tw, twiddles(n), c->stream));
    return QS_OK;
  }
  if (qs::gen_plan(n, &r.plan)) {  // prime factors 2, 3, 5, 7: direct mixed-radix passes
    r.gen = true;
    r.M = 0;
    r.rpb = qs::gen_rows_per_block(n);
    HIPCHK(c, upload(r.tw, twiddles(n), c->stream));
    return QS_OK;
  }
  // Bluestein through a power of two M >= 2n-1 up to 4096 (the optimised fixed-plan
  // transforms: measured faster than the LDS passes at the shorter length 2187 for
  // n = 1088), else through the smallest length M >= 2n-1 with prime factors 2, 3, 5, 7
  // that one block holds in LDS
  int M = 64;
  while (M < 2 * n - 1) M <<= 1;
  int Mg = 2 * n - 1;
  qs::GenPlan pg{};
  while (M > 4096 && Mg <= 8 * n && !qs::gen_plan(Mg, &pg)) Mg++;
  const bool use_gen = M > 4096 && Mg <= 8 * n && qs::gen_plan(Mg, &pg);
  if (use_gen) {
    r.gen = true;
    r.plan = pg;
    r.rpb = qs::gen_rows_per_block(Mg);
    M = Mg;
  }
  if (!r.gen && M > 4096)
    return fail(c, QS_ERR_UNSUPPORTED, "transform length " + std::to_string(n) +
                                           " needs a Bluestein length above 4096 (only sizes <= 2048 or powers of two "
                                           "<= 4096 are supported)");
  r.M = M;
  HIPCHK(c, upload(r.tw, twiddles(M), c->stream));
  for (int dir = -1; dir <= 1; dir += 2) {
    std::vector<float2> ch(n), vh(M);
    std::vector<std::complex<double>> cd(n), v(M, 0.0);
    for (int k = 0; k < n; k++) {
      const long long k2 = ((long long)k * k) % (2LL * n);
      const double ang = dir * M_PI * (double)k2 / n;
      cd[k] = std::complex<double>(cos(ang), sin(ang));
      ch[k] = make_float2((float)cd[k].real(), (float)cd[k].imag());
    }
    v[0] = std::conj(cd[0]);
    for (int k = 1; k < n; k++) v[k] = v[M - k] = std::conj(cd[k]);
    if (is_pow2(M)) host_fft_pow2(v, -1);
    else host_fft_any(v, -1);
    for (int k = 0; k < M; k++) vh[k] = make_float2((float)v[k].real(), (float)v[k].imag());
    HIPCHK(c, upload(dir < 0 ? r.chirp_f : r.chirp_b, ch, c->stream));
    HIPCHK(c, upload(dir < 0 ? r.vhat_f : r.vhat_b, vh, c->stream));
  }
  return QS_OK;
}

static void run_rowxform(qs_ctx* c, RowXform& r, float2* data, long nrows, int dir, float scale) {
  if (r.gen) {
    qs::GenRowXformArgs a{};
    a.data = data;
    a.nrows = nrows;
    a.n = r.n;
    a.M = r.M;
    a.rpb = r.rpb;
    a.scale = scale;
    a.plan = r.plan;
    a.tw = r.tw.as<float2>();
    a.chirp = (dir < 0 ? r.chirp_f : r.chirp_b).as<float2>();
    a.vhat = (dir < 0 ? r.vhat_f : r.vhat_b).as<float2>();
    qs::launch_gen_rowxform(a, dir, c->stream);
  } else if (r.M == 0) qs::launch_rows_fft(data, nrows, r.n, dir, r.tw.as<float2>(), scale, c->stream);
  else
    qs::launch_rows_bluestein(data, nrows, r.n, r.M, (dir < 0 ? r.chirp_f : r.chirp_b).as<float2>(),
                              (dir < 0 ? r.vhat_f : r.vhat_b).as<float2>(), r.tw.as<float2>(), scale, c->stream);
}

// k^2 limit of the transmission band limit (stemlib.cpp:2219-2262), float arithmetic
static float bl_k2max(const qs_params& p) {
  float k2max = p.potNx / (2.0F * p.potSizeX);
  const float t = p.potNy / (2.0F * p.potSizeY);
  if (t < k2max) k2max = t;
  k2max = (float)((2.0F / 3.0F) * k2max);
  return k2max * k2max;
}

// ---------------------------------------------------------------------------------
// Run without allocation.  The wave state, window offsets, detector partial sums
// and outputs of qs_run_stem_noalloc live in this array; each device holds its own
// copy (until the process ends), whose address hipGetSymbolAddress resolves once
// per device in qs_create.  Host code uses only that address, never the name.
static __device__ __align__(256) unsigned char scratch[QS_NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_ptr;        // device -> address of its copy of scratch
std::map<int, std::string> g_scratch_err;  // device -> why the address could not be resolved
}  // namespace

// Resolves the scratch address of device dev (the current device); once per device.
static void resolve_scratch(int dev) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr.count(dev) || g_scratch_err.count(dev)) return;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    g_scratch_err[dev] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
  else
    g_scratch_ptr[dev] = ptr;
}

static unsigned char* scratch_of(int dev, std::string& why) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it != g_scratch_ptr.end()) return static_cast<unsigned char*>(it->second);
  auto e = g_scratch_err.find(dev);
  why = e != g_scratch_err.end() ? e->second : "scratch address of this device not resolved";
  return nullptr;
}

extern "C" {

const char* qs_last_error(const qs_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int qs_create(const qs_params* p, qs_ctx** out) {
  if (!out) return QS_ERR_ARG;
  *out = nullptr;
  g_create_error.clear();
  if (!p) { g_create_error = "params is NULL"; return QS_ERR_ARG; }
  const bool fused = qs::supported_size(p->nx) && qs::supported_size(p->ny);
  if (!fused && (p->nx < 16 || p->ny < 16 || p->nx > 2048 || p->ny > 2048)) {
    g_create_error = "nx and ny must be powers of two in [64, 4096] or any size in [16, 2048]";
    return QS_ERR_UNSUPPORTED;
  }
  if (p->potNx < p->nx || p->potNy < p->ny) { g_create_error = "potNx/potNy smaller than nx/ny"; return QS_ERR_ARG; }
  if (p->ndet < 0 || p->ndet > QS_MAX_DET) { g_create_error = "ndet out of range"; return QS_ERR_ARG; }
  if (!(p->resolutionX > 0) || !(p->resolutionY > 0) || !(p->v0 > 0) || p->slices < 1) {
    g_create_error = "resolution, v0 and slices must be positive";
    return QS_ERR_ARG;
  }
  qs_ctx* c = new qs_ctx();
  c->p = *p;
  auto bad = [&](int code) {
    g_create_error = c->err;
    qs_destroy(c);
    return code;
  };
  if (hipStreamCreate(&c->stream) != hipSuccess || hipEventCreate(&c->ev0) != hipSuccess ||
      hipEventCreate(&c->ev1) != hipSuccess || hipStreamCreate(&c->stream2) != hipSuccess ||
      hipEventCreateWithFlags(&c->evfork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->evjoin, hipEventDisableTiming) != hipSuccess) {
    c->err = "hipStreamCreate/hipEventCreate failed (no HIP device?)";
    return bad(QS_ERR_HIP);
  }
  if (hipGetDevice(&c->device) != hipSuccess) {
    c->err = "hipGetDevice failed";
    return bad(QS_ERR_HIP);
  }
  resolve_scratch(c->device);  // address of qs_run_stem_noalloc's scratch on this device (once per device)
  if (const char* e = getenv("QS_ROWS_PER_WAVE")) c->rows_per_wave = atoi(e);  // tuning overrides
  if (const char* e = getenv("QS_TILES_PER_BLOCK")) c->tiles_per_block = atoi(e);
  if (const char* e = getenv("QS_STREAMS")) c->nstreams = atoi(e) == 1 ? 1 : 2;
  if (qs::set_col_lds_limits() != hipSuccess) {
    c->err = "hipFuncSetAttribute for the column sweep LDS size failed";
    return bad(QS_ERR_HIP);
  }
  const int nx = p->nx, ny = p->ny;
  // k^2 tables and band limit as propagate_slow builds them (stemlib.cpp:2782-2822), float arithmetic
  const float ax = p->resolutionX * nx, by = p->resolutionY * ny;
  c->kx2.resize(nx);
  c->ky2.resize(ny);
  for (int i = 0; i < nx; i++) {
    float k = (i > nx / 2) ? (float)(i - nx) / ax : (float)i / ax;
    c->kx2[i] = k * k;
  }
  for (int i = 0; i < ny; i++) {
    float k = (i > ny / 2) ? (float)(i - ny) / by : (float)i / by;
    c->ky2[i] = k * k;
  }
  float km = nx / (2.0F * ax);
  if (ny / (2.0F * by) < km) km = ny / (2.0F * by);
  km = (float)(2.0 / 3.0 * km);
  c->k2max = km * km;
  // band of ky columns that survive: [0, band_a] U [ny-band_b, ny-1]
  int a = -1;
  while (a + 1 <= ny / 2 && c->ky2[a + 1] < c->k2max) a++;
  int b = 0;
  while (b < ny / 2 - 1 && c->ky2[ny - 1 - b] < c->k2max) b++;
  for (int i = 0; i < ny; i++) {
    const bool in = (i <= a) || (i >= ny - b);
    if (in != (c->ky2[i] < c->k2max)) {
      c->err = "band-limit columns are not contiguous";
      return bad(QS_ERR_UNSUPPORTED);
    }
  }
  if (a < 0) { c->err = "band limit removes every column"; return bad(QS_ERR_ARG); }
  c->band_a = a;
  c->band_b = b;
  c->kc = a + 1 + b;
  c->generic = !fused;
  c->tc = fused ? qs::col_tile(nx) : 1;
  c->ntiles = (c->kc + c->tc - 1) / c->tc;
  // row pitch of the band-compact state: whole 128-byte lines, so that the two
  // halves of a line are adjacent column tiles (one block advances both)
  c->kcp = (c->ntiles * c->tc + 15) / 16 * 16;
  // propagator e^{-i pi dz lambda k^2}, double -> float, 1/(nx ny) folded into px
  const double wl = wavelength(p->v0), dz = p->sliceThickness;
  const double axd = (double)p->resolutionX * nx, byd = (double)p->resolutionY * ny;
  const double norm = 1.0 / ((double)nx * ny);
  std::vector<float2> hpx(nx), hpy(ny);
  for (int i = 0; i < nx; i++) {
    const double k = ((i > nx / 2) ? (double)(i - nx) : (double)i) / axd;
    const double t = M_PI * dz * wl * k * k;
    hpx[i] = make_float2((float)(cos(t) * norm), (float)(-sin(t) * norm));
  }
  for (int i = 0; i < ny; i++) {
    const double k = ((i > ny / 2) ? (double)(i - ny) : (double)i) / byd;
    const double t = M_PI * dz * wl * k * k;
    hpy[i] = make_float2((float)cos(t), (float)(-sin(t)));
  }
  // detector limits (stem3.cpp:1338-1347)
  for (int d = 0; d < p->ndet; d++) {
    float s0 = (float)(sin(p->det_rin[d] * 0.001) / wl);
    float s1 = (float)(sin(p->det_rout[d] * 0.001) / wl);
    c->det_k2in[d] = s0 * s0;
    c->det_k2out[d] = s1 * s1;
  }
  if (upload(c->twx, twiddles(nx), c->stream) != hipSuccess || upload(c->twy, twiddles(ny), c->stream) != hipSuccess ||
      upload(c->px, hpx, c->stream) != hipSuccess || upload(c->py, hpy, c->stream) != hipSuccess ||
      upload(c->dkx2, c->kx2, c->stream) != hipSuccess || upload(c->dky2, c->ky2, c->stream) != hipSuccess) {
    c->err = "device allocation of lookup tables failed";
    return bad(QS_ERR_HIP);
  }
  if (c->generic) {  // row transforms of the wave's lengths (Stockham or Bluestein)
    int rc = setup_rowxform(c, c->wx, nx);
    if (rc == QS_OK) rc = setup_rowxform(c, c->wy, ny);
    if (rc != QS_OK) return bad(rc);
    c->mixed = qs::gen_plan(nx, &c->gpx) && qs::gen_plan(ny, &c->gpy);
    if (c->mixed) {
      // compiled plans by default; QSTEM_COMPILED_PLANS=0 keeps the runtime-planned kernels
      const char* cp = getenv("QSTEM_COMPILED_PLANS");
      const bool comp = !(cp && strcmp(cp, "0") == 0);
      const int rc = set_mixed_shape(c, 0, 0, comp && qs::mix_compiled(nx) ? qs::kGenPlanCompiled : 0,
                                     comp && qs::mix_compiled(ny) ? qs::kGenPlanCompiled : 0);
      if (rc != QS_OK) return bad(rc);
    }
  }
  // transmission band-limit tables (stemlib.cpp:2219-2262), float arithmetic, Q1 optional
  if (p->bandlimittrans) {
    std::vector<float> kx2t(p->potNx), ky2t(p->potNy);
    for (int i = 0; i < p->potNx; i++) {
      float k = (i > p->potNx / 2) ? (float)(i - p->potNx) / p->potSizeX : (float)i / p->potSizeX;
      kx2t[i] = k * k;
    }
    const float byq = (p->compat & QS_COMPAT_Q1) ? p->by : p->potSizeY;
    for (int i = 0; i < p->potNy; i++) {
      float k = (i > p->potNy / 2) ? (float)(i - p->potNy) / byq : (float)i / p->potSizeY;
      ky2t[i] = k * k;
    }
    if (upload(c->pkx2, kx2t, c->stream) != hipSuccess || upload(c->pky2, ky2t, c->stream) != hipSuccess) {
      c->err = "device allocation failed";
      return bad(QS_ERR_HIP);
    }
    int rc = setup_rowxform(c, c->rx, p->potNx);
    if (rc == QS_OK) rc = setup_rowxform(c, c->ry, p->potNy);
    if (rc != QS_OK) return bad(rc);
    // band-compact sweeps when both lengths have Stockham plans and the ky band
    // is two contiguous column ranges, as for the wave
    if (qs::supported_size(p->potNx) && qs::supported_size(p->potNy)) {
      const float k2m = bl_k2max(*p);
      const int pny = p->potNy;
      int ba = -1;
      while (ba + 1 <= pny / 2 && ky2t[ba + 1] < k2m) ba++;
      int bb = 0;
      while (bb < pny / 2 - 1 && ky2t[pny - 1 - bb] < k2m) bb++;
      bool contiguous = ba >= 0;
      for (int i = 0; i < pny && contiguous; i++) contiguous = ((i <= ba) || (i >= pny - bb)) == (ky2t[i] < k2m);
      // the sweep kernels never compute the slots inside (N/3, N - N/3); with Q1 and
      // by > potSizeY the negative-ky half of the band reaches into them, and the
      // full-transform band limit below takes over
      const int third = pny / 3;
      contiguous = contiguous && ba <= third && bb <= third;
      if (contiguous) {
        c->bl.band_a = ba;
        c->bl.band_b = bb;
        c->bl.kc = ba + 1 + bb;
        const int tc = qs::col_tile(p->potNx);
        c->bl.ntiles = (c->bl.kc + tc - 1) / tc;
        c->bl.kcp = (c->bl.ntiles * tc + 15) / 16 * 16;
        const float s = (float)(1.0 / ((double)p->potNx * p->potNy));
        if (upload(c->blpx, std::vector<float2>(p->potNx, make_float2(s, 0.f)), c->stream) != hipSuccess ||
            upload(c->blpy, std::vector<float2>(pny, make_float2(1.f, 0.f)), c->stream) != hipSuccess) {
          c->err = "device allocation failed";
          return bad(QS_ERR_HIP);
        }
        c->bl.on = true;
      }
    }
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess) { c->err = "stream sync failed"; return bad(QS_ERR_HIP); }
  *out = c;
  return QS_OK;
}

static void drop_graph(qs_ctx* c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->gexec2) (void)hipGraphExecDestroy(c->gexec2);
  c->gexec = c->gexec2 = nullptr;
  c->gkey.clear();
}

// Called wherever the potential, probe or optics change: the captured graph's
// kernel arguments and the forward pass kept for the adjoint no longer hold.
static void invalidate(qs_ctx* c) {
  drop_graph(c);
  c->adj_saved = false;
}

void qs_destroy(qs_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graph(c);
  DevBuf* bufs[] = {&c->cfgidx, &c->twx, &c->twy, &c->px, &c->py, &c->dkx2, &c->dky2, &c->probe, &c->trans, &c->scratch, &c->V,
                    &c->state, &c->ipos, &c->partial, &c->partial2, &c->det, &c->diff, &c->exitw, &c->total, &c->pkx2, &c->pky2,
                    &c->probe_partial, &c->rx.tw, &c->rx.chirp_f, &c->rx.vhat_f, &c->rx.chirp_b, &c->rx.vhat_b,
                    &c->ry.tw, &c->ry.chirp_f, &c->ry.vhat_f, &c->ry.chirp_b, &c->ry.vhat_b,
                    &c->atoms, &c->kinds, &c->lutptrs, &c->spl, &c->luttmp, &c->lutscr,
                    &c->reach, &c->tcount, &c->tstart, &c->tlist,
                    &c->lx.tw, &c->lx.chirp_f, &c->lx.vhat_f, &c->lx.chirp_b, &c->lx.vhat_b,
                    &c->lz.tw, &c->lz.chirp_f, &c->lz.vhat_f, &c->lz.chirp_b, &c->lz.vhat_b,
                    &c->gx, &c->target, &c->lossp, &c->gT, &c->gprobe, &c->gV, &c->dsum,
                    &c->blpx, &c->blpy, &c->blstate, &c->blstate2, &c->adj_ipos,
                    &c->wx.tw, &c->wx.chirp_f, &c->wx.vhat_f, &c->wx.chirp_b, &c->wx.vhat_b,
                    &c->wy.tw, &c->wy.chirp_f, &c->wy.vhat_f, &c->wy.chirp_b, &c->wy.vhat_b, &c->gscratch, &c->gtiles};
  for (DevBuf* b : bufs) b->release();
  for (DevBuf& b : c->luts) b.release();
  for (DevBuf& b : c->oluts) b.release();
  c->olutptrs.release();
  for (DevBuf& b : c->ckpt) b.release();
  for (hipEvent_t e : c->evpool) (void)hipEventDestroy(e);
  for (int b = 0; b < 2; b++) {
    if (c->evrow[b]) (void)hipEventDestroy(c->evrow[b]);
    if (c->evbl[b]) (void)hipEventDestroy(c->evbl[b]);
  }
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->evfork) (void)hipEventDestroy(c->evfork);
  if (c->evjoin) (void)hipEventDestroy(c->evjoin);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  delete c;
}

// Band-limit scratch: band-compact rows of this many bytes at most per call chunk.
static constexpr size_t kBandScratchBytes = size_t(1) << 30;

// Band limit of initSTEMSlices (FFT, 2/3 mask with 1/(potNx potNy), IFFT) applied in
// place to `tr` [nslice][potNx][potNy]; the operator is self-adjoint, so the adjoint
// uses it unchanged on the transmission gradient.
// Band limit as band-compact sweeps, chunks of slices at a time, each slice a
// "probe" of the sweep kernels: row FFT into the ky band (rows mode 3), column
// FFT / propagator 1/(nx ny) inside the disk k^2 < k2max / column IFFT, row IFFT
// written back to `tr` (rows mode 2) or, with gv given, turned into
// dL/dV = sigma Im(g conj(exp(i sigma V))) of the slices' potential vpot (mode 4).
// rows_done (one slice, gv given): the first pass already ran (the adjoint's row sweep
// stored the row FFT band-compact into *scratch); tr is not used.
static int bl_sweeps(qs_ctx* c, float2* tr, int nslice, const float* vpot, float* gv, float sigma,
                     hipStream_t st = nullptr, DevBuf* scratch = nullptr, bool rows_done = false) {
  const qs_params& p = c->p;
  const int nx = p.potNx, ny = p.potNy;
  const size_t slice = (size_t)nx * ny;
  const size_t per = sizeof(float2) * (size_t)nx * c->bl.kcp;
  const int q = (int)std::max<size_t>(1, std::min<size_t>((size_t)nslice, kBandScratchBytes / per));
  if (!st) st = c->stream;
  if (!scratch) scratch = &c->blstate;
  if (rows_done && (nslice != 1 || !gv || !scratch->p || scratch->bytes < per)) {
    c->err = "band limit: the fused first pass needs one slice, dL/dV and its state";
    return QS_ERR_STATE;
  }
  HIPCHK(c, scratch->ensure(per * q));
  qs::SweepArgs a{};
  a.nx = nx;
  a.ny = ny;
  a.kc = c->bl.kc;
  a.kcp = c->bl.kcp;
  a.band_a = c->bl.band_a;
  a.band_b = c->bl.band_b;
  a.ntiles = c->bl.ntiles;
  a.potNy = ny;
  a.twx = c->rx.tw.as<float2>();
  a.twy = c->ry.tw.as<float2>();
  a.px = c->blpx.as<float2>();
  a.py = c->blpy.as<float2>();
  a.kx2 = c->pkx2.as<float>();
  a.ky2 = c->pky2.as<float>();
  a.k2max = bl_k2max(p);
  a.rec_bin = -1;
  a.state = scratch->as<float2>();
  a.gv_sigma = sigma;
  for (int s0 = 0; s0 < nslice; s0 += q) {
    a.nprobe = std::min(q, nslice - s0);
    a.probe = tr ? tr + (size_t)s0 * slice : nullptr;
    a.out_wave = tr ? tr + (size_t)s0 * slice : nullptr;
    if (gv) {
      a.vpot = vpot + (size_t)s0 * slice;
      a.gv = gv + (size_t)s0 * slice;
    }
    if (!rows_done) qs::launch_sweep_rows(a, 3, st);
    qs::launch_sweep_cols(a, false, false, st);
    qs::launch_sweep_rows(a, gv ? 4 : 2, st);
  }
  HIPCHK(c, hipGetLastError());
  return QS_OK;
}

static int bandlimit_buf(qs_ctx* c, float2* tr, int nslice) {
  const qs_params& p = c->p;
  const int nx = p.potNx, ny = p.potNy;
  const float k2max = bl_k2max(p);
  if (c->bl.on) return bl_sweeps(c, tr, nslice, nullptr, nullptr, 0.f);
  HIPCHK(c, c->scratch.ensure(sizeof(float2) * (size_t)nslice * nx * ny));
  float2* sc = c->scratch.as<float2>();
  run_rowxform(c, c->ry, tr, (long)nslice * nx, -1, 1.0f);
  qs::launch_transpose(tr, sc, nslice, nx, ny, c->stream);
  run_rowxform(c, c->rx, sc, (long)nslice * ny, -1, 1.0f);
  qs::launch_bandlimit_T(sc, nslice, ny, nx, c->pkx2.as<float>(), c->pky2.as<float>(), k2max,
                         (float)(1.0 / ((double)nx * ny)), c->stream);
  run_rowxform(c, c->rx, sc, (long)nslice * ny, +1, 1.0f);
  qs::launch_transpose(sc, tr, nslice, ny, nx, c->stream);
  run_rowxform(c, c->ry, tr, (long)nslice * nx, +1, 1.0f);
  HIPCHK(c, hipGetLastError());
  return QS_OK;
}

// The transmission functions a context runs on: its own, or those of the
// context it shares them with (qs_share_trans).
static qs_ctx* trans_owner(qs_ctx* c) { return c->trans_src ? c->trans_src : c; }
static float2* trans_base(qs_ctx* c) { return trans_owner(c)->trans.as<float2>(); }
static int trans_slices_of(qs_ctx* c) { return trans_owner(c)->trans_slices; }
static int nconfig_of(qs_ctx* c) { return trans_owner(c)->nconfig; }

// Room for nconfig configurations of nslice slices; returns the selected
// configuration's first slice.  Configuration 0 fixes the slice count (and may
// reallocate, dropping the others); the others must match it.
static int trans_slot(qs_ctx* c, int nslice, const char* who, float2** out) {
  const size_t per = (size_t)nslice * c->p.potNx * c->p.potNy;
  if (c->trans_src) return fail(c, QS_ERR_STATE, std::string(who) + ": this context uses another's transmission functions");
  if (c->cur_config > 0 && nslice != c->trans_slices)
    return fail(c, QS_ERR_ARG, std::string(who) + ": every configuration needs the slice count of configuration 0");
  if (c->cur_config == 0) {
    HIPCHK(c, c->trans.ensure(sizeof(float2) * per * c->nconfig));
    c->trans_slices = nslice;
  }
  *out = c->trans.as<float2>() + per * c->cur_config;
  return QS_OK;
}

// initSTEMSlices on the device-resident potential c->V, into the selected configuration.
static int build_trans(qs_ctx* c, int nslice) {
  const qs_params& p = c->p;
  const size_t n = (size_t)nslice * p.potNx * p.potNy;
  float2* t = nullptr;
  int rc = trans_slot(c, nslice, "potential", &t);
  if (rc != QS_OK) return rc;
  // sigma = gamma*lambda with gamma = 1 + v0/511 in float (stemlib.cpp:2271-2274)
  const float mm0 = 1.0F + p.v0 / 511.0F;
  const double scale = mm0 * wavelength(p.v0);
  qs::launch_phase_grating(c->V.as<float>(), t, n, scale, c->stream);
  HIPCHK(c, hipGetLastError());
  if (p.bandlimittrans) {
    rc = bandlimit_buf(c, t, nslice);
    if (rc != QS_OK) return rc;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return QS_OK;
}

int qs_set_potential(qs_ctx* c, const float* V, int nslice) {
  if (!c || !V || nslice < 1) return fail(c, QS_ERR_ARG, "qs_set_potential: bad arguments");
  invalidate(c);  // kernel arguments baked into a captured graph may change
  const qs_params& p = c->p;
  const size_t n = (size_t)nslice * p.potNx * p.potNy;
  HIPCHK(c, c->V.ensure(sizeof(float) * n));
  // host or device memory (hipMemcpyDefault: a torch tensor's storage on this device)
  HIPCHK(c, hipMemcpyAsync(c->V.p, V, sizeof(float) * n, hipMemcpyDefault, c->stream));
  c->built.valid = false;  // the potential no longer comes from the last qs_build_potential
  const int rc = build_trans(c, nslice);
  c->v_valid = rc == QS_OK;
  return rc;
}

// Akima spline coefficients (the fit of stem3/stemutil.cpp:975-1041), host set-up of 123 nodes.
static void akima(const double* x, const double* y, double* b, double* cc, double* d, int n) {
  const double small = 1.0e-25;
  double m5 = (y[3] - y[2]) / (x[3] - x[2]), m4 = (y[2] - y[1]) / (x[2] - x[1]), m3 = (y[1] - y[0]) / (x[1] - x[0]);
  double m2 = 2 * m3 - m4, m1 = 2 * m2 - m3;
  double m54 = fabs(m5 - m4), m43 = fabs(m4 - m3), m32 = fabs(m3 - m2), m21 = fabs(m2 - m1);
  double t1 = (m43 + m21 > small) ? (m43 * m2 + m21 * m3) / (m43 + m21) : 0.5 * (m2 + m3);
  for (int i = 0; i < n - 1; i++) {
    const double t2 = (m54 + m32 > small) ? (m54 * m3 + m32 * m4) / (m54 + m32) : 0.5 * (m3 + m4);
    const double h = x[i + 1] - x[i];
    b[i] = t1;
    cc[i] = (3 * m3 - 2 * t1 - t2) / h;
    d[i] = (t1 + t2 - 2 * m3) / (h * h);
    m1 = m2; m2 = m3; m3 = m4; m4 = m5;
    m5 = (i < n - 4) ? (y[i + 4] - y[i + 3]) / (x[i + 4] - x[i + 3]) : 2 * m4 - m3;
    m21 = m32; m32 = m43; m43 = m54; m54 = fabs(m5 - m4);
    t1 = t2;
  }
  b[n - 1] = cc[n - 1] = d[n - 1] = 0;
}

int qs_set_scattering_table(qs_ctx* c, const double* table) {
  if (!c || !table) return fail(c, QS_ERR_ARG, "qs_set_scattering_table: bad arguments");
  c->scat.assign(table, table + 93 * 123);
  return QS_OK;
}

int qs_set_ion_scattering_table(qs_ctx* c, const double* table) {
  if (!c || !table) return fail(c, QS_ERR_ARG, "qs_set_ion_scattering_table: bad arguments");
  c->scat_offs.assign(table, table + 93 * 123);
  return QS_OK;
}

// Taper of a scattering-factor table's s grid at smax (stemlib.cpp:1390-1413, :1598-1620):
// the last three nodes become 1.0, 1.1, 1.2 smax and nodes beyond smax step down with zero
// rows.
static void taper_table(std::vector<double>& tab, double smax) {
  double* sgrid = tab.data();
  sgrid[122] = 1.2 * smax;
  sgrid[121] = 1.1 * smax;
  sgrid[120] = smax;
  if (sgrid[119] > sgrid[120]) {
    for (int ix = 0; ix < 123 - 10; ix++) {
      if (sgrid[119 - ix] < sgrid[120] - 0.001 * (ix + 1)) break;
      sgrid[119 - ix] = sgrid[120] - 0.001 * (ix + 1);
      for (int r = 1; r < 93; r++) tab[(size_t)r * 123 + 119 - ix] = 0;
    }
  }
}

int qs_build_potential(qs_ctx* c, const qs_atom* atoms, int natom, const qs_pot_params* pp, float* V_out) {
  if (!c || !pp || natom < 0 || (natom > 0 && !atoms)) return fail(c, QS_ERR_ARG, "qs_build_potential: bad arguments");
  invalidate(c);  // kernel arguments baked into a captured graph may change
  if (c->scat.size() != 93 * 123) return fail(c, QS_ERR_STATE, "qs_build_potential: scattering table not set");
  const qs_params& p = c->p;
  c->built.valid = false;
  bool charged = false;
  for (int i = 0; i < natom; i++) {
    if (atoms[i].Znum < 0 || atoms[i].Znum > 92) return fail(c, QS_ERR_ARG, "qs_build_potential: Z out of range");
    if (atoms[i].Znum > 0 && atoms[i].q != 0.f) charged = true;
  }
  if (charged && c->scat_offs.size() != 93 * 123)
    return fail(c, QS_ERR_STATE, "qs_build_potential: charged atoms need qs_set_ion_scattering_table");
  // LUT geometry (stemlib.cpp:1366-1388), float products as the reference
  const float R = pp->atomRadius, dx = p.resolutionX, dz = p.sliceThickness;
  LutArgsHost g;
  g.nx = 2 * 2 * (int)ceil(R / dx);
  int nzp = (int)floor(2 * dz / dx);
  if (2.0 * (nzp >> 1) == nzp) nzp += 1;
  g.nzp = nzp;
  g.nz = (2 * (int)ceil(R / dz)) * nzp;
  g.dkx = 0.5 * 2 / (double)(float)(g.nx * dx);
  g.dkz = nzp / (double)(float)(g.nz * dz);
  const double smax = 0.5 * g.nx * g.dkx / 2.0;
  if (g.nx > 2048 || g.nz > 2048) return fail(c, QS_ERR_UNSUPPORTED, "qs_build_potential: LUT grid above 2048");
  // taper of the shared s grid (stemlib.cpp:1390-1413), and of the ion table's (:1598-1620)
  std::vector<double> tab = c->scat;
  taper_table(tab, smax);
  const double* sgrid = tab.data();
  std::vector<double> otab = c->scat_offs;
  if (charged) taper_table(otab, smax);
  // distinct LUT kinds: (Z, B) keys, or Z only with the B of the lowest atom in z (compat Q6)
  std::vector<std::pair<int, float>> keys;
  std::vector<int> kind(natom, 0);
  std::vector<int> order(natom);
  for (int i = 0; i < natom; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return atoms[a].z < atoms[b].z; });
  for (int oi = 0; oi < natom; oi++) {
    const int i = order[oi];
    if (atoms[i].Znum == 0) continue;
    const float B = pp->tds ? 0.f : atoms[i].dw;
    int k = -1;
    for (size_t j = 0; j < keys.size(); j++)
      if (keys[j].first == atoms[i].Znum && (pp->compat_q6 || keys[j].second == B)) { k = (int)j; break; }
    if (k < 0) { keys.push_back({atoms[i].Znum, B}); k = (int)keys.size() - 1; }
    kind[i] = k;
  }
  std::vector<char> need_off(keys.size(), 0);  // kinds with a charged atom get an offset LUT
  for (int i = 0; i < natom; i++)
    if (atoms[i].Znum > 0 && atoms[i].q != 0.f) need_off[kind[i]] = 1;
  int rc = setup_rowxform(c, c->lx, g.nx);
  if (rc == QS_OK) rc = setup_rowxform(c, c->lz, g.nz);
  if (rc != QS_OK) return rc;
  qs::LutArgs la{};
  la.nx = g.nx; la.nz = g.nz; la.nzPerSlice = nzp;
  la.dkx = g.dkx; la.dkz = g.dkz; la.smax2 = smax * smax;
  la.zPos = -2.0 * M_PI * ((double)dz / nzp * ((nzp - 1) / 2));
  la.dky = g.dkx;
  // normalisation switch: slice sum of an atom = f_e(0) instead of 47.8658 f_e(0) / (4 L_z)
  la.scale = pp->lut_norm_fix ? 4.0 * g.nz * ((double)dz / nzp) / 47.8658 : 1.0;
  const int Nr = g.nx / 2, Nz = g.nz / 2;
  c->luts.resize(keys.size());
  for (DevBuf& b : c->oluts) b.release();
  c->oluts.resize(keys.size());
  HIPCHK(c, c->spl.ensure(sizeof(double) * 7 * 123));
  HIPCHK(c, c->luttmp.ensure(sizeof(float2) * (size_t)g.nx * g.nz));
  HIPCHK(c, c->lutscr.ensure(sizeof(float2) * (size_t)g.nx * g.nz));
  std::vector<const float*> lutp(keys.size()), olutp(keys.size(), nullptr);
  std::vector<double> spl(7 * 123);
  // one LUT: spline fitted to (sg, row) for the (kx, kz) term, q_y sum over the nodes
  // (qsg, qrow) with that spline's coefficients and step dky (stemlib.cpp:1419-1532)
  auto make_lut = [&](const double* sg, const double* row, const double* qsg, const double* qrow, double dky,
                      double B, DevBuf& out) -> int {
    std::copy(sg, sg + 123, spl.begin());
    std::copy(row, row + 123, spl.begin() + 123);
    akima(&spl[0], &spl[123], &spl[246], &spl[369], &spl[492], 123);
    std::copy(qsg, qsg + 123, spl.begin() + 615);
    std::copy(qrow, qrow + 123, spl.begin() + 738);
    HIPCHK(c, hipMemcpyAsync(c->spl.p, spl.data(), sizeof(double) * spl.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, out.ensure(sizeof(float) * (size_t)Nr * Nz));
    float2* tmp = c->luttmp.as<float2>();
    float2* scr = c->lutscr.as<float2>();
    qs::LutArgs lk = la;
    lk.dky = dky;
    qs::launch_lut_fk(c->spl.as<double>(), lk, B, tmp, c->stream);
    run_rowxform(c, c->lx, tmp, g.nz, +1, 1.0f);
    qs::launch_transpose(tmp, scr, 1, g.nz, g.nx, c->stream);
    run_rowxform(c, c->lz, scr, g.nx, +1, 1.0f);
    qs::launch_transpose(scr, tmp, 1, g.nx, g.nz, c->stream);
    qs::launch_lut_extract(tmp, lk, out.as<float>(), c->stream);
    HIPCHK(c, hipStreamSynchronize(c->stream));  // spl is reused by the next LUT
    return QS_OK;
  };
  // getAtomPotentialOffset3D's q_y step uses resolutionY (stemlib.cpp:1594)
  const int lny = 2 * 2 * (int)ceil(R / p.resolutionY);
  const double dky_off = 0.5 * 2 / (double)(float)(lny * p.resolutionY);
  for (size_t k = 0; k < keys.size(); k++) {
    const double* row = tab.data() + (size_t)keys[k].first * 123;
    rc = make_lut(sgrid, row, sgrid, row, g.dkx, (double)keys[k].second, c->luts[k]);
    if (rc != QS_OK) return rc;
    lutp[k] = c->luts[k].as<float>();
    if (need_off[k]) {
      const double* orow = otab.data() + (size_t)keys[k].first * 123;
      const double* qsg = pp->compat_q5 ? sgrid : otab.data();
      const double* qrow = pp->compat_q5 ? row : orow;
      rc = make_lut(otab.data(), orow, qsg, qrow, dky_off, (double)keys[k].second, c->oluts[k]);
      if (rc != QS_OK) return rc;
      olutp[k] = c->oluts[k].as<float>();
    }
  }
  // atoms, kinds, LUT pointers
  const size_t nV = (size_t)p.slices * p.potNx * p.potNy;
  HIPCHK(c, c->V.ensure(sizeof(float) * nV));
  HIPCHK(c, hipMemsetAsync(c->V.p, 0, sizeof(float) * nV, c->stream));
  if (natom > 0 && !keys.empty()) {
    HIPCHK(c, c->atoms.ensure(sizeof(qs_atom) * natom));
    HIPCHK(c, c->kinds.ensure(sizeof(int) * natom));
    HIPCHK(c, c->lutptrs.ensure(sizeof(float*) * keys.size()));
    HIPCHK(c, hipMemcpyAsync(c->atoms.p, atoms, sizeof(qs_atom) * natom, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->kinds.p, kind.data(), sizeof(int) * natom, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->lutptrs.p, lutp.data(), sizeof(float*) * keys.size(), hipMemcpyHostToDevice, c->stream));
    if (charged) {
      HIPCHK(c, c->olutptrs.ensure(sizeof(float*) * keys.size()));
      HIPCHK(c, hipMemcpyAsync(c->olutptrs.p, olutp.data(), sizeof(float*) * keys.size(), hipMemcpyHostToDevice,
                               c->stream));
    }
    qs::ScatterArgs sa{};
    sa.potNx = p.potNx; sa.potNy = p.potNy; sa.slices = p.slices;
    sa.dx = p.resolutionX; sa.dy = p.resolutionY; sa.sliceThickness = dz;
    sa.potOffsetX = pp->potOffsetX; sa.potOffsetY = pp->potOffsetY; sa.czOffset = pp->czOffset;
    sa.atomRadius = R; sa.centerSlices = pp->centerSlices; sa.periodic = pp->periodicXY;
    sa.iRadX = (int)ceil(R / p.resolutionX); sa.iRadY = (int)ceil(R / p.resolutionY); sa.iRadZ = (int)ceil(R / dz);
    sa.Nr = Nr; sa.Nz = Nz; sa.nzSub = nzp;
    sa.dr = pp->lut_dr_fix ? p.resolutionX : p.resolutionX / 2;
    // gather: per-atom reach, count and list the atoms of each bin (pixel tile, block
    // of slices), then one thread per voxel sums them (one store per voxel instead of
    // an atomic per sample)
    const int tiles_x = (p.potNx + qs::kPotTile - 1) / qs::kPotTile;
    const int tiles_y = (p.potNy + qs::kPotTile - 1) / qs::kPotTile;
    const int zblocks = (p.slices + qs::kPotZc - 1) / qs::kPotZc;
    const int ntile = tiles_x * tiles_y * zblocks;  // bins: pixel tile x block of slices
    HIPCHK(c, c->reach.ensure(sizeof(qs::AtomReach) * natom));
    HIPCHK(c, c->tcount.ensure(sizeof(int) * ntile));
    HIPCHK(c, c->tstart.ensure(sizeof(int) * (ntile + 1)));
    qs::AtomReach* reach = c->reach.as<qs::AtomReach>();
    qs::launch_atom_reach(c->atoms.as<qs::QsAtomDev>(), natom, c->kinds.as<int>(), sa, reach, c->stream);
    HIPCHK(c, hipMemsetAsync(c->tcount.p, 0, sizeof(int) * ntile, c->stream));
    qs::launch_bin_atoms(reach, natom, sa, zblocks, c->tcount.as<int>(), nullptr, false, c->stream);
    std::vector<int> cnt(ntile), start(ntile + 1, 0);
    HIPCHK(c, hipMemcpyAsync(cnt.data(), c->tcount.p, sizeof(int) * ntile, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    long long total = 0;
    for (int t = 0; t < ntile; t++) {
      start[t] = (int)total;
      total += cnt[t];
      if (total > 0x7fffffffLL) return fail(c, QS_ERR_ARG, "qs_build_potential: too many atom-tile pairs");
    }
    start[ntile] = (int)total;
    HIPCHK(c, c->tlist.ensure(sizeof(int) * (size_t)std::max<long long>(1, total)));
    HIPCHK(c, hipMemcpyAsync(c->tstart.p, start.data(), sizeof(int) * (ntile + 1), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->tcount.p, start.data(), sizeof(int) * ntile, hipMemcpyHostToDevice, c->stream));
    qs::launch_bin_atoms(reach, natom, sa, zblocks, c->tcount.as<int>(), c->tlist.as<int>(), true, c->stream);
    qs::launch_gather(reach, c->tstart.as<int>(), c->tlist.as<int>(), c->lutptrs.as<const float*>(),
                      charged ? c->olutptrs.as<const float*>() : nullptr, sa, c->V.as<float>(), c->stream);
    HIPCHK(c, hipStreamSynchronize(c->stream));  // the host arrays above are copied from
  }
  HIPCHK(c, hipGetLastError());
  if (V_out) HIPCHK(c, hipMemcpyAsync(V_out, c->V.p, sizeof(float) * nV, hipMemcpyDeviceToHost, c->stream));
  rc = build_trans(c, p.slices);
  c->v_valid = rc == QS_OK;
  if (rc == QS_OK) {
    c->built.Z.resize(natom);
    c->built.q.resize(natom);
    for (int i = 0; i < natom; i++) {
      c->built.Z[i] = atoms[i].Znum;
      c->built.q[i] = atoms[i].Znum > 0 ? atoms[i].q : 0.f;
    }
    c->built.pp = *pp;
    c->built.valid = true;
  }
  return rc;
}

int qs_set_trans(qs_ctx* c, const float* trans, int nslice) {
  if (!c || !trans || nslice < 1) return fail(c, QS_ERR_ARG, "qs_set_trans: bad arguments");
  invalidate(c);  // kernel arguments baked into a captured graph may change
  const size_t n = (size_t)nslice * c->p.potNx * c->p.potNy;
  float2* t = nullptr;
  const int rc = trans_slot(c, nslice, "qs_set_trans", &t);
  if (rc != QS_OK) return rc;
  HIPCHK(c, hipMemcpyAsync(t, trans, sizeof(float2) * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->v_valid = false;
  c->built.valid = false;
  return QS_OK;
}

int qs_get_trans(qs_ctx* c, float* out) {
  if (!c || !out) return fail(c, QS_ERR_ARG, "qs_get_trans: bad arguments");
  if (trans_slices_of(c) < 1) return fail(c, QS_ERR_STATE, "no transmission function set");
  const size_t n = (size_t)trans_slices_of(c) * c->p.potNx * c->p.potNy;
  const int k = c->trans_src ? 0 : c->cur_config;
  HIPCHK(c, hipMemcpyAsync(out, trans_base(c) + n * k, sizeof(float2) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return QS_OK;
}

int qs_make_probe_at(qs_ctx* c, double x, double y, float* out) {
  if (!c) return QS_ERR_ARG;
  if (!std::isfinite(x) || !std::isfinite(y)) return fail(c, QS_ERR_ARG, "qs_make_probe_at: position not finite");
  invalidate(c);  // kernel arguments baked into a captured graph may change
  const qs_params& p = c->p;
  const int nx = p.nx, ny = p.ny;
  ProbeArgs a{};
  a.nx = nx;
  a.ny = ny;
  a.ax = nx * (double)p.resolutionX;
  a.by = ny * (double)p.resolutionY;
  a.rx = 1.0 / a.ax;
  a.ry = 1.0 / a.by;
  a.rx2 = a.rx * a.rx;
  a.ry2 = a.ry * a.ry;
  const double v0 = p.v0;
  a.wavlen = 12.26 / sqrt(v0 * 1.e3 + v0 * v0 * 0.9788);  // probe's own formula, stemlib.cpp:2014
  const double alpha = 0.001 * p.alpha;
  a.k2max = sin(alpha) / a.wavlen;
  a.k2max *= a.k2max;
  a.pixel = a.rx2 + a.ry2;
  a.scale = (double)(float)(1.0 / sqrt((double)nx * (double)ny));  // "(float) scale" at stemlib.cpp:2101
  a.df_eff = (double)p.df0 + (double)p.Cc * (double)p.dE_E;
  a.astigMag = p.astigMag;
  a.astigAngle = p.astigAngle;
  a.a33 = p.a33; a.a31 = p.a31; a.a44 = p.a44; a.a42 = p.a42; a.a55 = p.a55;
  a.a53 = p.a53; a.a51 = p.a51; a.a66 = p.a66; a.a64 = p.a64;
  a.a62 = (p.compat & QS_COMPAT_Q2) ? 0.0 : (double)p.a62;
  a.phi33 = p.phi33; a.phi31 = p.phi31; a.phi44 = p.phi44; a.phi42 = p.phi42; a.phi55 = p.phi55;
  a.phi53 = p.phi53; a.phi51 = p.phi51; a.phi66 = p.phi66; a.phi64 = p.phi64; a.phi62 = p.phi62;
  a.Cs = p.Cs;
  a.C5 = p.C5;
  a.on3 = p.a33 > 0 || p.a31 > 0;
  a.on4 = p.a44 > 0 || p.a42 > 0 || p.Cs != 0;
  a.on5 = p.a55 > 0 || p.a53 > 0 || p.a51 > 0;
  a.on6 = p.a66 > 0 || p.a64 > 0 || p.C5 != 0 || (!(p.compat & QS_COMPAT_Q2) && p.a62 > 0);
  a.sx = a.ax - x;  // probe(): dx = ax - dx (stemlib.cpp:1974-1975)
  a.sy = a.by - y;
  a.ismoth = p.ismoth;
  a.gaussFlag = p.gaussFlag;
  a.gaussScale = p.gaussScale;
  a.aAIS = p.aAIS;
  a.resX = p.resolutionX;
  a.resY = p.resolutionY;
  a.edge = 5.0 * sqrt(0.5 * ((double)p.resolutionX * p.resolutionX + (double)p.resolutionY * p.resolutionY));
  const size_t n = (size_t)nx * ny;
  HIPCHK(c, c->probe.ensure(sizeof(float2) * n));
  HIPCHK(c, c->scratch.ensure(sizeof(float2) * n > c->scratch.bytes ? sizeof(float2) * n : c->scratch.bytes));
  HIPCHK(c, c->probe_partial.ensure(sizeof(double) * 1024));
  float2* pr = c->probe.as<float2>();
  float2* sc = c->scratch.as<float2>();
  qs::launch_probe_kspace(a, pr, c->stream);
  // 2D inverse FFT (unnormalised): rows over ky, transpose, rows over kx, transpose back
  if (c->generic) run_rowxform(c, c->wy, pr, nx, +1, 1.0f);
  else qs::launch_rows_fft(pr, nx, ny, +1, c->twy.as<float2>(), 1.0f, c->stream);
  qs::launch_transpose(pr, sc, 1, nx, ny, c->stream);
  if (c->generic) run_rowxform(c, c->wx, sc, ny, +1, 1.0f);
  else qs::launch_rows_fft(sc, ny, nx, +1, c->twx.as<float2>(), 1.0f, c->stream);
  qs::launch_transpose(sc, pr, 1, ny, nx, c->stream);
  qs::launch_probe_window(a, pr, c->probe_partial.as<double>(), c->stream);
  HIPCHK(c, hipGetLastError());
  if (out) HIPCHK(c, hipMemcpyAsync(out, pr, sizeof(float2) * n, hipMemcpyDeviceToHost, c->stream));
  std::vector<double> parts(qs::probe_window_blocks(nx * ny));
  HIPCHK(c, hipMemcpyAsync(parts.data(), c->probe_partial.p, sizeof(double) * parts.size(), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double tot = 0;
  for (double v : parts) tot += v;
  c->probe_norm = sqrt((double)n / tot);
  c->probe_args = a;
  c->have_probe = true;
  c->custom_wave = false;
  return QS_OK;
}

int qs_make_probe(qs_ctx* c, float* out) {
  if (!c) return QS_ERR_ARG;
  // the driver's window centre, nx/2*resolutionX in float (stem3.cpp:2838)
  const float cx = (float)(c->p.nx / 2) * c->p.resolutionX, cy = (float)(c->p.ny / 2) * c->p.resolutionY;
  return qs_make_probe_at(c, (double)cx, (double)cy, out);
}

int qs_set_optics(qs_ctx* c, const qs_params* p) {
  if (!c || !p) return fail(c, QS_ERR_ARG, "qs_set_optics: bad arguments");
  invalidate(c);  // kernel arguments baked into a captured graph may change
  qs_params& q = c->p;
  q.Cs = p->Cs; q.C5 = p->C5; q.df0 = p->df0; q.astigMag = p->astigMag; q.astigAngle = p->astigAngle;
  q.a33 = p->a33; q.a31 = p->a31; q.a44 = p->a44; q.a42 = p->a42; q.a55 = p->a55;
  q.a53 = p->a53; q.a51 = p->a51; q.a66 = p->a66; q.a64 = p->a64; q.a62 = p->a62;
  q.phi33 = p->phi33; q.phi31 = p->phi31; q.phi44 = p->phi44; q.phi42 = p->phi42; q.phi55 = p->phi55;
  q.phi53 = p->phi53; q.phi51 = p->phi51; q.phi66 = p->phi66; q.phi64 = p->phi64; q.phi62 = p->phi62;
  q.alpha = p->alpha; q.aAIS = p->aAIS; q.Cc = p->Cc; q.dE_E = p->dE_E;
  q.ismoth = p->ismoth; q.gaussFlag = p->gaussFlag; q.gaussScale = p->gaussScale;
  q.compat = (q.compat & ~QS_COMPAT_Q2) | (p->compat & QS_COMPAT_Q2);
  c->have_probe = false;
  return QS_OK;
}

int qs_set_incident_wave(qs_ctx* c, const float* psi) {
  if (!c || !psi) return fail(c, QS_ERR_ARG, "qs_set_incident_wave: bad arguments");
  invalidate(c);
  const size_t n = (size_t)c->p.nx * c->p.ny;
  HIPCHK(c, c->probe.ensure(sizeof(float2) * n));
  HIPCHK(c, hipMemcpyAsync(c->probe.p, psi, sizeof(float2) * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->have_probe = true;
  c->custom_wave = true;  // no aberration function behind it: no probe gradients
  return QS_OK;
}

int qs_device_memory(long long* current_bytes, long long* peak_bytes, int reset_peak) {
  if (current_bytes) *current_bytes = g_dev_bytes.load();
  if (peak_bytes) *peak_bytes = g_dev_peak.load();
  if (reset_peak) g_dev_peak.store(g_dev_bytes.load());
  return QS_OK;
}

int qs_record_bins_repeat(int slices, int cellDiv, int outputInterval, int totalSliceCount, int mulsRepeat1,
                          int* bin_of_step) {
  if (slices < 1 || cellDiv < 1 || outputInterval < 0 || mulsRepeat1 < 1 || !bin_of_step) return -1;
  const int total = slices * cellDiv;
  const int tCount = outputInterval > 0 ? total / outputInterval : 0;
  const int nb = tCount + 1;
  std::vector<int> last(nb, -1);
  for (int k = 0; k < slices * mulsRepeat1; k++) {
    // runMulsSTEM records step (repeat r, slice s) under slice totalSliceCount + s (1 + r)
    const int r = k / slices, s = k % slices;
    const long long slice = (long long)totalSliceCount + (long long)s * (1 + r);
    int t;
    if (outputInterval == 0) t = 0;
    else if (slice < total - 1) t = (int)(slice / outputInterval);
    else t = tCount;
    bin_of_step[k] = -1;
    if (t >= 0 && t < nb) last[t] = k;
  }
  for (int t = 0; t < nb; t++)
    if (last[t] >= 0) bin_of_step[last[t]] = t;
  return nb;
}

int qs_record_bins(int slices, int cellDiv, int outputInterval, int totalSliceCount, int* bin_of_slice) {
  return qs_record_bins_repeat(slices, cellDiv, outputInterval, totalSliceCount, 1, bin_of_slice);
}

int qs_set_batch(qs_ctx* c, int batch) {
  if (!c || batch < 0) return fail(c, QS_ERR_ARG, "qs_set_batch: bad arguments");
  c->batch = batch;
  return QS_OK;
}

int qs_set_streams(qs_ctx* c, int nstreams) {
  if (!c || (nstreams != 1 && nstreams != 2)) return fail(c, QS_ERR_ARG, "qs_set_streams: nstreams must be 1 or 2");
  if (c->nstreams != nstreams) drop_graph(c);
  c->nstreams = nstreams;
  return QS_OK;
}

int qs_get_streams(const qs_ctx* c) { return c ? c->nstreams : QS_ERR_ARG; }

int qs_set_graph(qs_ctx* c, int on) {
  if (!c) return QS_ERR_ARG;
  c->use_graph = on != 0;
  if (!c->use_graph) drop_graph(c);
  return QS_OK;
}

int qs_set_configs(qs_ctx* c, int nconfig) {
  if (!c || nconfig < 1) return fail(c, QS_ERR_ARG, "qs_set_configs: bad arguments");
  if (c->trans_src) return fail(c, QS_ERR_STATE, "qs_set_configs: this context uses another's transmission functions");
  invalidate(c);
  c->nconfig = nconfig;
  c->cur_config = 0;
  c->trans_slices = 0;  // every configuration is (re)loaded, configuration 0 first
  c->v_valid = false;
  c->built.valid = false;
  return QS_OK;
}

int qs_select_config(qs_ctx* c, int config) {
  if (!c || config < 0 || config >= c->nconfig) return fail(c, QS_ERR_ARG, "qs_select_config: index out of range");
  if (config > 0 && c->trans_slices < 1) return fail(c, QS_ERR_STATE, "qs_select_config: load configuration 0 first");
  c->cur_config = config;
  return QS_OK;
}

int qs_share_trans(qs_ctx* dst, qs_ctx* src) {
  if (!dst || dst == src) return fail(dst, QS_ERR_ARG, "qs_share_trans: bad arguments");
  invalidate(dst);
  if (!src) {
    dst->trans_src = nullptr;
    return QS_OK;
  }
  if (src->trans_src) return fail(dst, QS_ERR_ARG, "qs_share_trans: the source shares another context's functions");
  const qs_params &a = dst->p, &b = src->p;
  if (a.potNx != b.potNx || a.potNy != b.potNy || a.resolutionX != b.resolutionX || a.resolutionY != b.resolutionY ||
      a.v0 != b.v0 || a.bandlimittrans != b.bandlimittrans || dst->k2max != src->k2max)
    return fail(dst, QS_ERR_ARG,
                "qs_share_trans: potential grid, sampling, voltage and band limit must match (window sizes may differ)");
  dst->trans_src = src;
  dst->v_valid = false;
  return QS_OK;
}

int qs_set_tuning(qs_ctx* c, int rows_per_wave, int tiles_per_block) {
  if (!c || rows_per_wave < 0 || tiles_per_block < 0 || (rows_per_wave & (rows_per_wave - 1)) != 0)
    return fail(c, QS_ERR_ARG, "qs_set_tuning: rows_per_wave must be 0 or a power of two, tiles_per_block >= 0");
  if (c->mixed) return set_mixed_shape(c, rows_per_wave, tiles_per_block, c->gvx, c->gvy);
  c->rows_per_wave = rows_per_wave;
  c->tiles_per_block = tiles_per_block;
  return QS_OK;
}

int qs_get_tuning(const qs_ctx* c, int* rows_per_wave, int* tiles_per_block, int* batch) {
  if (!c || !rows_per_wave || !tiles_per_block || !batch) return QS_ERR_ARG;
  *rows_per_wave = c->mixed ? c->grpb : c->rows_per_wave;
  *tiles_per_block = c->mixed ? c->gtc : c->tiles_per_block;
  *batch = c->batch;
  return QS_OK;
}

int qs_get_batch(const qs_ctx* c) { return c ? c->batch : -1; }

float qs_last_run_ms(const qs_ctx* c) { return c ? c->last_ms : 0.f; }

int qs_set_profiling(qs_ctx* c, int on) {
  if (!c) return QS_ERR_ARG;
  c->profiling = on != 0;
  c->ms_rows = c->ms_cols = 0;
  c->n_rows = c->n_cols = 0;
  return QS_OK;
}

int qs_kernel_times(const qs_ctx* c, double* ms_rows, double* ms_cols, long* n_rows, long* n_cols) {
  if (!c || !ms_rows || !ms_cols || !n_rows || !n_cols) return QS_ERR_ARG;
  *ms_rows = c->ms_rows;
  *ms_cols = c->ms_cols;
  *n_rows = c->n_rows;
  *n_cols = c->n_cols;
  return QS_OK;
}

// ---------------------------------------------------------------------------------
// Multislice runs.  A RunPlan holds one context's validated call: device buffers,
// kernel arguments, and the probe-group size; enqueue_slice issues the launches of
// one probe group at one slice.  qs_run_stem advances each probe group through all
// slices; qs_run_groups interleaves the groups of several contexts slice by slice.
struct RunPlan {
  qs_ctx* c = nullptr;
  int nprobe = 0, start = 0, nslice = 0, nbins = 0, B = 0;
  int slice0 = 0;  // first transmission slice of the call
  const int* bins = nullptr;
  bool want_det = false, has_cfg = false;
  float *det = nullptr, *diff = nullptr, *exit_wave = nullptr, *intensity = nullptr;
  // device buffers of the outputs: the context's own, or carved from the static
  // scratch array by qs_run_stem_noalloc
  float *det_d = nullptr, *diff_d = nullptr, *total_d = nullptr;
  float2* exit_d = nullptr;
  size_t slice_elems = 0;
  SweepArgs a{};
};

// Argument checks shared by the run entry points.
static int check_run(qs_ctx* c, const char* who, const int* ipos, const int* config, int nprobe, int start,
                     const int* bin_of_slice, int nslice, int nbins, const float* det, int slice0 = 0) {
  const qs_params& p = c->p;
  const std::string w(who);
  if (!ipos || nprobe < 1 || nslice < 1 || (start != 0 && start != 1)) return fail(c, QS_ERR_ARG, w + ": bad arguments");
  if (slice0 < 0) return fail(c, QS_ERR_ARG, w + ": negative first slice");
  if ((long long)slice0 + nslice > trans_slices_of(c))
    return fail(c, QS_ERR_STATE, w + ": fewer transmission slices than first slice + nslice");
  if (start == 0 && !c->have_probe) return fail(c, QS_ERR_STATE, w + ": qs_make_probe not called");
  if (start == 1 && (!c->state_valid || c->state_probes != nprobe))
    return fail(c, QS_ERR_STATE, w + ": no wave state to continue for this probe count");
  for (int i = 0; i < nprobe; i++) {
    const int x = ipos[2 * i], y = ipos[2 * i + 1];
    if (x < 0 || y < 0 || x + p.nx > p.potNx || y + p.ny > p.potNy)
      return fail(c, QS_ERR_ARG, w + ": window offset outside the transmission array");
    if (config && (config[i] < 0 || config[i] >= nconfig_of(c)))
      return fail(c, QS_ERR_ARG, w + ": configuration index out of range");
  }
  if (det && p.ndet > 0) {
    if (!bin_of_slice || nbins < 1) return fail(c, QS_ERR_ARG, w + ": det needs bin_of_slice and nbins");
    for (int s = 0; s < nslice; s++)
      if (bin_of_slice[s] >= nbins) return fail(c, QS_ERR_ARG, w + ": bin index >= nbins");
  }
  return QS_OK;
}

// Probe-group size: the library default keeps the wave state of the groups in
// flight (two sweeps per slice read and write it) inside the 256 MB Infinity
// Cache: one group of ~180 MB, or with two streams two groups of ~96 MB
// (measured best for 512x512 waves: 128 probes, 2 x 64 probes).
static int group_size(const qs_ctx* c, int nprobe, int nstreams) {
  const size_t per_probe = sizeof(float2) * (size_t)c->p.nx * c->kcp;
  const size_t budget = (size_t)(nstreams > 1 ? 96u : 180u) << 20;
  int Bdef = (int)std::max<size_t>(32, std::min<size_t>(1024, budget / per_probe));
  Bdef = std::max(32, (Bdef / 16) * 16);
  if (nstreams > 1 && nprobe > 1) Bdef = std::min(Bdef, (nprobe + 1) / 2);  // both streams get work
  return c->batch > 0 ? std::min(c->batch, nprobe) : std::min(Bdef, nprobe);
}

// Kernel arguments of a run whose device buffers (state, ipos, partial, outputs) are set in r.
static void fill_sweep_args(qs_ctx* c, RunPlan& r, const int* d_ipos, const int* d_cfg, float2* d_state,
                            float* d_partial) {
  const qs_params& p = c->p;
  SweepArgs& a = r.a;
  a.nx = p.nx;
  a.ny = p.ny;
  a.kc = c->kc;
  a.kcp = c->kcp;
  a.band_a = c->band_a;
  a.band_b = c->band_b;
  a.ntiles = c->ntiles;
  a.potNy = p.potNy;
  a.rw = c->rows_per_wave;
  a.tpb = c->tiles_per_block;
  a.probe = c->probe.as<float2>();
  a.ipos = reinterpret_cast<const int2*>(d_ipos);
  a.cfg = d_cfg;
  a.cfg_stride = (long long)trans_slices_of(c) * (long long)r.slice_elems;
  a.state = d_state;
  a.out_wave = r.exit_d;
  a.twx = c->twx.as<float2>();
  a.twy = c->twy.as<float2>();
  a.px = c->px.as<float2>();
  a.py = c->py.as<float2>();
  a.kx2 = c->dkx2.as<float>();
  a.ky2 = c->dky2.as<float>();
  a.k2max = c->k2max;
  a.ndet = p.ndet;
  for (int d = 0; d < p.ndet; d++) {
    a.det_sx[d] = (int)p.det_shiftx[d];
    a.det_sy[d] = (int)p.det_shifty[d];
    a.det_k2in[d] = c->det_k2in[d];
    a.det_k2out[d] = c->det_k2out[d];
  }
  a.partial = d_partial;
  a.diff_scale = (float)pow((double)p.nx * p.ny, 1.5);
}

static int plan_run(qs_ctx* c, const char* who, const int* ipos, const int* config, int nprobe, int start,
                    const int* bin_of_slice, int nslice, int nbins, float* det, float* diff, float* exit_wave,
                    float* int_intensity, RunPlan& r, int slice0 = 0) {
  const qs_params& p = c->p;
  const int rc = check_run(c, who, ipos, config, nprobe, start, bin_of_slice, nslice, nbins, det, slice0);
  if (rc != QS_OK) return rc;
  const int nx = p.nx, ny = p.ny;
  r = RunPlan{};
  r.c = c;
  r.nprobe = nprobe;
  r.start = start;
  r.slice0 = slice0;
  r.nslice = nslice;
  r.nbins = nbins;
  r.bins = bin_of_slice;
  r.B = group_size(c, nprobe, c->nstreams);
  r.want_det = det && p.ndet > 0;
  r.det = det;
  r.diff = diff;
  r.exit_wave = exit_wave;
  r.intensity = int_intensity;
  r.has_cfg = config != nullptr;
  r.slice_elems = (size_t)p.potNx * p.potNy;
  const size_t state_elems = (size_t)nprobe * nx * c->kcp;
  HIPCHK(c, c->state.ensure(sizeof(float2) * state_elems));
  if (start == 0) qs::launch_fill_zero(c->state.as<float2>(), state_elems, c->stream);  // padding columns stay zero
  HIPCHK(c, c->ipos.ensure(sizeof(int) * 2 * nprobe));
  HIPCHK(c, hipMemcpyAsync(c->ipos.p, ipos, sizeof(int) * 2 * nprobe, hipMemcpyHostToDevice, c->stream));
  if (config) {
    HIPCHK(c, c->cfgidx.ensure(sizeof(int) * nprobe));
    HIPCHK(c, hipMemcpyAsync(c->cfgidx.p, config, sizeof(int) * nprobe, hipMemcpyHostToDevice, c->stream));
  }
  HIPCHK(c, c->partial.ensure(sizeof(float) * (size_t)r.B * c->ntiles * (qs::kMaxDet + 1)));
  if (c->nstreams > 1 && r.nprobe > r.B)
    HIPCHK(c, c->partial2.ensure(sizeof(float) * (size_t)r.B * c->ntiles * (qs::kMaxDet + 1)));
  if (r.want_det) {
    const size_t nd = (size_t)nbins * p.ndet * nprobe;
    HIPCHK(c, c->det.ensure(sizeof(float) * nd));
    HIPCHK(c, hipMemcpyAsync(c->det.p, det, sizeof(float) * nd, hipMemcpyHostToDevice, c->stream));
  }
  if (diff) {
    HIPCHK(c, c->diff.ensure(sizeof(float) * (size_t)nprobe * nx * ny));
    HIPCHK(c, hipMemsetAsync(c->diff.p, 0, sizeof(float) * (size_t)nprobe * nx * ny, c->stream));
  }
  if (exit_wave) HIPCHK(c, c->exitw.ensure(sizeof(float2) * (size_t)nprobe * nx * ny));
  if (int_intensity) HIPCHK(c, c->total.ensure(sizeof(float) * nprobe));
  r.det_d = c->det.as<float>();
  r.diff_d = c->diff.as<float>();
  r.total_d = c->total.as<float>();
  r.exit_d = c->exitw.as<float2>();

  fill_sweep_args(c, r, c->ipos.as<int>(), config ? c->cfgidx.as<int>() : nullptr, c->state.as<float2>(),
                  c->partial.as<float>());
  return QS_OK;
}

// Event pairs around every sweep launch (qs_set_profiling), accumulated after the run.
struct LaunchMarks {
  qs_ctx* c;
  std::vector<std::pair<int, int>> marks;  // (kind 0=rows 1=cols, first event index)
  size_t nev = 0;
  hipError_t mark(int kind, bool begin, hipStream_t st) {
    if (!c->profiling) return hipSuccess;
    if (nev >= c->evpool.size()) {
      hipEvent_t e;
      hipError_t r = hipEventCreate(&e);
      if (r != hipSuccess) return r;
      c->evpool.push_back(e);
    }
    if (begin) marks.push_back({kind, (int)nev});
    return hipEventRecord(c->evpool[nev++], st);
  }
};

// Launches of probe group [p0, p0 + B) at slice s on stream st, with detector partial
// sums in `partial`; the exit-wave sweep follows the last slice.
static hipError_t enqueue_slice(RunPlan& r, int p0, int s, hipStream_t st, float* partial, LaunchMarks* lm) {
  qs_ctx* c = r.c;
  const qs_params& p = c->p;
  SweepArgs& a = r.a;
  const int nb = std::min(r.B, r.nprobe - p0);
  a.nprobe = nb;
  a.probe0 = p0;
  a.trans = trans_base(c) + (size_t)(r.slice0 + s) * r.slice_elems;
  const bool last = (s == r.nslice - 1);
  hipError_t e = lm ? lm->mark(0, true, st) : hipSuccess;
  if (e != hipSuccess) return e;
  qs::launch_sweep_rows(a, (s == 0 && r.start == 0) ? 0 : 1, st);
  if (lm && (e = lm->mark(0, false, st)) != hipSuccess) return e;
  const int bin = (r.want_det && r.bins) ? r.bins[s] : -1;
  const bool tot = last && r.intensity;
  a.rec_bin = (bin >= 0 || tot) ? (bin >= 0 ? bin : 0) : -1;
  a.diff = (last && r.diff) ? r.diff_d : nullptr;
  a.partial = partial;
  if (lm && (e = lm->mark(1, true, st)) != hipSuccess) return e;
  qs::launch_sweep_cols(a, a.rec_bin >= 0, a.diff != nullptr, st);
  if (lm && (e = lm->mark(1, false, st)) != hipSuccess) return e;
  if (bin >= 0 || tot) {
    float* dbin = bin >= 0 ? r.det_d + (size_t)bin * p.ndet * r.nprobe : nullptr;
    qs::launch_reduce_partials(partial, nb, c->ntiles, bin >= 0 ? p.ndet : 0, p.electronScale, dbin,
                               r.nprobe, tot ? r.total_d : nullptr, p0, st);
  }
  if (last && r.exit_wave) qs::launch_sweep_rows(a, 2, st);
  return hipSuccess;
}

// Device-to-host copies of the requested outputs on stream st (then synchronised).
static int finish_run(RunPlan& r, hipStream_t st) {
  qs_ctx* c = r.c;
  const qs_params& p = c->p;
  const size_t n2 = (size_t)r.nprobe * p.nx * p.ny;
  if (r.want_det)
    HIPCHK(c, hipMemcpyAsync(r.det, r.det_d, sizeof(float) * (size_t)r.nbins * p.ndet * r.nprobe,
                             hipMemcpyDeviceToHost, st));
  if (r.diff) HIPCHK(c, hipMemcpyAsync(r.diff, r.diff_d, sizeof(float) * n2, hipMemcpyDeviceToHost, st));
  if (r.exit_wave) HIPCHK(c, hipMemcpyAsync(r.exit_wave, r.exit_d, sizeof(float2) * n2, hipMemcpyDeviceToHost, st));
  if (r.intensity)
    HIPCHK(c, hipMemcpyAsync(r.intensity, r.total_d, sizeof(float) * r.nprobe, hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipStreamSynchronize(st));
  c->state_valid = true;
  c->state_probes = r.nprobe;
  return QS_OK;
}

// qs_run_stem for wave grids whose sizes factor into 2, 3, 5 and 7 (not both powers
// of two): per probe group and slice, k_gen_rows (y-IFFT of slice s-1, transmit,
// y-FFT) and k_gen_cols (x-FFT, propagator, detector sums, x-IFFT) over the
// y-spectrum state [probe][x][ky] (in-band columns only); the exit wave by a final
// y-IFFT (qs_generic.hip).
static int run_mixed(qs_ctx* c, const int* ipos, const int* config, int nprobe, int start, int slice0,
                     const int* bins, int nslice, int nbins, float* det, float* diff, float* exit_wave,
                     float* int_intensity) {
  int rc = check_run(c, "qs_run_stem", ipos, config, nprobe, start, bins, nslice, nbins, det, slice0);
  if (rc != QS_OK) return rc;
  const qs_params& p = c->p;
  const int nx = p.nx, ny = p.ny;
  const size_t n2 = (size_t)nx * ny, slice = (size_t)p.potNx * p.potNy;
  // probe group: its state at most 160 MB, resident in the Infinity Cache across the two kernels
  const int B = c->batch > 0 ? std::min(c->batch, nprobe)
                             : (int)std::max<size_t>(1, std::min<size_t>(nprobe, (size_t(160) << 20) / (8 * n2)));
  const bool want_det = det && p.ndet > 0;
  hipStream_t st = c->stream;
  HIPCHK(c, c->state.ensure(sizeof(float2) * n2 * nprobe));
  HIPCHK(c, c->ipos.ensure(sizeof(int) * 2 * nprobe));
  HIPCHK(c, hipMemcpyAsync(c->ipos.p, ipos, sizeof(int) * 2 * nprobe, hipMemcpyHostToDevice, st));
  if (config) {
    HIPCHK(c, c->cfgidx.ensure(sizeof(int) * nprobe));
    HIPCHK(c, hipMemcpyAsync(c->cfgidx.p, config, sizeof(int) * nprobe, hipMemcpyHostToDevice, st));
  }
  HIPCHK(c, c->partial.ensure(sizeof(float) * (size_t)B * c->gntiles * (qs::kMaxDet + 1)));
  const size_t nd = want_det ? (size_t)nbins * p.ndet * nprobe : 0;
  if (want_det) {
    HIPCHK(c, c->det.ensure(sizeof(float) * nd));
    HIPCHK(c, hipMemcpyAsync(c->det.p, det, sizeof(float) * nd, hipMemcpyHostToDevice, st));
  }
  if (diff) {  // columns outside the band are never visited: their pattern is zero
    HIPCHK(c, c->diff.ensure(sizeof(float) * n2 * nprobe));
    HIPCHK(c, hipMemsetAsync(c->diff.p, 0, sizeof(float) * n2 * nprobe, st));
  }
  if (exit_wave) HIPCHK(c, c->exitw.ensure(sizeof(float2) * n2 * nprobe));
  if (int_intensity) HIPCHK(c, c->total.ensure(sizeof(float) * nprobe));
  HIPCHK(c, hipEventRecord(c->ev0, st));
  qs::GenRowArgs ra{};
  ra.nx = nx;
  ra.ny = ny;
  ra.potNy = p.potNy;
  ra.rpb = c->grpb;
  ra.probe = c->probe.as<float2>();
  ra.ipos = c->ipos.as<int2>();
  ra.cfg = config ? c->cfgidx.as<int>() : nullptr;
  ra.cfg_stride = (long long)trans_slices_of(c) * (long long)slice;
  ra.state = c->state.as<float2>();
  ra.exitw = c->exitw.as<float2>();
  ra.ky2 = c->dky2.as<float>();
  ra.k2max = c->k2max;
  ra.plan = c->gpy;
  ra.tw = c->twy.as<float2>();
  qs::GenColArgs ca{};
  ca.nx = nx;
  ca.ny = ny;
  ca.tc = c->gvx == qs::kGenPlanCompiled ? c->gkc : c->gtc;
  ca.state = c->state.as<float2>();
  ca.tiles = c->gtiles.as<int>();
  ca.px = c->px.as<float2>();
  ca.py = c->py.as<float2>();
  ca.kx2 = c->dkx2.as<float>();
  ca.ky2 = c->dky2.as<float>();
  ca.k2max = c->k2max;
  ca.ndet = p.ndet;
  for (int d = 0; d < p.ndet; d++) {  // shifts reduced to [0, nx) and [0, ny)
    ca.det_sx[d] = ((int)p.det_shiftx[d] % nx + nx) % nx;
    ca.det_sy[d] = ((int)p.det_shifty[d] % ny + ny) % ny;
    ca.det_k2in[d] = c->det_k2in[d];
    ca.det_k2out[d] = c->det_k2out[d];
  }
  ca.partial = c->partial.as<float>();
  ca.diff_scale = (float)pow((double)n2, 1.5);
  ca.plan = c->gpx;
  ca.tw = c->twx.as<float2>();
  for (int p0 = 0; p0 < nprobe; p0 += B) {
    const int nb = std::min(B, nprobe - p0);
    ra.nprobe = ca.nprobe = nb;
    ra.probe0 = ca.probe0 = p0;
    for (int s = 0; s < nslice; s++) {
      ra.mode = (s == 0 && start == 0) ? 0 : 1;
      ra.trans = trans_base(c) + (size_t)(slice0 + s) * slice;
      if (c->gvy == qs::kGenPlanCompiled) qs::launch_mix_rows(ra, st);
      else qs::launch_gen_rows(ra, st);
      const int bin = want_det && bins ? bins[s] : -1;
      const bool last = s == nslice - 1, tot = last && int_intensity;
      ca.rec = (bin >= 0 || tot) ? 1 : 0;
      ca.diff = (last && diff) ? c->diff.as<float>() : nullptr;
      if (c->gvx == qs::kGenPlanCompiled) qs::launch_mix_cols(ca, c->gntiles, st);
      else qs::launch_gen_cols(ca, c->gntiles, st);
      if (ca.rec)
        qs::launch_reduce_partials(ca.partial, nb, c->gntiles, bin >= 0 ? p.ndet : 0, p.electronScale,
                                   bin >= 0 ? c->det.as<float>() + (size_t)bin * p.ndet * nprobe : nullptr, nprobe,
                                   tot ? c->total.as<float>() : nullptr, p0, st);
    }
    if (exit_wave) {  // the real-space exit waves: y-IFFT of the state, no transmit
      ra.mode = 2;
      if (c->gvy == qs::kGenPlanCompiled) qs::launch_mix_rows(ra, st);
      else qs::launch_gen_rows(ra, st);
    }
  }
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipEventRecord(c->ev1, st));
  if (want_det) HIPCHK(c, hipMemcpyAsync(det, c->det.p, sizeof(float) * nd, hipMemcpyDeviceToHost, st));
  if (diff) HIPCHK(c, hipMemcpyAsync(diff, c->diff.p, sizeof(float) * n2 * nprobe, hipMemcpyDeviceToHost, st));
  if (exit_wave)
    HIPCHK(c, hipMemcpyAsync(exit_wave, c->exitw.p, sizeof(float2) * n2 * nprobe, hipMemcpyDeviceToHost, st));
  if (int_intensity)
    HIPCHK(c, hipMemcpyAsync(int_intensity, c->total.p, sizeof(float) * nprobe, hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipStreamSynchronize(st));
  HIPCHK(c, hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
  c->state_valid = true;
  c->state_probes = nprobe;
  return QS_OK;
}

// qs_run_stem for other wave grids that are not powers of two (a prime factor above
// 7): per slice and probe group, transmit, forward 2D FFT (rows, transpose, rows),
// propagator and detector sums, inverse 2D FFT, on one stream over a full-size state
// (qs_generic.hip).
static int run_generic(qs_ctx* c, const int* ipos, const int* config, int nprobe, int start, int slice0,
                       const int* bins, int nslice, int nbins, float* det, float* diff, float* exit_wave,
                       float* int_intensity) {
  if (c->mixed)
    return run_mixed(c, ipos, config, nprobe, start, slice0, bins, nslice, nbins, det, diff, exit_wave,
                     int_intensity);
  const char* who = "qs_run_stem";
  int rc = check_run(c, who, ipos, config, nprobe, start, bins, nslice, nbins, det, slice0);
  if (rc != QS_OK) return rc;
  const qs_params& p = c->p;
  const int nx = p.nx, ny = p.ny;
  const size_t n2 = (size_t)nx * ny, slice = (size_t)p.potNx * p.potNy;
  // probe group: its transposed copy at most 256 MB
  const int B = c->batch > 0 ? std::min(c->batch, nprobe)
                             : (int)std::max<size_t>(1, std::min<size_t>(nprobe, (size_t(256) << 20) / (8 * n2)));
  const int nchunk = qs::gen_chunks(ny);
  const bool want_det = det && p.ndet > 0;
  hipStream_t st = c->stream;
  HIPCHK(c, c->state.ensure(sizeof(float2) * n2 * nprobe));
  HIPCHK(c, c->gscratch.ensure(sizeof(float2) * n2 * B));
  HIPCHK(c, c->ipos.ensure(sizeof(int) * 2 * nprobe));
  HIPCHK(c, hipMemcpyAsync(c->ipos.p, ipos, sizeof(int) * 2 * nprobe, hipMemcpyHostToDevice, st));
  if (config) {
    HIPCHK(c, c->cfgidx.ensure(sizeof(int) * nprobe));
    HIPCHK(c, hipMemcpyAsync(c->cfgidx.p, config, sizeof(int) * nprobe, hipMemcpyHostToDevice, st));
  }
  HIPCHK(c, c->partial.ensure(sizeof(float) * (size_t)B * nchunk * (qs::kMaxDet + 1)));
  const size_t nd = want_det ? (size_t)nbins * p.ndet * nprobe : 0;
  if (want_det) {
    HIPCHK(c, c->det.ensure(sizeof(float) * nd));
    HIPCHK(c, hipMemcpyAsync(c->det.p, det, sizeof(float) * nd, hipMemcpyHostToDevice, st));
  }
  if (diff) HIPCHK(c, c->diff.ensure(sizeof(float) * n2 * nprobe));
  if (int_intensity) HIPCHK(c, c->total.ensure(sizeof(float) * nprobe));
  HIPCHK(c, hipEventRecord(c->ev0, st));
  qs::GenPropArgs pa{};
  pa.nx = nx;
  pa.ny = ny;
  pa.px = c->px.as<float2>();
  pa.py = c->py.as<float2>();
  pa.kx2 = c->dkx2.as<float>();
  pa.ky2 = c->dky2.as<float>();
  pa.k2max = c->k2max;
  pa.ndet = p.ndet;
  for (int d = 0; d < p.ndet; d++) {
    pa.det_sx[d] = (int)p.det_shiftx[d];
    pa.det_sy[d] = (int)p.det_shifty[d];
    pa.det_k2in[d] = c->det_k2in[d];
    pa.det_k2out[d] = c->det_k2out[d];
  }
  pa.partial = c->partial.as<float>();
  pa.diff_scale = (float)pow((double)n2, 1.5);
  float2* sc = c->gscratch.as<float2>();
  for (int p0 = 0; p0 < nprobe; p0 += B) {
    const int nb = std::min(B, nprobe - p0);
    float2* stt = c->state.as<float2>() + (size_t)p0 * n2;
    for (int s = 0; s < nslice; s++) {
      qs::GenArgs g{};
      g.nx = nx;
      g.ny = ny;
      g.nprobe = nb;
      g.probe0 = p0;
      g.potNy = p.potNy;
      g.probe = c->probe.as<float2>();
      g.trans = trans_base(c) + (size_t)(slice0 + s) * slice;
      g.ipos = c->ipos.as<int2>();
      g.cfg = config ? c->cfgidx.as<int>() : nullptr;
      g.cfg_stride = (long long)trans_slices_of(c) * (long long)slice;
      g.state = stt;
      qs::launch_gen_transmit(g, s == 0 && start == 0, st);
      run_rowxform(c, c->wy, stt, (long)nb * nx, -1, 1.0f);
      qs::launch_transpose(stt, sc, nb, nx, ny, st);
      run_rowxform(c, c->wx, sc, (long)nb * ny, -1, 1.0f);
      const int bin = want_det && bins ? bins[s] : -1;
      const bool last = s == nslice - 1, tot = last && int_intensity;
      pa.data = sc;
      pa.probe0 = p0;
      pa.rec = (bin >= 0 || tot) ? 1 : 0;
      pa.diff = (last && diff) ? c->diff.as<float>() : nullptr;
      qs::launch_gen_propagate(pa, nb, st);
      if (pa.rec)
        qs::launch_reduce_partials(pa.partial, nb, nchunk, bin >= 0 ? p.ndet : 0, p.electronScale,
                                   bin >= 0 ? c->det.as<float>() + (size_t)bin * p.ndet * nprobe : nullptr, nprobe,
                                   tot ? c->total.as<float>() : nullptr, p0, st);
      run_rowxform(c, c->wx, sc, (long)nb * ny, +1, 1.0f);
      qs::launch_transpose(sc, stt, nb, ny, nx, st);
      run_rowxform(c, c->wy, stt, (long)nb * nx, +1, 1.0f);
    }
  }
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipEventRecord(c->ev1, st));
  if (want_det) HIPCHK(c, hipMemcpyAsync(det, c->det.p, sizeof(float) * nd, hipMemcpyDeviceToHost, st));
  if (diff) HIPCHK(c, hipMemcpyAsync(diff, c->diff.p, sizeof(float) * n2 * nprobe, hipMemcpyDeviceToHost, st));
  if (exit_wave)
    HIPCHK(c, hipMemcpyAsync(exit_wave, c->state.p, sizeof(float2) * n2 * nprobe, hipMemcpyDeviceToHost, st));
  if (int_intensity)
    HIPCHK(c, hipMemcpyAsync(int_intensity, c->total.p, sizeof(float) * nprobe, hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipStreamSynchronize(st));
  HIPCHK(c, hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
  c->state_valid = true;
  c->state_probes = nprobe;
  return QS_OK;
}

int qs_run_slices(qs_ctx* c, const int* ipos, const int* config, int nprobe, int start, int first_slice,
                  const int* bin_of_slice, int nslice, int nbins, float* det, float* diff, float* exit_wave,
                  float* int_intensity) {
  if (!c) return QS_ERR_ARG;
  if (c->generic)
    return run_generic(c, ipos, config, nprobe, start, first_slice, bin_of_slice, nslice, nbins, det, diff,
                       exit_wave, int_intensity);
  RunPlan r;
  int rc = plan_run(c, "qs_run_stem", ipos, config, nprobe, start, bin_of_slice, nslice, nbins, det, diff, exit_wave,
                    int_intensity, r, first_slice);
  if (rc != QS_OK) return rc;
  LaunchMarks lm{c};
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  // probe group g runs on stream g % 2 (one stream for a single group or nstreams 1);
  // the groups of one stream follow each other, and the launches of a pair of
  // groups are issued slice by slice so that both streams stay fed
  const int ngroup = (r.nprobe + r.B - 1) / r.B;
  const int nst = (c->nstreams > 1 && ngroup > 1) ? 2 : 1;
  auto fork = [&]() -> hipError_t {  // the second stream starts after the uploads on the first
    const hipError_t e = hipEventRecord(c->evfork, c->stream);
    return e != hipSuccess ? e : hipStreamWaitEvent(c->stream2, c->evfork, 0);
  };
  auto join = [&]() -> hipError_t {  // the first stream waits for the second
    const hipError_t e = hipEventRecord(c->evjoin, c->stream2);
    return e != hipSuccess ? e : hipStreamWaitEvent(c->stream, c->evjoin, 0);
  };
  // launches of stream `only` (or of both streams, only = -1)
  auto enqueue = [&](int only) -> hipError_t {
    hipError_t e = hipSuccess;
    for (int g = 0; g < ngroup; g += nst)
      for (int s = 0; s < r.nslice; s++)
        for (int k = 0; k < nst && g + k < ngroup; k++) {
          if (only >= 0 && k != only) continue;
          hipStream_t st = k ? c->stream2 : c->stream;
          float* part = (k ? c->partial2 : c->partial).as<float>();
          if ((e = enqueue_slice(r, (g + k) * r.B, s, st, part, &lm)) != hipSuccess) return e;
        }
    return hipGetLastError();
  };
  if (c->use_graph && !c->profiling) {
    // the graph bakes in every kernel argument: key on the call shape and the buffers
    std::vector<long long> key = {nprobe, r.B, nslice, start, r.slice0, nbins, r.want_det, diff != nullptr,
                                  exit_wave != nullptr, int_intensity != nullptr, r.has_cfg,
                                  (long long)trans_base(c), (long long)c->state.p, (long long)c->det.p,
                                  (long long)c->diff.p, (long long)c->exitw.p, (long long)c->total.p,
                                  (long long)c->partial.p, (long long)c->partial2.p, nst, (long long)c->ipos.p,
                                  (long long)c->cfgidx.p, (long long)c->probe.p, c->rows_per_wave,
                                  c->tiles_per_block};
    for (int s = 0; s < nslice; s++) key.push_back(bin_of_slice ? bin_of_slice[s] : -1);
    if (!c->gexec || key != c->gkey) {
      // one graph per stream, each the linear chain of that stream's launches (measured
      // faster than one graph with two parallel branches)
      drop_graph(c);
      for (int k = 0; k < nst; k++) {
        hipStream_t st = k ? c->stream2 : c->stream;
        hipGraphExec_t& ge = k ? c->gexec2 : c->gexec;
        HIPCHK(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const hipError_t le = enqueue(k);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(st, &g);
        if (le != hipSuccess || ce != hipSuccess) {
          if (g) (void)hipGraphDestroy(g);
          drop_graph(c);
          return fail(c, QS_ERR_HIP, std::string("qs_run_stem: graph capture failed: ") +
                                         hipGetErrorString(le != hipSuccess ? le : ce));
        }
        const hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) {
          ge = nullptr;
          drop_graph(c);
          return fail(c, QS_ERR_HIP, std::string("qs_run_stem: hipGraphInstantiate: ") + hipGetErrorString(ie));
        }
      }
      c->gkey = key;
    }
    if (nst == 2) HIPCHK(c, fork());
    HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
    if (nst == 2) {
      HIPCHK(c, hipGraphLaunch(c->gexec2, c->stream2));
      HIPCHK(c, join());
    }
  } else {
    if (nst == 2) HIPCHK(c, fork());
    HIPCHK(c, enqueue(-1));
    if (nst == 2) HIPCHK(c, join());
  }
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  rc = finish_run(r, c->stream);
  if (rc != QS_OK) return rc;
  HIPCHK(c, hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
  for (auto& m : lm.marks) {
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->evpool[m.second], c->evpool[m.second + 1]));
    if (m.first == 0) { c->ms_rows += ms; c->n_rows++; }
    else { c->ms_cols += ms; c->n_cols++; }
  }
  return QS_OK;
}

int qs_run_stem_cfg(qs_ctx* c, const int* ipos, const int* config, int nprobe, int start, const int* bin_of_slice,
                    int nslice, int nbins, float* det, float* diff, float* exit_wave, float* int_intensity) {
  return qs_run_slices(c, ipos, config, nprobe, start, 0, bin_of_slice, nslice, nbins, det, diff, exit_wave,
                       int_intensity);
}

int qs_run_stem(qs_ctx* c, const int* ipos, int nprobe, int start, const int* bin_of_slice, int nslice, int nbins,
                float* det, float* diff, float* exit_wave, float* int_intensity) {
  return qs_run_stem_cfg(c, ipos, nullptr, nprobe, start, bin_of_slice, nslice, nbins, det, diff, exit_wave,
                         int_intensity);
}

// qs_run_stem from the probe, with every device buffer of the call in the static scratch array.
int qs_run_stem_noalloc(qs_ctx* c, const int* ipos, int nprobe, const int* bin_of_slice, int nslice, int nbins,
                        float* det, float* int_intensity) {
  if (!c) return QS_ERR_ARG;
  const char* who = "qs_run_stem_noalloc";
  if (c->generic)
    return fail(c, QS_ERR_UNSUPPORTED, std::string(who) + ": wave grids that are not powers of two use qs_run_stem");
  int rc = check_run(c, who, ipos, nullptr, nprobe, 0, bin_of_slice, nslice, nbins, det);
  if (rc != QS_OK) return rc;
  const qs_params& p = c->p;
  std::string why;
  unsigned char* base = scratch_of(c->device, why);
  if (!base) return fail(c, QS_ERR_HIP, std::string(who) + ": " + why);
  RunPlan r;
  r.c = c;
  r.nprobe = nprobe;
  r.start = 0;
  r.nslice = nslice;
  r.nbins = nbins;
  r.bins = bin_of_slice;
  r.B = group_size(c, nprobe, 1);
  r.want_det = det && p.ndet > 0;
  r.det = det;
  r.intensity = int_intensity;
  r.slice_elems = (size_t)p.potNx * p.potNy;
  // layout of the scratch array for this call, 256-byte aligned pieces
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  };
  const size_t state_elems = (size_t)nprobe * p.nx * c->kcp;
  const size_t o_ipos = carve(sizeof(int) * 2 * (size_t)nprobe);
  const size_t o_state = carve(sizeof(float2) * state_elems);
  const size_t o_part = carve(sizeof(float) * (size_t)r.B * c->ntiles * (qs::kMaxDet + 1));
  const size_t nd = r.want_det ? (size_t)nbins * p.ndet * nprobe : 0;
  const size_t o_det = carve(sizeof(float) * nd);
  const size_t o_tot = carve(int_intensity ? sizeof(float) * (size_t)nprobe : 0);
  if (off > (size_t)QS_NOALLOC_SCRATCH_BYTES)
    return fail(c, QS_ERR_ARG, std::string(who) + ": " + std::to_string(nprobe) + " probes need " +
                                   std::to_string(off) + " bytes of scratch, more than the " +
                                   std::to_string((long long)QS_NOALLOC_SCRATCH_BYTES) + " the library holds");
  int* d_ipos = reinterpret_cast<int*>(base + o_ipos);
  float2* d_state = reinterpret_cast<float2*>(base + o_state);
  float* d_part = reinterpret_cast<float*>(base + o_part);
  r.det_d = r.want_det ? reinterpret_cast<float*>(base + o_det) : nullptr;
  r.total_d = int_intensity ? reinterpret_cast<float*>(base + o_tot) : nullptr;
  hipStream_t st = c->stream;
  HIPCHK(c, hipEventRecord(c->ev0, st));
  // the array keeps the previous call's contents: the state's padding columns must
  // start at zero, the detector bins start from the caller's values (in/out)
  qs::launch_fill_zero(d_state, state_elems, st);
  HIPCHK(c, hipMemcpyAsync(d_ipos, ipos, sizeof(int) * 2 * (size_t)nprobe, hipMemcpyHostToDevice, st));
  if (r.want_det) HIPCHK(c, hipMemcpyAsync(r.det_d, det, sizeof(float) * nd, hipMemcpyHostToDevice, st));
  fill_sweep_args(c, r, d_ipos, nullptr, d_state, d_part);
  for (int p0 = 0; p0 < nprobe; p0 += r.B)
    for (int s = 0; s < nslice; s++) HIPCHK(c, enqueue_slice(r, p0, s, st, d_part, nullptr));
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipEventRecord(c->ev1, st));
  if (r.want_det) HIPCHK(c, hipMemcpyAsync(det, r.det_d, sizeof(float) * nd, hipMemcpyDeviceToHost, st));
  if (int_intensity)
    HIPCHK(c, hipMemcpyAsync(int_intensity, r.total_d, sizeof(float) * nprobe, hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipStreamSynchronize(st));
  HIPCHK(c, hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
  return QS_OK;
}

int qs_run_groups(const qs_group* groups, int ngroup, const int* bin_of_slice, int nslice, int nbins) {
  if (!groups || ngroup < 1 || !groups[0].ctx) return QS_ERR_ARG;
  qs_ctx* c0 = groups[0].ctx;
  const float2* t0 = trans_base(c0);
  for (int g = 0; g < ngroup; g++) {
    qs_ctx* c = groups[g].ctx;
    if (!c) return fail(c0, QS_ERR_ARG, "qs_run_groups: null context");
    if (c->generic)
      return fail(c0, QS_ERR_UNSUPPORTED, "qs_run_groups: wave grids that are not powers of two use qs_run_stem");
    if (trans_base(c) != t0 || trans_slices_of(c) != trans_slices_of(c0))
      return fail(c0, QS_ERR_ARG, "qs_run_groups: contexts must share one transmission function (qs_share_trans)");
    for (int h = 0; h < g; h++)
      if (groups[h].ctx == c) return fail(c0, QS_ERR_ARG, "qs_run_groups: a context appears twice");
  }
  std::vector<RunPlan> plans(ngroup);
  for (int g = 0; g < ngroup; g++) {
    const qs_group& G = groups[g];
    const int rc = plan_run(G.ctx, "qs_run_groups", G.ipos, G.config, G.nitem, 0, bin_of_slice, nslice, nbins, G.det,
                            nullptr, nullptr, G.int_intensity, plans[g]);
    if (rc != QS_OK) {
      if (G.ctx != c0) c0->err = G.ctx->err;
      return rc;
    }
    HIPCHK(c0, hipStreamSynchronize(G.ctx->stream));  // uploads done before the shared stream reads them
  }
  // round j advances probe group j of every context through all slices; within a
  // slice the contexts follow each other, so slice s of the shared transmission
  // function is read by every window size while it is cache-resident
  int rounds = 0;
  for (auto& r : plans) rounds = std::max(rounds, (r.nprobe + r.B - 1) / r.B);
  // with two streams each context's launches go to the stream with less work so far
  // (work ~ items x rows x band columns), so that one stream's launches fill the
  // other's ramp-up and drain
  std::vector<int> on2(ngroup, 0);
  bool two = false;
  if (c0->nstreams > 1 && ngroup > 1) {
    std::vector<int> order(ngroup);
    for (int g = 0; g < ngroup; g++) order[g] = g;
    auto work = [&](int g) { return (double)plans[g].nprobe * groups[g].ctx->p.nx * groups[g].ctx->kcp; };
    std::sort(order.begin(), order.end(), [&](int u, int v) { return work(u) > work(v); });
    double w[2] = {0.0, 0.0};
    for (int g : order) {
      on2[g] = w[1] < w[0] ? 1 : 0;
      w[on2[g]] += work(g);
    }
    two = w[1] > 0.0;
  }
  HIPCHK(c0, hipEventRecord(c0->ev0, c0->stream));
  if (two) {
    HIPCHK(c0, hipEventRecord(c0->evfork, c0->stream));
    HIPCHK(c0, hipStreamWaitEvent(c0->stream2, c0->evfork, 0));
  }
  for (int j = 0; j < rounds; j++)
    for (int s = 0; s < nslice; s++)
      for (int g = 0; g < ngroup; g++) {
        RunPlan& r = plans[g];
        if ((long)j * r.B < r.nprobe)
          HIPCHK(c0, enqueue_slice(r, j * r.B, s, on2[g] ? c0->stream2 : c0->stream, r.c->partial.as<float>(), nullptr));
      }
  HIPCHK(c0, hipGetLastError());
  if (two) {
    HIPCHK(c0, hipEventRecord(c0->evjoin, c0->stream2));
    HIPCHK(c0, hipStreamWaitEvent(c0->stream, c0->evjoin, 0));
  }
  HIPCHK(c0, hipEventRecord(c0->ev1, c0->stream));
  HIPCHK(c0, hipStreamSynchronize(c0->stream));
  for (auto& r : plans) {
    const int rc = finish_run(r, r.c->stream);
    if (rc != QS_OK) return rc;
  }
  HIPCHK(c0, hipEventElapsedTime(&c0->last_ms, c0->ev0, c0->ev1));
  return QS_OK;
}


// ---------------------------------------------------------------------------------
// Adjoint multislice with slice checkpointing.
static void sweep_args(qs_ctx* c, SweepArgs& a, int nprobe) {
  const qs_params& p = c->p;
  a = SweepArgs{};
  a.nx = p.nx; a.ny = p.ny; a.nprobe = nprobe; a.probe0 = 0;
  a.kc = c->kc; a.kcp = c->kcp; a.band_a = c->band_a; a.band_b = c->band_b; a.ntiles = c->ntiles;
  a.potNy = p.potNy;
  a.rw = c->rows_per_wave;
  a.tpb = c->tiles_per_block;
  a.probe = c->probe.as<float2>();
  a.ipos = c->ipos.as<int2>();
  a.twx = c->twx.as<float2>(); a.twy = c->twy.as<float2>();
  a.px = c->px.as<float2>(); a.py = c->py.as<float2>();
  a.kx2 = c->dkx2.as<float>(); a.ky2 = c->dky2.as<float>();
  a.k2max = c->k2max;
  a.rec_bin = -1;
  a.ndet = 0;
  a.partial = c->partial.as<float>();
}

// Reverse-pass arguments shared by the two halves of the adjoint.
static qs::AdjArgs adj_args(qs_ctx* c, int nprobe) {
  const qs_params& p = c->p;
  SweepArgs a;
  sweep_args(c, a, nprobe);
  a.ipos = c->adj_ipos.as<int2>();
  a.state = c->state.as<float2>();
  qs::AdjArgs A{};
  A.f = a;
  A.gx = c->gx.as<float2>();
  A.target = c->target.as<float>();
  A.loss_partial = c->lossp.as<float>();
  A.gprobe = c->gprobe.as<float>();
  // a single window: one writer per pixel and launch; covering the whole slice:
  // every gT element is stored once per slice
  const bool whole = nprobe == 1 && p.nx == p.potNx && p.ny == p.potNy;
  A.single_writer = nprobe == 1 ? (whole ? 2 : 1) : 0;
  return A;
}

// First half of the adjoint: the forward pass with checkpoints (X_{jK-1} in
// ckpt[j], written in place by the row sweep that produces it; X_{S-1} in
// c->state), the reverse column sweep of the last slice (loss partials and the
// adjoint seed in gx) and the loss.
// Default checkpoint interval K for S slices of st_bytes wave states: the smallest K
// whose ceil(S/K) + K states fit in half the device memory that is free (counting
// this context's checkpoint buffers, which are reused) -- K = 1 keeps every slice and
// recomputes nothing -- and at most ceil(sqrt(S)), the interval of least memory.
static int ckpt_interval(const qs_ctx* c, int S, size_t st_bytes) {
  const int kmax = (int)ceil(sqrt((double)S));
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return kmax;
  for (const DevBuf& b : c->ckpt) free_b += b.bytes;
  for (int K = 1; K < kmax; K++)
    if ((size_t)((S + K - 1) / K + K) * st_bytes <= free_b / 2) return K;
  return kmax;
}

static int adj_forward(qs_ctx* c, const char* who, const int* ipos, int nprobe, const float* target, int ckpt_every,
                       float* loss_out) {
  if (!c) return QS_ERR_ARG;
  c->adj_saved = false;
  if (!ipos || nprobe < 1 || !target) return fail(c, QS_ERR_ARG, std::string(who) + ": bad arguments");
  if (c->generic)
    return fail(c, QS_ERR_UNSUPPORTED, std::string(who) + ": the adjoint needs power-of-two wave grids");
  const qs_params& p = c->p;
  const int S = trans_slices_of(c);
  if (S < 1) return fail(c, QS_ERR_STATE, std::string(who) + ": no transmission functions");
  if (nconfig_of(c) > 1 || c->trans_src)
    return fail(c, QS_ERR_UNSUPPORTED,
                std::string(who) + ": needs one configuration of the context's own transmission function");
  if (!c->have_probe) return fail(c, QS_ERR_STATE, std::string(who) + ": qs_make_probe not called");
  for (int i = 0; i < nprobe; i++) {
    const int x = ipos[2 * i], y = ipos[2 * i + 1];
    if (x < 0 || y < 0 || x + p.nx > p.potNx || y + p.ny > p.potNy)
      return fail(c, QS_ERR_ARG, std::string(who) + ": window offset outside the transmission array");
  }
  const size_t st_elems = (size_t)nprobe * p.nx * c->kcp;
  const size_t st_bytes = sizeof(float2) * st_elems;
  const int K = ckpt_every > 0 ? std::min(ckpt_every, S) : ckpt_interval(c, S, st_bytes);
  const int nseg = (S + K - 1) / K;
  const size_t slice_elems = (size_t)p.potNx * p.potNy;
  HIPCHK(c, c->state.ensure(st_bytes));
  HIPCHK(c, c->gx.ensure(st_bytes));
  if ((int)c->ckpt.size() < nseg + K) c->ckpt.resize(nseg + K);
  for (int j = 0; j < nseg + K; j++) HIPCHK(c, c->ckpt[j].ensure(st_bytes));  // [0,nseg): X_{jK-1}; [nseg,nseg+K): segment
  HIPCHK(c, c->adj_ipos.ensure(sizeof(int) * 2 * nprobe));
  HIPCHK(c, hipMemcpyAsync(c->adj_ipos.p, ipos, sizeof(int) * 2 * nprobe, hipMemcpyHostToDevice, c->stream));
  const size_t nt = (size_t)nprobe * p.nx * p.ny;
  HIPCHK(c, c->target.ensure(sizeof(float) * nt));
  HIPCHK(c, hipMemcpyAsync(c->target.p, target, sizeof(float) * nt, hipMemcpyDefault, c->stream));
  HIPCHK(c, c->partial.ensure(sizeof(float) * (size_t)nprobe * c->ntiles * (qs::kMaxDet + 1)));
  HIPCHK(c, c->lossp.ensure(sizeof(float) * (size_t)nprobe * c->ntiles));
  HIPCHK(c, c->dsum.ensure(sizeof(double) * 4));
  qs::launch_cplx_fill_zero(reinterpret_cast<float*>(c->dsum.p), 8, c->stream);
  qs::AdjArgs A = adj_args(c, nprobe);
  SweepArgs a = A.f;
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  qs::launch_fill_zero(a.state, st_elems, c->stream);
  float2* cur = c->state.as<float2>();  // X_{s-1}
  for (int s = 0; s < S; s++) {
    float2* out = ((s + 1) % K == 0 && s + 1 < S) ? c->ckpt[(s + 1) / K].as<float2>() : c->state.as<float2>();
    a.trans = trans_base(c) + (size_t)s * slice_elems;
    a.state = out;
    a.state_in = (s > 0 && cur != out) ? cur : nullptr;
    qs::launch_sweep_rows(a, s == 0 ? 0 : 1, c->stream);
    qs::launch_sweep_cols(a, false, false, c->stream);
    cur = out;
  }
  // reverse column sweep of the last slice: loss partials, adjoint seed
  qs::launch_fill_zero(A.gx, st_elems, c->stream);
  A.f.trans = trans_base(c) + (size_t)(S - 1) * slice_elems;
  A.xprev = c->state.as<float2>();
  qs::launch_adj_cols(A, true, c->stream);
  // loss = sum over band pixels (partials) + sum of target^2 outside the band
  qs::launch_target_outside_band(c->target.as<float>(), nprobe, p.nx, p.ny, c->band_a, c->band_b,
                                 c->dsum.as<double>() + 2, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  std::vector<float> lp((size_t)nprobe * c->ntiles);
  double ds[4];
  HIPCHK(c, hipMemcpyAsync(lp.data(), c->lossp.p, sizeof(float) * lp.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(ds, c->dsum.p, sizeof(double) * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
  double loss = ds[2];
  for (float v : lp) loss += v;
  if (loss_out) *loss_out = (float)loss;
  c->adj_loss = (float)loss;
  c->adj_nprobe = nprobe;
  c->adj_K = K;
  c->adj_saved = true;
  c->state_valid = false;
  return QS_OK;
}

// Second half: segment by segment, recompute the forward states inside the
// segment from its checkpoint, then the reverse sweeps of its slices (the last
// slice's column sweep ran in the first half); dL/dV and dL/d(df0, Cs).
static int adj_backward(qs_ctx* c, const char* who, float* grad_V, float* grad_probe) {
  if (!c) return QS_ERR_ARG;
  if (!c->adj_saved)
    return fail(c, QS_ERR_STATE, std::string(who) + ": no saved forward pass (qs_forward_save), or the potential, "
                                                     "probe or optics changed since");
  const qs_params& p = c->p;
  if (grad_V && !c->v_valid)
    return fail(c, QS_ERR_STATE, std::string(who) + ": grad_V needs qs_set_potential/qs_build_potential");
  if (grad_probe && c->custom_wave)
    return fail(c, QS_ERR_UNSUPPORTED, std::string(who) + ": probe gradients need the probe of qs_make_probe");
  if (grad_probe && (p.gaussFlag || p.aAIS > 0))
    return fail(c, QS_ERR_UNSUPPORTED, std::string(who) + ": probe gradients need gaussian: no and no AIS aperture");
  c->adj_saved = false;  // the reverse pass consumes the seed and the segment slots
  const int S = trans_slices_of(c), nprobe = c->adj_nprobe, K = c->adj_K;
  const int nseg = (S + K - 1) / K;
  const size_t slice_elems = (size_t)p.potNx * p.potNy;
  // dL/dt of a slice is complete once that slice's reverse row sweep has run (all
  // probes are in the launch), so dL/dV is formed slice by slice from a one-slice
  // dL/dt buffer; only the band limit by row transforms and transposes (grids that
  // are not powers of two) keeps dL/dt of every slice and works on all at once.
  const bool grad_each = grad_V && (!p.bandlimittrans || c->bl.on);
  const int gslices = grad_V && !grad_each ? S : 1;
  const float mm0 = 1.0F + p.v0 / 511.0F;
  const double sigma = mm0 * wavelength(p.v0);
  if (grad_V) HIPCHK(c, c->gV.ensure(sizeof(float) * slice_elems * S));
  HIPCHK(c, c->gprobe.ensure(sizeof(float2) * (size_t)p.nx * p.ny));
  qs::AdjArgs A = adj_args(c, nprobe);
  const bool whole = A.single_writer == 2;
  // fuse: one window covers the slice and the band limit runs as band-compact sweeps.
  // Each row of dL/dt is then complete in the reverse row sweep, which also does the
  // band limit's first pass (row FFT into the band-compact state), so dL/dt is never
  // stored at full size; the band limit's other two sweeps of slice s run on the second
  // stream beside the reverse sweeps of slice s-1 (overlap), from two band-compact
  // buffers used in turn.  Its buffers are allocated before the loop.
  const bool fuse = grad_each && p.bandlimittrans && c->bl.on && whole;
  if (!fuse) HIPCHK(c, c->gT.ensure(sizeof(float2) * slice_elems * gslices));
  const bool overlap = fuse && c->nstreams > 1 && S > 1;
  bool pending[2] = {false, false};  // the band limit on the second stream still reads buffer b
  if (fuse) {
    const size_t bl_per = sizeof(float2) * (size_t)p.potNx * c->bl.kcp;
    HIPCHK(c, c->blstate.ensure(std::max(c->blstate.bytes, bl_per)));
    if (overlap) {
      HIPCHK(c, c->blstate2.ensure(bl_per));
      for (int b = 0; b < 2; b++) {
        if (!c->evrow[b]) HIPCHK(c, hipEventCreateWithFlags(&c->evrow[b], hipEventDisableTiming));
        if (!c->evbl[b]) HIPCHK(c, hipEventCreateWithFlags(&c->evbl[b], hipEventDisableTiming));
      }
    }
    A.bl_band_a = c->bl.band_a;
    A.bl_gap = p.potNy - c->bl.band_b - c->bl.band_a - 1;
    A.bl_kcp = c->bl.kcp;
  }
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  if (!whole && gslices > 1) qs::launch_cplx_fill_zero(c->gT.as<float>(), 2 * slice_elems * S, c->stream);
  qs::launch_cplx_fill_zero(c->gprobe.as<float>(), 2 * (size_t)p.nx * p.ny, c->stream);
  for (int j = nseg - 1; j >= 0; j--) {
    const int s0 = j * K, s1 = std::min(S, s0 + K) - 1;
    // recompute X_s, s in [s0, s1-1], straight into segment slots nseg + (s - s0):
    // each row sweep reads the previous slot (checkpoint j first) and writes the next
    SweepArgs w = A.f;
    const float2* prev = s0 > 0 ? c->ckpt[j].as<float2>() : nullptr;
    for (int s = s0; s < s1; s++) {
      w.trans = trans_base(c) + (size_t)s * slice_elems;
      w.state = c->ckpt[nseg + (s - s0)].as<float2>();
      w.state_in = prev;
      qs::launch_sweep_rows(w, s == 0 ? 0 : 1, c->stream);
      qs::launch_sweep_cols(w, false, false, c->stream);
      prev = w.state;
    }
    for (int s = s1; s >= s0; s--) {
      const int b = overlap ? (s & 1) : 0;
      DevBuf& bls = b ? c->blstate2 : c->blstate;
      A.f.trans = trans_base(c) + (size_t)s * slice_elems;
      A.gT = fuse ? nullptr : c->gT.as<float>() + (gslices > 1 ? 2 * (size_t)s * slice_elems : 0);
      A.blstate = fuse ? bls.as<float2>() : nullptr;
      A.xprev = nullptr;
      if (s < S - 1) qs::launch_adj_cols(A, false, c->stream);
      if (s > 0) A.xprev = (s - 1 >= s0) ? c->ckpt[nseg + (s - 1 - s0)].as<float2>() : c->ckpt[j].as<float2>();
      if (grad_each && !whole) qs::launch_cplx_fill_zero(A.gT, 2 * slice_elems, c->stream);
      if (pending[b]) HIPCHK(c, hipStreamWaitEvent(c->stream, c->evbl[b], 0));  // slice s+2's band limit read it
      qs::launch_adj_rows(A, s == 0, c->stream);
      if (grad_each) {
        const float* vs = c->V.as<float>() + (size_t)s * slice_elems;
        float* gs = c->gV.as<float>() + (size_t)s * slice_elems;
        if (overlap) {
          HIPCHK(c, hipEventRecord(c->evrow[b], c->stream));
          HIPCHK(c, hipStreamWaitEvent(c->stream2, c->evrow[b], 0));
          const int rc = bl_sweeps(c, nullptr, 1, vs, gs, (float)sigma, c->stream2, &bls, true);
          if (rc != QS_OK) return rc;
          HIPCHK(c, hipEventRecord(c->evbl[b], c->stream2));
          pending[b] = true;
        } else if (p.bandlimittrans) {
          const int rc = bl_sweeps(c, fuse ? nullptr : reinterpret_cast<float2*>(A.gT), 1, vs, gs, (float)sigma,
                                   nullptr, nullptr, fuse);
          if (rc != QS_OK) return rc;
        } else {
          qs::launch_grad_v(reinterpret_cast<const float2*>(A.gT), vs, gs, slice_elems, sigma, c->stream);
        }
      }
    }
  }
  HIPCHK(c, hipGetLastError());
  if (overlap) {  // join: the first stream waits for the last band limits
    HIPCHK(c, hipEventRecord(c->evjoin, c->stream2));
    HIPCHK(c, hipStreamWaitEvent(c->stream, c->evjoin, 0));
  }
  // dL/dV of every slice at once: band limit (self-adjoint) of dL/dt, then sigma Im(g conj(u))
  if (grad_V && !grad_each) {
    int rc = bandlimit_buf(c, c->gT.as<float2>(), S);
    if (rc != QS_OK) return rc;
    qs::launch_grad_v(c->gT.as<float2>(), c->V.as<float>(), c->gV.as<float>(), slice_elems * S, sigma, c->stream);
  }
  if (grad_probe) {  // G = c * FFT(sum_p g_psi0)
    float2* g = c->gprobe.as<float2>();
    HIPCHK(c, c->scratch.ensure(std::max(c->scratch.bytes, sizeof(float2) * (size_t)p.nx * p.ny)));
    float2* sc = c->scratch.as<float2>();
    qs::launch_rows_fft(g, p.nx, p.ny, -1, c->twy.as<float2>(), 1.0f, c->stream);
    qs::launch_transpose(g, sc, 1, p.nx, p.ny, c->stream);
    qs::launch_rows_fft(sc, p.ny, p.nx, -1, c->twx.as<float2>(), 1.0f, c->stream);
    qs::launch_transpose(sc, g, 1, p.ny, p.nx, c->stream);
    qs::launch_probe_grad(c->probe_args, g, c->probe_norm, c->probe_args.wavlen, c->dsum.as<double>(), c->stream);
  }
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  double ds[4];
  HIPCHK(c, hipMemcpyAsync(ds, c->dsum.p, sizeof(double) * 4, hipMemcpyDeviceToHost, c->stream));
  if (grad_V)
    HIPCHK(c, hipMemcpyAsync(grad_V, c->gV.p, sizeof(float) * slice_elems * S, hipMemcpyDefault, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
  if (grad_probe) {
    grad_probe[0] = (float)ds[0];
    grad_probe[1] = (float)ds[1];
  }
  return QS_OK;
}

int qs_last_ckpt_interval(const qs_ctx* c) { return c ? c->adj_K : QS_ERR_ARG; }

int qs_forward_save(qs_ctx* c, const int* ipos, int nprobe, const float* target, int ckpt_every, float* loss) {
  return adj_forward(c, "qs_forward_save", ipos, nprobe, target, ckpt_every, loss);
}

int qs_adjoint_saved(qs_ctx* c, float* grad_V, float* grad_probe) {
  return adj_backward(c, "qs_adjoint_saved", grad_V, grad_probe);
}

int qs_adjoint(qs_ctx* c, const int* ipos, int nprobe, const float* target, int ckpt_every, float* loss_out,
               float* grad_V, float* grad_probe) {
  if (!c) return QS_ERR_ARG;
  if (grad_V && !c->v_valid) return fail(c, QS_ERR_STATE, "qs_adjoint: grad_V needs qs_set_potential/qs_build_potential");
  if (grad_probe && c->custom_wave)
    return fail(c, QS_ERR_UNSUPPORTED, "qs_adjoint: probe gradients need the probe of qs_make_probe");
  if (grad_probe && (c->p.gaussFlag || c->p.aAIS > 0))
    return fail(c, QS_ERR_UNSUPPORTED, "qs_adjoint: probe gradients need gaussian: no and no AIS aperture");
  int rc = adj_forward(c, "qs_adjoint", ipos, nprobe, target, ckpt_every, loss_out);
  if (rc != QS_OK) return rc;
  const float fwd_ms = c->last_ms;
  rc = adj_backward(c, "qs_adjoint", grad_V, grad_probe);
  if (rc != QS_OK) return rc;
  c->last_ms += fwd_ms;
  return QS_OK;
}

int qs_atom_gradients(qs_ctx* c, const qs_atom* atoms, int natom, const qs_pot_params* pp, float* grad_xyz) {
  if (!c || !pp || !grad_xyz || natom < 0 || (natom > 0 && !atoms)) return fail(c, QS_ERR_ARG, "qs_atom_gradients: bad arguments");
  if (c->gV.bytes == 0) return fail(c, QS_ERR_STATE, "qs_atom_gradients: run qs_adjoint with grad_V first");
  // the kinds and LUTs of the last qs_build_potential call are reused: the same atoms
  // (count, Z, charge) and potential parameters are required, checked here on the host
  // (the kernel indexes the kinds by atom and the LUTs by the geometry)
  if (!c->built.valid)
    return fail(c, QS_ERR_STATE, "qs_atom_gradients: the potential does not come from qs_build_potential");
  if (natom != (int)c->built.Z.size())
    return fail(c, QS_ERR_ARG, "qs_atom_gradients: atom count differs from the last qs_build_potential");
  for (int i = 0; i < natom; i++)
    if (atoms[i].Znum != c->built.Z[i] || (atoms[i].Znum > 0 && atoms[i].q != c->built.q[i]))
      return fail(c, QS_ERR_ARG, "qs_atom_gradients: atom Z or charge differs from the last qs_build_potential");
  {
    const qs_pot_params& b = c->built.pp;
    if (pp->atomRadius != b.atomRadius || pp->potOffsetX != b.potOffsetX || pp->potOffsetY != b.potOffsetY ||
        pp->czOffset != b.czOffset || pp->centerSlices != b.centerSlices || pp->periodicXY != b.periodicXY ||
        pp->tds != b.tds || pp->compat_q6 != b.compat_q6 || pp->lut_dr_fix != b.lut_dr_fix ||
        pp->compat_q5 != b.compat_q5 || pp->lut_norm_fix != b.lut_norm_fix)
      return fail(c, QS_ERR_ARG, "qs_atom_gradients: potential parameters differ from the last qs_build_potential");
  }
  if (c->luts.empty() || natom == 0) {
    memset(grad_xyz, 0, sizeof(float) * 3 * natom);
    return QS_OK;
  }
  const qs_params& p = c->p;
  bool charged = false;
  for (int i = 0; i < natom; i++) charged = charged || (atoms[i].Znum > 0 && atoms[i].q != 0.f);
  qs::ScatterArgs sa{};
  const float R = pp->atomRadius, dz = p.sliceThickness;
  int nzp = (int)floor(2 * dz / p.resolutionX);
  if (2.0 * (nzp >> 1) == nzp) nzp += 1;
  const int lnx = 2 * 2 * (int)ceil(R / p.resolutionX);
  const int lnz = (2 * (int)ceil(R / dz)) * nzp;
  sa.potNx = p.potNx; sa.potNy = p.potNy; sa.slices = p.slices;
  sa.dx = p.resolutionX; sa.dy = p.resolutionY; sa.sliceThickness = dz;
  sa.potOffsetX = pp->potOffsetX; sa.potOffsetY = pp->potOffsetY; sa.czOffset = pp->czOffset;
  sa.atomRadius = R; sa.centerSlices = pp->centerSlices; sa.periodic = pp->periodicXY;
  sa.iRadX = (int)ceil(R / p.resolutionX); sa.iRadY = (int)ceil(R / p.resolutionY); sa.iRadZ = (int)ceil(R / dz);
  sa.Nr = lnx / 2; sa.Nz = lnz / 2; sa.nzSub = nzp;
  sa.dr = pp->lut_dr_fix ? p.resolutionX : p.resolutionX / 2;
  HIPCHK(c, c->atoms.ensure(sizeof(qs_atom) * natom));
  HIPCHK(c, hipMemcpyAsync(c->atoms.p, atoms, sizeof(qs_atom) * natom, hipMemcpyHostToDevice, c->stream));
  DevBuf g;
  HIPCHK(c, g.ensure(sizeof(float) * 3 * natom));
  qs::launch_atom_grad(c->atoms.as<qs::QsAtomDev>(), natom, c->lutptrs.as<const float*>(),
                       charged ? c->olutptrs.as<const float*>() : nullptr, c->kinds.as<int>(), sa, c->gV.as<float>(),
                       g.as<float>(), c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(grad_xyz, g.p, sizeof(float) * 3 * natom, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  g.release();
  return QS_OK;
}
}  // extern "C"

// ---------------------------------------------------------------------------------
// Autotuner: picks the probe group size (batch) and the launch shape of the two
// sweeps by timing the slice loop on the context's own transmission function and
// probe, over the first 8 slices and up to 512 of the caller's window offsets.
// Coordinate descent: group size first (library launch defaults), then row-sweep
// steps per wavefront, then column tiles per block; a launch shape replaces the
// current best only when 2% faster (timing noise).  Results are cached per
// device model and wave geometry for the life of the process and, when the
// environment variable QSTEM_TUNE_CACHE names a file, in that JSON file across
// processes (read once per process, rewritten with the union of its entries and
// this process's after every new measurement; entries of another kernel version
// are ignored).
namespace {
// bumped whenever the sweep kernels or their launch parameters change meaning
const char* const kTuneVersion = "sweeps-4";
struct TuneKey {
  std::string device;  // gcnArchName / name / compute units
  int nx, ny, kcp, potNy, nstreams;
  bool operator<(const TuneKey& o) const {
    return std::tie(device, nx, ny, kcp, potNy, nstreams) <
           std::tie(o.device, o.nx, o.ny, o.kcp, o.potNy, o.nstreams);
  }
};
struct TuneVal {
  int batch, rw, tpb;
  float rate;
  int vx, vy;  // mixed-radix plan variants (0 for the power-of-two sweeps)
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneVal> g_tune_cache;
bool g_tune_file_read = false;

std::string tune_file() {
  const char* e = getenv("QSTEM_TUNE_CACHE");
  return e ? std::string(e) : std::string();
}

// One entry per line, as tune_save writes them; other lines are skipped.
void tune_load_locked(const std::string& path) {
  FILE* f = fopen(path.c_str(), "r");
  if (!f) return;
  char line[1024];
  while (fgets(line, sizeof line, f)) {
    char dev[256] = {0}, ver[64] = {0};
    TuneKey k;
    TuneVal v{};
    if (sscanf(line,
               " {\"version\": \"%63[^\"]\", \"device\": \"%255[^\"]\", \"nx\": %d, \"ny\": %d, \"kcp\": %d, "
               "\"potNy\": %d, \"streams\": %d, \"batch\": %d, \"rows_per_wave\": %d, \"tiles_per_block\": %d, "
               "\"plan_x\": %d, \"plan_y\": %d, \"rate\": %f}",
               ver, dev, &k.nx, &k.ny, &k.kcp, &k.potNy, &k.nstreams, &v.batch, &v.rw, &v.tpb, &v.vx, &v.vy,
               &v.rate) != 13)
      continue;
    if (strcmp(ver, kTuneVersion) != 0 || v.batch < 1) continue;
    k.device = dev;
    g_tune_cache.emplace(k, v);  // this process's own measurements win
  }
  fclose(f);
}

// Rewrites the file with the union of its entries and the in-memory ones (a write
// to a temporary file in the same directory, then a rename).
void tune_save_locked(const std::string& path) {
  tune_load_locked(path);
  const std::string tmp = path + ".tmp" + std::to_string((long)getpid());
  FILE* f = fopen(tmp.c_str(), "w");
  if (!f) return;
  fprintf(f, "{\"qstem_amd_autotune\": [\n");
  size_t i = 0;
  for (const auto& kv : g_tune_cache) {
    const TuneKey& k = kv.first;
    const TuneVal& v = kv.second;
    fprintf(f,
            " {\"version\": \"%s\", \"device\": \"%s\", \"nx\": %d, \"ny\": %d, \"kcp\": %d, \"potNy\": %d, "
            "\"streams\": %d, \"batch\": %d, \"rows_per_wave\": %d, \"tiles_per_block\": %d, \"plan_x\": %d, "
            "\"plan_y\": %d, \"rate\": %.1f}%s\n",
            kTuneVersion, k.device.c_str(), k.nx, k.ny, k.kcp, k.potNy, k.nstreams, v.batch, v.rw, v.tpb, v.vx, v.vy,
            v.rate, ++i < g_tune_cache.size() ? "," : "");
  }
  fprintf(f, "]}\n");
  const bool ok = fclose(f) == 0;
  if (!ok || rename(tmp.c_str(), path.c_str()) != 0) remove(tmp.c_str());
}
}  // namespace

int qs_autotune(qs_ctx* c, const int* ipos, int nprobe, int* batch, int* rows_per_wave, int* tiles_per_block,
                float* rate) {
  if (!c || !ipos || nprobe < 1) return fail(c, QS_ERR_ARG, "qs_autotune: bad arguments");
  if (c->generic && !c->mixed)
    return fail(c, QS_ERR_UNSUPPORTED,
                "qs_autotune: tunes the power-of-two sweeps and the mixed-radix path (prime factors 2, 3, 5, 7)");
  if (trans_slices_of(c) < 1 || !c->have_probe)
    return fail(c, QS_ERR_STATE, "qs_autotune: set the potential and the probe first");
  int dev = 0;
  HIPCHK(c, hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIPCHK(c, hipGetDeviceProperties(&prop, dev));
  std::string dname = std::string(prop.gcnArchName) + "/" + prop.name + "/" + std::to_string(prop.multiProcessorCount);
  for (char& ch : dname)
    if (ch == '"' || ch == '\\' || (unsigned char)ch < 32) ch = '_';
  const TuneKey key{dname, c->p.nx, c->p.ny, c->mixed ? -1 : c->kcp, c->p.potNy, c->nstreams};
  const std::string file = tune_file();
  TuneVal best{0, 0, 0, 0.f, 0, 0};
  bool cached = false;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (!file.empty() && !g_tune_file_read) {
      tune_load_locked(file);
      g_tune_file_read = true;
    }
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      best = it->second;
      cached = true;
    }
  }
  if (!cached) {
    const int n = std::min(nprobe, 512);
    const int S = std::min(8, trans_slices_of(c));
    const bool graph = c->use_graph, prof = c->profiling;
    c->use_graph = false;
    c->profiling = false;
    // mixed-radix path: rw = rows per block, tpb = columns per tile, vx/vy = plan variants
    int vx = c->gvx, vy = c->gvy;
    auto measure = [&](int B, int rw, int tpb, float* out) -> int {
      c->batch = B;
      if (c->mixed) {
        const int rc = set_mixed_shape(c, rw, tpb, vx, vy);
        if (rc != QS_OK) return rc;
      } else {
        c->rows_per_wave = rw;
        c->tiles_per_block = tpb;
      }
      float tmin = 1e30f;
      for (int r = 0; r < 3; r++) {  // first run also pays first-touch costs
        const int rc = qs_run_stem(c, ipos, n, 0, nullptr, S, 0, nullptr, nullptr, nullptr, nullptr);
        if (rc != QS_OK) return rc;
        if (r > 0) tmin = std::min(tmin, c->last_ms);
      }
      *out = (float)((double)n * S / (tmin * 1e-3));
      return QS_OK;
    };
    auto consider = [&](int B, int rw, int tpb) -> int {
      float r = 0.f;
      const int rc = measure(B, rw, tpb, &r);
      if (rc != QS_OK) return rc;
      if (r > best.rate * (best.batch == B ? 1.02f : 1.0f))  // launch shapes: 2% hysteresis
        best = TuneVal{B, rw, tpb, r, vx, vy};
      return QS_OK;
    };
    int rc = QS_OK;
    const int Bs[] = {16, 32, 64, 96, 128, 192, 256, 384, 512};
    for (int B : Bs)
      if (rc == QS_OK && (B <= n || B == Bs[0])) rc = consider(std::min(B, n), 0, 0);
    const int B0 = best.batch;
    if (c->mixed) {
      for (int rw : {1, 2, 4, 8, 16})
        if (rc == QS_OK && rw <= qs::gen_max_vectors(c->p.ny)) rc = consider(B0, rw, best.tpb);
      for (int tc : {2, 4, 8, 16})  // (the compiled column kernel has no tiles)
        if (rc == QS_OK && vx != qs::kGenPlanCompiled && tc <= qs::gen_max_vectors(c->p.nx))
          rc = consider(B0, best.rw, tc);
      qs::GenPlan pl{};
      for (int v = 0; v < qs::kGenPlanVariants && rc == QS_OK; v++) {  // x plan, then y plan
        if (!qs::gen_plan_variant(c->p.nx, v, &pl)) continue;
        vx = v;
        vy = best.vy;
        rc = consider(B0, best.rw, best.tpb);
      }
      for (int v = 0; v < qs::kGenPlanVariants && rc == QS_OK; v++) {
        if (!qs::gen_plan_variant(c->p.ny, v, &pl)) continue;
        vx = best.vx;
        vy = v;
        // rows per block of the runtime kernels and rows per slot of the compiled one
        // differ in meaning: each plan with its own default
        rc = consider(B0, (v == qs::kGenPlanCompiled) != (best.vy == qs::kGenPlanCompiled) ? 0 : best.rw, best.tpb);
      }
    } else {
      for (int rw : {1, 2, 4, 8})
        if (rc == QS_OK) rc = consider(B0, rw, best.tpb);
      const int rw0 = best.rw;
      for (int tpb : {1, 2, 4})
        if (rc == QS_OK) rc = consider(B0, rw0, tpb);
    }
    c->use_graph = graph;
    c->profiling = prof;
    if (rc != QS_OK) return rc;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = best;
    if (!file.empty()) tune_save_locked(file);
  }
  c->batch = best.batch;
  if (c->mixed) {
    const int rc = set_mixed_shape(c, best.rw, best.tpb, best.vx, best.vy);
    if (rc != QS_OK) return rc;
  } else {
    c->rows_per_wave = best.rw;
    c->tiles_per_block = best.tpb;
  }
  drop_graph(c);
  if (batch) *batch = best.batch;
  if (rows_per_wave) *rows_per_wave = best.rw;
  if (tiles_per_block) *tiles_per_block = best.tpb;
  if (rate) *rate = best.rate;
  return QS_OK;
}
