// qs_sizes.hip -- per-size instantiation of the sweep kernels.
// Compiled once per transform length with -DQS_N=<n> so the sizes build in parallel.
#include "qs_sweep.hpp"

#ifndef QS_N
#error "compile with -DQS_N=<power of two>"
#endif
#define QS_CAT2(a, b) a##b
#define QS_CAT(a, b) QS_CAT2(a, b)

namespace qs {
void QS_CAT(sweep_rows_, QS_N)(const SweepArgs& a, int mode, hipStream_t st) { launch_rows_mode<QS_N>(a, mode, st); }
void QS_CAT(sweep_cols_, QS_N)(const SweepArgs& a, bool rec, bool diff, hipStream_t st) {
  launch_cols_t<QS_N>(a, rec, diff, st);
}
hipError_t QS_CAT(set_col_lds_, QS_N)() {
  hipError_t e = set_col_lds_t<QS_N>();
  return e == hipSuccess ? set_adj_lds_t<QS_N>() : e;
}
void QS_CAT(adj_rows_, QS_N)(const AdjArgs& a, bool first, hipStream_t st) { launch_adj_rows_t<QS_N>(a, first, st); }
void QS_CAT(adj_cols_, QS_N)(const AdjArgs& a, bool last, hipStream_t st) { launch_adj_cols_t<QS_N>(a, last, st); }
int QS_CAT(col_tile_, QS_N)() { return ColCfg<QS_N>::TC; }
void QS_CAT(rows_fft_, QS_N)(float2* d, long nrows, int dir, const float2* tw, float scale, hipStream_t st) {
  if (dir < 0) rows_fft_t<QS_N, -1>(d, nrows, tw, scale, st);
  else rows_fft_t<QS_N, +1>(d, nrows, tw, scale, st);
}
void QS_CAT(bluestein_, QS_N)(float2* d, long nrows, int n, const float2* chirp, const float2* vhat, const float2* tw,
                               float scale, hipStream_t st) {
  blue_t<QS_N>(d, nrows, n, chirp, vhat, tw, scale, st);
}
}  // namespace qs
