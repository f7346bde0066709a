// qs_mix.hip -- per-length instantiation of the compiled-plan sweeps (qs_mix.hpp).
// Compiled once per length of QS_MIX_SIZES with -DQS_MIX_N=<n> so the lengths build in
// parallel.
#include "qs_mix.hpp"

#ifndef QS_MIX_N
#error "compile with -DQS_MIX_N=<a length of QS_MIX_SIZES>"
#endif
#define QS_CAT2(a, b) a##b
#define QS_CAT(a, b) QS_CAT2(a, b)

namespace qs {
static_assert(MixPlan<QS_MIX_N>::E > 0, "no compiled plan for this length");
void QS_CAT(mix_rows_, QS_MIX_N)(const GenRowArgs& a, hipStream_t st) { launch_mix_rows_t<QS_MIX_N>(a, st); }
void QS_CAT(mix_cols_, QS_MIX_N)(const GenColArgs& a, int nblocks, hipStream_t st) {
  launch_mix_cols_t<QS_MIX_N>(a, nblocks, st);
}
int QS_CAT(mix_cpb_, QS_MIX_N)() { return mix_cpb<QS_MIX_N>(); }
}  // namespace qs
