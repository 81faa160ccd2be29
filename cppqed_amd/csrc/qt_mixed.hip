// qt_mixed.hip -- launch table of the mixed-batch kernel (several ensembles in one launch).
// The interaction-picture and GSL-capable instances are compiled in their own translation
// units (qt_mixed_td.hip, qt_mixed_gsl.hip, qt_mixed_td_gsl.hip) by the same table macro.
#include "qt_mixed.hpp"

namespace qt {

QT_MIXED_TABLE(plain, false, false)

bool mixed_supported(int nwb, int rb, int whb, int rw, int whw) {
#define QT_CASE(A, B, C, D, F) if (nwb == A && rb == B && whb == C && rw == D && whw == F) return true;
  QT_MIXED(QT_CASE)
#undef QT_CASE
  return false;
}

// compiled mixed variants as rows of 5 ints; returns their number
int mixed_variants(int* table, int cap) {
  int n = 0;
#define QT_CASE(A, B, C, D, F) \
  if (n < cap && table) { int* r = table + 5 * n; r[0] = A; r[1] = B; r[2] = C; r[3] = D; r[4] = F; } ++n;
  QT_MIXED(QT_CASE)
#undef QT_CASE
  return n;
}

int mixed_blocks_per_cu(int nwb, int rb, int whb, int rw, int whw, bool f32, bool td, bool gsl) {
  if (td && gsl) return mixed_blocks_td_gsl(nwb, rb, whb, rw, whw, f32);
  if (td) return mixed_blocks_td(nwb, rb, whb, rw, whw, f32);
  if (gsl) return mixed_blocks_gsl(nwb, rb, whb, rw, whw, f32);
  return mixed_blocks_plain(nwb, rb, whb, rw, whw, f32);
}

hipError_t launch_mixed(int nwb, int rb, int whb, int rw, int whw, bool f32, bool td, bool gsl, int grid,
                        const MixDev* M, hipStream_t st) {
  if (td && gsl) return launch_mixed_td_gsl(nwb, rb, whb, rw, whw, f32, grid, M, st);
  if (td) return launch_mixed_td(nwb, rb, whb, rw, whw, f32, grid, M, st);
  if (gsl) return launch_mixed_gsl(nwb, rb, whb, rw, whw, f32, grid, M, st);
  return launch_mixed_plain(nwb, rb, whb, rw, whw, f32, grid, M, st);
}

}  // namespace qt
