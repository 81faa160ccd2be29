// qt_mixed.hip -- the dispatch over the mixed-batch kernel's instances (several ensembles in one
// launch); the launch tables are compiled in their own units
// (qt_mixed_{plain,td,gsl,td_gsl}_{f32,f64}.hip) by one table macro.
#include "qt_mixed.hpp"

namespace qt {

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

// ph: 1 for the work-group-team kernel, 6 for the quarter- and wavefront-team kernel
int mixed_blocks_per_cu(int nwb, int rb, int whb, int rw, int whw, int ph, bool f32, bool td, bool gsl) {
  if (td && gsl)
    return f32 ? mixed_blocks_td_gsl_f32(nwb, rb, whb, rw, whw, ph) : mixed_blocks_td_gsl_f64(nwb, rb, whb, rw, whw, ph);
  if (td) return f32 ? mixed_blocks_td_f32(nwb, rb, whb, rw, whw, ph) : mixed_blocks_td_f64(nwb, rb, whb, rw, whw, ph);
  if (gsl) return f32 ? mixed_blocks_gsl_f32(nwb, rb, whb, rw, whw, ph) : mixed_blocks_gsl_f64(nwb, rb, whb, rw, whw, ph);
  return f32 ? mixed_blocks_plain_f32(nwb, rb, whb, rw, whw, ph) : mixed_blocks_plain_f64(nwb, rb, whb, rw, whw, ph);
}

hipError_t launch_mixed(int nwb, int rb, int whb, int rw, int whw, int ph, bool f32, bool td, bool gsl, int grid,
                        const MixDev* M, hipStream_t st) {
  if (td && gsl)
    return f32 ? launch_mixed_td_gsl_f32(nwb, rb, whb, rw, whw, ph, grid, M, st)
               : launch_mixed_td_gsl_f64(nwb, rb, whb, rw, whw, ph, grid, M, st);
  if (td)
    return f32 ? launch_mixed_td_f32(nwb, rb, whb, rw, whw, ph, grid, M, st)
               : launch_mixed_td_f64(nwb, rb, whb, rw, whw, ph, grid, M, st);
  if (gsl)
    return f32 ? launch_mixed_gsl_f32(nwb, rb, whb, rw, whw, ph, grid, M, st)
               : launch_mixed_gsl_f64(nwb, rb, whb, rw, whw, ph, grid, M, st);
  return f32 ? launch_mixed_plain_f32(nwb, rb, whb, rw, whw, ph, grid, M, st)
             : launch_mixed_plain_f64(nwb, rb, whb, rw, whw, ph, grid, M, st);
}

}  // namespace qt
