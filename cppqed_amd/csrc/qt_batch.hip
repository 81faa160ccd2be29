// qt_batch.hip -- the dispatch over every batched variant; the launch tables are compiled in
// their own units (qt_batch_{sch,td,gsl,gsl_td}_{f32,f64}.hip).
#include "qt_engine.hpp"

namespace qt {

bool layout_supported(int rows, int nw, int wh) {
#define QT_CASE(R, N, W) if (rows == R && nw == N && wh == W) return true;
  QT_LAYOUTS(QT_CASE)
#undef QT_CASE
  return false;
}

bool batch_supported(int rows, int nw, int wh, int minw, bool td, bool gsl) {
#define QT_CASE(R, N, W, M) if (rows == R && nw == N && wh == W && minw == M) return true;
  if (gsl) { QT_BATCH_GSL(QT_CASE) }
  else if (td) { QT_BATCH_TD(QT_CASE) }
  else { QT_BATCH(QT_CASE) }
#undef QT_CASE
  return false;
}

// resident work groups of a batched variant per compute unit (registers and LDS permitting)
int batch_blocks_per_cu(int rows, int nw, int wh, int minw, bool f32, bool td, bool gsl) {
  if (gsl && td) return f32 ? batch_gsl_td_f32_blocks(rows, nw, wh, minw) : batch_gsl_td_f64_blocks(rows, nw, wh, minw);
  if (gsl) return f32 ? batch_gsl_f32_blocks(rows, nw, wh, minw) : batch_gsl_f64_blocks(rows, nw, wh, minw);
  if (td) return f32 ? batch_td_f32_blocks(rows, nw, wh, minw) : batch_td_f64_blocks(rows, nw, wh, minw);
  return f32 ? batch_sch_f32_blocks(rows, nw, wh, minw) : batch_sch_f64_blocks(rows, nw, wh, minw);
}

hipError_t launch_batch(int rows, int nw, int wh, int minw, bool f32, bool td, bool gsl, int grid, const SysDev& S,
                        const EnsDev& E, hipStream_t st) {
  if (gsl && td)
    return f32 ? batch_gsl_td_f32_launch(rows, nw, wh, minw, grid, S, E, st)
               : batch_gsl_td_f64_launch(rows, nw, wh, minw, grid, S, E, st);
  if (gsl)
    return f32 ? batch_gsl_f32_launch(rows, nw, wh, minw, grid, S, E, st)
               : batch_gsl_f64_launch(rows, nw, wh, minw, grid, S, E, st);
  if (td)
    return f32 ? batch_td_f32_launch(rows, nw, wh, minw, grid, S, E, st)
               : batch_td_f64_launch(rows, nw, wh, minw, grid, S, E, st);
  return f32 ? batch_sch_f32_launch(rows, nw, wh, minw, grid, S, E, st)
             : batch_sch_f64_launch(rows, nw, wh, minw, grid, S, E, st);
}

}  // namespace qt
