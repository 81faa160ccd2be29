// qt_batch.hip -- launch table of the Schroedinger-picture batched MCWF kernels (odeint
// controller) and the dispatch over every batched variant.
#include "qt_engine.hpp"

namespace qt {

QT_BATCH_TABLE(batch_sch, QT_BATCH, false, false)

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
  if (gsl && td) return batch_gsl_td_blocks(rows, nw, wh, minw, f32);
  if (gsl) return batch_gsl_blocks(rows, nw, wh, minw, f32);
  if (td) return batch_td_blocks(rows, nw, wh, minw, f32);
  return batch_sch_blocks(rows, nw, wh, minw, f32);
}

hipError_t launch_batch(int rows, int nw, int wh, int minw, bool f32, bool td, bool gsl, int grid, const SysDev& S,
                        const EnsDev& E, hipStream_t st) {
  if (gsl && td) return batch_gsl_td_launch(rows, nw, wh, minw, f32, grid, S, E, st);
  if (gsl) return batch_gsl_launch(rows, nw, wh, minw, f32, grid, S, E, st);
  if (td) return batch_td_launch(rows, nw, wh, minw, f32, grid, S, E, st);
  return batch_sch_launch(rows, nw, wh, minw, f32, grid, S, E, st);
}

}  // namespace qt
