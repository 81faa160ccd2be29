// qt_batch_td_f64.hip -- launch table of the batched MCWF kernel instances:
// interaction picture capable, odeint's controller, FP64.
// One unit per (picture, controller, precision), so that the instantiations compile in parallel.
#define QT_TBL_F32 false
#define QT_TBL_TD true
#define QT_TBL_GSL false
#include "qt_engine.hpp"

namespace qt {
QT_BATCH_TABLE(batch_td_f64, QT_BATCH_TD)
}  // namespace qt
