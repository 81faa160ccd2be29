// qt_batch_sch_f64.hip -- launch table of the batched MCWF kernel instances:
// Schroedinger picture capable, odeint's controller, FP64.
// One unit per (picture, controller, precision), so that the instantiations compile in parallel.
#define QT_TBL_F32 false
#define QT_TBL_TD false
#define QT_TBL_GSL false
#include "qt_engine.hpp"

namespace qt {
QT_BATCH_TABLE(batch_sch_f64, QT_BATCH)
}  // namespace qt
