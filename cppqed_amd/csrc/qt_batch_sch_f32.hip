// qt_batch_sch_f32.hip -- launch table of the batched MCWF kernel instances:
// Schroedinger picture capable, odeint's controller, FP32 derivatives.
// One unit per (picture, controller, precision), so that the instantiations compile in parallel.
#define QT_TBL_F32 true
#define QT_TBL_TD false
#define QT_TBL_GSL false
#include "qt_engine.hpp"

namespace qt {
QT_BATCH_TABLE(batch_sch_f32, QT_BATCH)
}  // namespace qt
