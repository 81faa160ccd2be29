// qt_mixed_td_f64.hip -- launch table of the mixed-batch kernel instances:
// interaction picture capable, odeint's controller, FP64.
// One unit per (picture, controller, precision), so that the instantiations compile in parallel.
#define QT_TBL_F32 false
#define QT_TBL_TD true
#define QT_TBL_GSL false
#include "qt_mixed.hpp"

namespace qt {
QT_MIXED_TABLE(td_f64)
}  // namespace qt
