// qt_mixed_plain_f64.hip -- launch table of the mixed-batch kernel instances:
// Schroedinger picture capable, odeint's controller, FP64.
// One unit per (picture, controller, precision), so that the instantiations compile in parallel.
#define QT_TBL_F32 false
#define QT_TBL_TD false
#define QT_TBL_GSL false
#include "qt_mixed.hpp"

namespace qt {
QT_MIXED_TABLE(plain_f64)
}  // namespace qt
