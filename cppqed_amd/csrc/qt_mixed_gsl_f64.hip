// qt_mixed_gsl_f64.hip -- launch table of the mixed-batch kernel instances:
// Schroedinger picture capable, GSL's controller, FP64.
// One unit per (picture, controller, precision), so that the instantiations compile in parallel.
#define QT_TBL_F32 false
#define QT_TBL_TD false
#define QT_TBL_GSL true
#include "qt_mixed.hpp"

namespace qt {
QT_MIXED_TABLE(gsl_f64)
}  // namespace qt
