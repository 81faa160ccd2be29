// qt_mixed_gsl.hip -- mixed-batch kernel instances (interaction picture: false, GSL controller: true).
#include "qt_mixed.hpp"

namespace qt {
QT_MIXED_TABLE(gsl, false, true)
}  // namespace qt
