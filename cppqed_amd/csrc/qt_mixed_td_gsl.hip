// qt_mixed_td_gsl.hip -- mixed-batch kernel instances (interaction picture: true, GSL controller: true).
#include "qt_mixed.hpp"

namespace qt {
QT_MIXED_TABLE(td_gsl, true, true)
}  // namespace qt
