// qt_mixed_td.hip -- mixed-batch kernel instances (interaction picture: true, GSL controller: false).
#include "qt_mixed.hpp"

namespace qt {
QT_MIXED_TABLE(td, true, false)
}  // namespace qt
