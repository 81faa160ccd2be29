// qt_batch_td.hip -- launch table of the interaction-picture variants of the batched MCWF kernel.
#include "qt_engine.hpp"

namespace qt {

QT_BATCH_TABLE(batch_td, QT_BATCH_TD, true, false)

}  // namespace qt
