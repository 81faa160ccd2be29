// qt_batch_gsl.hip -- launch tables of the GSL-controller variants of the batched MCWF kernel
// (Schroedinger and interaction picture).
#include "qt_engine.hpp"

namespace qt {

QT_BATCH_TABLE(batch_gsl, QT_BATCH_GSL, false, true)
QT_BATCH_TABLE(batch_gsl_td, QT_BATCH_GSL, true, true)

}  // namespace qt
