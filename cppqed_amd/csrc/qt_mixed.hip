// qt_mixed.hip -- launch table of the mixed-batch kernel (several ensembles in one launch).
#include "qt_engine.hpp"

namespace qt {

// Mixed-batch variants: (waves per work group, rows per lane of the work-group team, its ELL
// register width, rows per lane of the wavefront team, its ELL register width).
#define QT_MIXED(X) \
  X(16, 4, 16, 1, 8) X(16, 4, 16, 2, 8) X(16, 2, 16, 1, 8) X(16, 1, 16, 1, 8)

bool mixed_supported(int nwb, int rb, int whb, int rw, int whw) {
#define QT_CASE(A, B, C, D, F) if (nwb == A && rb == B && whb == C && rw == D && whw == F) return true;
  QT_MIXED(QT_CASE)
#undef QT_CASE
  return false;
}

// compiled mixed variants as rows of 5 ints; returns their number
int mixed_variants(int* table, int cap) {
  int n = 0;
#define QT_CASE(A, B, C, D, F) \
  if (n < cap && table) { int* r = table + 5 * n; r[0] = A; r[1] = B; r[2] = C; r[3] = D; r[4] = F; } ++n;
  QT_MIXED(QT_CASE)
#undef QT_CASE
  return n;
}

int mixed_blocks_per_cu(int nwb, int rb, int whb, int rw, int whw, bool f32) {
  int n = 0;
#define QT_CASE(A, B, C, D, F)                                                                                      \
  if (nwb == A && rb == B && whb == C && rw == D && whw == F) {                                                     \
    if (f32) (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, mcwf_mixed_kernel<A, B, C, D, F, true>, 64 * A, 0); \
    else (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, mcwf_mixed_kernel<A, B, C, D, F, false>, 64 * A, 0);   \
    return n;                                                                                                       \
  }
  QT_MIXED(QT_CASE)
#undef QT_CASE
  return 0;
}

hipError_t launch_mixed(int nwb, int rb, int whb, int rw, int whw, bool f32, int grid, const MixDev* M, hipStream_t st) {
#define QT_CASE(A, B, C, D, F)                                                                                          \
  if (nwb == A && rb == B && whb == C && rw == D && whw == F) {                                                         \
    if (f32) hipLaunchKernelGGL((mcwf_mixed_kernel<A, B, C, D, F, true>), dim3(grid), dim3(64 * A), 0, st, M);          \
    else hipLaunchKernelGGL((mcwf_mixed_kernel<A, B, C, D, F, false>), dim3(grid), dim3(64 * A), 0, st, M);             \
    return hipGetLastError();                                                                                           \
  }
  QT_MIXED(QT_CASE)
#undef QT_CASE
  return hipErrorInvalidValue;
}

}  // namespace qt
