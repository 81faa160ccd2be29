// qt_mixed.hpp -- compiled variants and launch tables of the mixed-batch kernel.
#pragma once
#include "qt_engine.hpp"

namespace qt {

// Mixed-batch variants: (waves per work group, rows per lane of the work-group team, its ELL
// register width, rows per lane of the wavefront team, its ELL register width).
#define QT_MIXED(X) \
  X(16, 4, 16, 1, 8) X(16, 4, 16, 2, 8) X(16, 2, 16, 1, 8) X(16, 1, 16, 1, 8)

// resident work groups per compute unit and the launch of every variant, for one choice of
// (interaction-picture capable, GSL capable) instances
#define QT_MIXED_TABLE(NAME, TDC, GSLC)                                                                         \
  int mixed_blocks_##NAME(int nwb, int rb, int whb, int rw, int whw, bool f32) {                               \
    int n = 0;                                                                                                  \
    QT_MIXED(QT_MIXED_B_##TDC##_##GSLC)                                                                         \
    return 0;                                                                                                   \
  }                                                                                                             \
  hipError_t launch_mixed_##NAME(int nwb, int rb, int whb, int rw, int whw, bool f32, int grid, const MixDev* M, \
                                 hipStream_t st) {                                                              \
    QT_MIXED(QT_MIXED_L_##TDC##_##GSLC)                                                                         \
    return hipErrorInvalidValue;                                                                                \
  }
#define QT_MIXED_B(A, B, C, D, F, TDC, GSLC)                                                                    \
  if (nwb == A && rb == B && whb == C && rw == D && whw == F) {                                                 \
    if (f32) (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, mcwf_mixed_kernel<A, B, C, D, F, true, TDC, GSLC>, 64 * A, 0); \
    else (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, mcwf_mixed_kernel<A, B, C, D, F, false, TDC, GSLC>, 64 * A, 0);   \
    return n;                                                                                                   \
  }
#define QT_MIXED_L(A, B, C, D, F, TDC, GSLC)                                                                    \
  if (nwb == A && rb == B && whb == C && rw == D && whw == F) {                                                 \
    if (f32) hipLaunchKernelGGL((mcwf_mixed_kernel<A, B, C, D, F, true, TDC, GSLC>), dim3(grid), dim3(64 * A), 0, st, M); \
    else hipLaunchKernelGGL((mcwf_mixed_kernel<A, B, C, D, F, false, TDC, GSLC>), dim3(grid), dim3(64 * A), 0, st, M);    \
    return hipGetLastError();                                                                                   \
  }
#define QT_MIXED_B_false_false(A, B, C, D, F) QT_MIXED_B(A, B, C, D, F, false, false)
#define QT_MIXED_B_true_false(A, B, C, D, F) QT_MIXED_B(A, B, C, D, F, true, false)
#define QT_MIXED_B_false_true(A, B, C, D, F) QT_MIXED_B(A, B, C, D, F, false, true)
#define QT_MIXED_B_true_true(A, B, C, D, F) QT_MIXED_B(A, B, C, D, F, true, true)
#define QT_MIXED_L_false_false(A, B, C, D, F) QT_MIXED_L(A, B, C, D, F, false, false)
#define QT_MIXED_L_true_false(A, B, C, D, F) QT_MIXED_L(A, B, C, D, F, true, false)
#define QT_MIXED_L_false_true(A, B, C, D, F) QT_MIXED_L(A, B, C, D, F, false, true)
#define QT_MIXED_L_true_true(A, B, C, D, F) QT_MIXED_L(A, B, C, D, F, true, true)

#define QT_MIXED_DECL(NAME)                                                                                     \
  int mixed_blocks_##NAME(int nwb, int rb, int whb, int rw, int whw, bool f32);                                 \
  hipError_t launch_mixed_##NAME(int nwb, int rb, int whb, int rw, int whw, bool f32, int grid, const MixDev* M, \
                                 hipStream_t st);
QT_MIXED_DECL(plain)
QT_MIXED_DECL(td)
QT_MIXED_DECL(gsl)
QT_MIXED_DECL(td_gsl)

}  // namespace qt
