// qt_mixed.hpp -- compiled variants and launch tables of the mixed-batch kernel.
#pragma once
#include "qt_engine.hpp"

namespace qt {

// Mixed-batch variants: (waves per work group, rows per lane of the work-group team, its ELL
// register width, rows per lane of the wavefront team, its ELL register width).
#define QT_MIXED(X) \
  X(16, 4, 16, 1, 8) X(16, 4, 16, 2, 8) X(16, 2, 16, 1, 8) X(16, 1, 16, 1, 8)

// resident work groups per compute unit and the launch of every variant, each as two kernels
// (ph 1: work-group teams; ph 6: quarter and wavefront teams, one quarter team per work group of
// A / 4 wavefronts, mcwf_mixed_kernel's PH 14), for one choice of
// (interaction-picture capable, GSL capable, precision) instances: the translation unit that
// invokes it defines QT_TBL_F32, QT_TBL_TD and QT_TBL_GSL (each choice has a unit of its own, so
// that the heavy instantiations compile in parallel)
#define QT_MIXED_TABLE(NAME)                                                                                    \
  int mixed_blocks_##NAME(int nwb, int rb, int whb, int rw, int whw, int ph) {                                  \
    int n = 0;                                                                                                  \
    QT_MIXED(QT_MIXED_B)                                                                                        \
    return 0;                                                                                                   \
  }                                                                                                             \
  hipError_t launch_mixed_##NAME(int nwb, int rb, int whb, int rw, int whw, int ph, int grid, const MixDev* M,  \
                                 hipStream_t st) {                                                              \
    QT_MIXED(QT_MIXED_L)                                                                                        \
    return hipErrorInvalidValue;                                                                                \
  }
#define QT_MIXED_B(A, B, C, D, F)                                                                               \
  if (nwb == A && rb == B && whb == C && rw == D && whw == F) {                                                 \
    if (ph == 1)                                                                                                \
      (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(                                                       \
          &n, mcwf_mixed_kernel<A, B, C, D, F, QT_TBL_F32, QT_TBL_TD, QT_TBL_GSL, 1>, 64 * A, 0);               \
    else                                                                                                        \
      (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(                                                       \
          &n, mcwf_mixed_kernel<A, B, C, D, F, QT_TBL_F32, QT_TBL_TD, QT_TBL_GSL, 14>, 16 * A, 0);              \
    return n;                                                                                                   \
  }
#define QT_MIXED_L(A, B, C, D, F)                                                                               \
  if (nwb == A && rb == B && whb == C && rw == D && whw == F) {                                                 \
    if (ph == 1)                                                                                                \
      hipLaunchKernelGGL((mcwf_mixed_kernel<A, B, C, D, F, QT_TBL_F32, QT_TBL_TD, QT_TBL_GSL, 1>), dim3(grid),  \
                         dim3(64 * A), 0, st, M);                                                               \
    else                                                                                                        \
      hipLaunchKernelGGL((mcwf_mixed_kernel<A, B, C, D, F, QT_TBL_F32, QT_TBL_TD, QT_TBL_GSL, 14>), dim3(grid), \
                         dim3(16 * A), 0, st, M);                                                               \
    return hipGetLastError();                                                                                   \
  }

#define QT_MIXED_DECL(NAME)                                                                                     \
  int mixed_blocks_##NAME(int nwb, int rb, int whb, int rw, int whw, int ph);                                   \
  hipError_t launch_mixed_##NAME(int nwb, int rb, int whb, int rw, int whw, int ph, int grid, const MixDev* M,  \
                                 hipStream_t st);
#define QT_MIXED_DECL2(NAME) QT_MIXED_DECL(NAME##_f32) QT_MIXED_DECL(NAME##_f64)
QT_MIXED_DECL2(plain)
QT_MIXED_DECL2(td)
QT_MIXED_DECL2(gsl)
QT_MIXED_DECL2(td_gsl)

}  // namespace qt
